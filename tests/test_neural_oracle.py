"""CPU pinning of the neural-field oracle (oracle/neural.py): bf16 rounding
against torch's conversion, the backward pass against central finite
differences of the unrounded float64 forward pass (a dropped term, a wrong
mask or a transposed operand fails it), the hash index against its
definition, and Adam against a closed-form first step."""
import numpy as np
import pytest

from oracle import neural as nn


def test_bf16_rounding_matches_torch():
    import torch
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 10000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.0078125, 3.3895314e38], np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(nn.bf16(x), ref)


def small_problem(seed=0, n=96, levels=4, log2T=8):
    cfg = nn.Config(levels=levels, log2_table=log2T, n_min=4, n_max=40)
    rng = np.random.default_rng(seed)
    table = rng.uniform(-0.5, 0.5, (levels, 1 << log2T, 2))
    W1 = rng.uniform(-0.4, 0.4, (32, 64))
    W2 = rng.uniform(-0.3, 0.3, (64, 64))
    w3 = rng.uniform(-0.3, 0.3, 64)
    x = rng.uniform(0.02, 0.98, (n, 3))
    t = rng.uniform(0, 1, n)
    return cfg, [table, W1, W2, w3], x, t


def test_backward_matches_finite_differences():
    cfg, params, x, t = small_problem()
    loss, grads = nn.loss_and_grads(cfg, params, x, t, rnd=nn.same, dt=np.float64)
    rng = np.random.default_rng(1)
    eps = 1e-6
    for k, (p, g) in enumerate(zip(params, grads)):
        flat = p.reshape(-1)
        gflat = g.reshape(-1)
        # probe entries with non-zero gradient plus random ones
        nz = np.flatnonzero(np.abs(gflat) > 1e-8)
        idx = np.concatenate([rng.choice(nz, min(12, len(nz)), replace=False), rng.integers(0, flat.size, 4)])
        for i in idx:
            old = flat[i]
            flat[i] = old + eps
            lp = nn.loss_and_grads(cfg, params, x, t, rnd=nn.same, dt=np.float64)[0]
            flat[i] = old - eps
            lm = nn.loss_and_grads(cfg, params, x, t, rnd=nn.same, dt=np.float64)[0]
            flat[i] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - gflat[i]) <= 1e-6 + 1e-4 * abs(fd), (k, i, fd, gflat[i])


def test_hash_and_dense_index():
    cfg = nn.Config(levels=16, log2_table=12, n_min=16, n_max=512)
    g = np.array([[3, 5, 7], [0, 0, 0], [511, 200, 17]], np.int64)
    for l in range(16):
        n = cfg.res[l]
        idx = nn.corner_index(cfg, l, np.minimum(g, n))
        gg = np.minimum(g, n)
        if (n + 1) ** 3 <= cfg.T:
            assert np.array_equal(idx, gg[:, 0] + gg[:, 1] * (n + 1) + gg[:, 2] * (n + 1) ** 2)
        else:
            for row, v in zip(gg, idx):
                h = (int(row[0]) * 1) ^ (int(row[1]) * 2654435761) ^ (int(row[2]) * 805459861)
                assert v == (h & 0xFFFFFFFF) % cfg.T
    # resolutions follow N_min * b^l with b = (N_max / N_min)^(1/(L-1))
    assert cfg.res[0] == 16 and cfg.res[-1] == 512


def test_encoding_reproduces_affine_table():
    """A table that stores an affine function of the grid vertex at a dense
    level interpolates to the same affine function of x (trilinear is exact
    for affine fields)."""
    cfg = nn.Config(levels=1, log2_table=12, n_min=8, n_max=8)
    n = cfg.res[0]
    table = np.zeros((1, cfg.T, 2))
    v = np.array(np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij")).reshape(3, -1).T
    idx = nn.corner_index(cfg, 0, v)
    table[0, idx, 0] = 0.25 * v[:, 0] - 0.5 * v[:, 1] + 0.125 * v[:, 2] + 1.0
    table[0, idx, 1] = v[:, 2] * 2.0
    x = np.random.default_rng(3).uniform(0, 1, (500, 3))
    enc, _ = nn.encode(cfg, table, x, rnd=nn.same, dt=np.float64)
    p = x * n
    assert np.allclose(enc[:, 0], 0.25 * p[:, 0] - 0.5 * p[:, 1] + 0.125 * p[:, 2] + 1.0, atol=1e-12)
    assert np.allclose(enc[:, 1], 2.0 * p[:, 2], atol=1e-12)
    assert np.all(enc[:, 2:] == 0)


def test_adam_first_step_closed_form():
    """After one step from zero moments, m^ = g and v^ = g^2, so every
    parameter moves by -lr * g / (|g| + eps)."""
    rng = np.random.default_rng(5)
    p = rng.standard_normal(100).astype(np.float32)
    g = rng.standard_normal(100).astype(np.float32)
    m = np.zeros(100, np.float32)
    v = np.zeros(100, np.float32)
    p0 = p.copy()
    nn.adam(p, g, m, v, 1, 0.01)
    assert np.allclose(p, p0 - 0.01 * g / (np.abs(g) + 1e-10), rtol=1e-5, atol=1e-7)
