"""CPU oracle of the neural field sampler (hash-grid encoding + MLP).

TEST INFRASTRUCTURE ONLY: imported by tests/ to check the GPU kernels; the
product never loads it.  The reference has no neural sampler; this restates
the published method the product implements:

- multiresolution hash encoding (Mueller, Evans, Schied, Keller, "Instant
  Neural Graphics Primitives with a Multiresolution Hash Encoding", SIGGRAPH
  2022): L levels, level l has resolution N_l = floor(N_min * b^l) with
  b = exp((ln N_max - ln N_min) / (L - 1)); a position x in [0,1]^3 is scaled
  to x * N_l, the 8 surrounding grid vertices index a table of T entries of F
  features (dense index when (N_l + 1)^3 <= T, else the spatial hash
  (x * 1) ^ (y * 2654435761) ^ (z * 805459861) mod T), and the features are
  trilinearly interpolated;
- an MLP 32 -> 64 -> 64 -> 1 without biases, ReLU on the hidden layers;
- loss = mean squared error on the value normalised to the value range;
- Adam (Kingma & Ba 2015) with bias correction.

Mixed precision exactly as the GPU kernels: the MLP inputs of every matrix
product (encodings, weights, hidden activations, their gradients) are rounded
to bfloat16 (round to nearest even) and products are accumulated in float32;
the encoding and the hash-table gradients are float32.
"""
import numpy as np

PRIMES = (1, 2654435761, 805459861)


def bf16(x):
    """Round float32 to the nearest bfloat16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


class Config:
    def __init__(self, levels=16, log2_table=19, features=2, n_min=16, n_max=512, hidden=64):
        assert features == 2 and levels <= 16
        self.L, self.T, self.F = levels, 1 << log2_table, features
        self.n_min, self.n_max, self.hidden = n_min, n_max, hidden
        b = np.exp((np.log(np.float64(n_max)) - np.log(np.float64(n_min))) / (levels - 1)) if levels > 1 else 1.0
        self.res = [int(np.floor(n_min * b ** l)) for l in range(levels)]

    @property
    def n_in(self):
        return 32  # the MLP input width; levels beyond L contribute zeros


def corner_index(cfg, l, g):
    """g: (n, 3) int64 grid vertex coordinates at level l -> table index."""
    n = cfg.res[l]
    if (n + 1) ** 3 <= cfg.T:
        return (g[:, 0] + g[:, 1] * (n + 1) + g[:, 2] * (n + 1) * (n + 1)).astype(np.int64)
    h = (g[:, 0].astype(np.uint64) * np.uint64(PRIMES[0])) ^ (g[:, 1].astype(np.uint64) * np.uint64(PRIMES[1])) \
        ^ (g[:, 2].astype(np.uint64) * np.uint64(PRIMES[2]))
    return ((h & np.uint64(0xFFFFFFFF)) % np.uint64(cfg.T)).astype(np.int64)


def same(x):
    return x


def encode(cfg, table, x, rnd=bf16, dt=np.float32):
    """table: (L, T, F) master values (read through rnd); x: (n, 3) in [0,1].
    Returns (n, 32) and the per-level corner indices / weights for the
    backward pass.  rnd=same, dt=float64 gives the unrounded reference."""
    n = len(x)
    out = np.zeros((n, 32), dt)
    tb = rnd(table).astype(dt)
    cache = []
    for l in range(cfg.L):
        nl = dt(cfg.res[l])
        p = x.astype(dt) * nl
        g0f = np.clip(np.floor(p), 0, nl - 1).astype(dt)
        g0 = g0f.astype(np.int64)
        w = (p - g0f).astype(dt)
        idxs, wts = [], []
        for c in range(8):
            off = np.array([(c >> 0) & 1, (c >> 1) & 1, (c >> 2) & 1])
            idx = corner_index(cfg, l, g0 + off)
            wx = np.where(off[0], w[:, 0], dt(1) - w[:, 0])
            wy = np.where(off[1], w[:, 1], dt(1) - w[:, 1])
            wz = np.where(off[2], w[:, 2], dt(1) - w[:, 2])
            wc = ((wx * wy).astype(dt) * wz).astype(dt)
            out[:, l * cfg.F:(l + 1) * cfg.F] += wc[:, None] * tb[l, idx]
            idxs.append(idx)
            wts.append(wc)
        cache.append((idxs, wts))
    return out, cache


def forward(cfg, params, x, rnd=bf16, dt=np.float32):
    table, W1, W2, w3 = params
    enc, cache = encode(cfg, table, x, rnd, dt)
    X = rnd(enc).astype(dt)
    H1 = rnd(np.maximum(X @ rnd(W1).astype(dt), 0).astype(dt)).astype(dt)
    H2 = rnd(np.maximum(H1 @ rnd(W2).astype(dt), 0).astype(dt)).astype(dt)
    y = (H2 @ rnd(w3).astype(dt)).astype(dt)
    return y, (enc, cache, X, H1, H2)


def loss_and_grads(cfg, params, x, target, weight=None, rnd=bf16, dt=np.float32):
    """MSE over the batch (weight: 0/1 per sample, counted in the mean)."""
    table, W1, W2, w3 = params
    y, (enc, cache, X, H1, H2) = forward(cfg, params, x, rnd, dt)
    n = len(x)
    wgt = np.ones(n, dt) if weight is None else weight.astype(dt)
    r = (y - target.astype(dt)) * wgt
    loss = float(np.sum(r.astype(np.float64) ** 2) / n)
    gy = (dt(2.0 / n) * r).astype(dt)
    dw3 = (H2.T @ gy).astype(dt)
    dH2 = rnd(np.where(H2 > 0, gy[:, None] * rnd(w3).astype(dt)[None, :], 0).astype(dt)).astype(dt)
    dW2 = (H1.T @ dH2).astype(dt)
    dH1 = rnd(np.where(H1 > 0, dH2 @ rnd(W2).astype(dt).T, 0).astype(dt)).astype(dt)
    dW1 = (X.T @ dH1).astype(dt)
    dX = (dH1 @ rnd(W1).astype(dt).T).astype(dt)
    dtable = np.zeros(table.shape, dt)
    for l, (idxs, wts) in enumerate(cache):
        for idx, wc in zip(idxs, wts):
            for f in range(cfg.F):
                np.add.at(dtable[l, :, f], idx, wc * dX[:, l * cfg.F + f])
    return loss, (dtable, dW1, dW2, dw3)


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.99, eps=1e-10):
    """One Adam step (step counts from 1); updates p, m, v in place."""
    m *= np.float32(b1)
    m += np.float32(1 - b1) * g
    v *= np.float32(b2)
    v += np.float32(1 - b2) * g * g
    mh = m / np.float32(1 - b1 ** step)
    vh = v / np.float32(1 - b2 ** step)
    p -= np.float32(lr) * mh / (np.sqrt(vh) + np.float32(eps))
