"""Neural field sampler on the GPU against the CPU oracle (oracle/neural.py):
forward values, loss and every gradient (bf16 MFMA with float32
accumulation vs the same rounding on the CPU), one Adam step; then training,
per-macrocell validation and rendering with the network in the element stage.

Tolerances: the GPU and the oracle round the same values to bf16 and sum in
float32 in different orders; gradients agree to about 1e-7 whole-array
relative error, so they are held to 1e-4 (a 1% scale error in one layer
fails), forward values to 2e-3 (a bf16 rounding flips now and then).
Training and rendering are judged against the exact sampler at the
configuration's validation bound, 1e-3 of the value range: validated
macrocells' majorants are the TF maxima over the exact value range widened
by the bound, and the image with the network stays within the noise of two
exact images.
"""
import numpy as np
import pytest

from oracle import neural as onn
from owlexastitcher_amd import NeuralField, OWLRenderer, RENDER_PERSISTENT, camera_look_at
from tests import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg1():
    return scenes.config1()


@pytest.fixture(scope="module")
def renderer(cfg1, device):
    r = OWLRenderer(cfg1.model, (32, 32, 32))
    W, H = 64, 48
    b = r.modelBounds
    center = 0.5 * (b[:3] + b[3:])
    span = b[3:] - b[:3]
    org, d00, du, dv = camera_look_at(center + np.array([-0.3, 0.7, 1.0]) * span, center, (0, 1, 0), 70.0, W, H)
    r.resize(W, H)
    r.setCamera(org, d00, du, dv)
    r.setColorMap(scenes.tf128())
    vr = r.valueRange
    r.setRange(vr[0], vr[1])
    r.setRelDomain(0.0, 100.0)
    r.setOpacityScale(115.0)
    r.setMaxBounces(4)
    r.setDirectionalLight(0, (0.3, 1.0, 0.5), 2.0, True)
    r.setScheme(RENDER_PERSISTENT)
    return r


def oracle_cfg(field):
    c = field.config()
    cfg = onn.Config(levels=c["levels"], log2_table=c["log2T"])
    cfg.res = c["res"]
    return cfg


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


def random_params(cfg, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-0.5, 0.5, (cfg.L, cfg.T, 2)).astype(np.float32),
            rng.uniform(-0.4, 0.4, (32, 64)).astype(np.float32),
            rng.uniform(-0.3, 0.3, (64, 64)).astype(np.float32),
            rng.uniform(-0.3, 0.3, 64).astype(np.float32)]


def test_gradients_match_oracle(renderer):
    print("case: neural gradients, 16 levels x 2^12, 256 rows (some without target)", flush=True)
    f = NeuralField(renderer, levels=16, log2T=12, n_min=16, n_max=128, seed=3)
    cfg = oracle_cfg(f)
    params = random_params(cfg, 1)
    f.setParams(*params)
    rng = np.random.default_rng(2)
    n = 256
    x = rng.uniform(0.0, 1.0, (n, 3)).astype(np.float32)
    t = rng.uniform(0.0, 1.0, n).astype(np.float32)
    t[rng.integers(0, n, 20)] = np.nan
    loss, (dt, d1, d2, d3) = f.gradients(x, t)
    valid = ~np.isnan(t)
    loss_o, (dt_o, d1_o, d2_o, d3_o) = onn.loss_and_grads(cfg, params, x, np.where(valid, t, 0), valid)
    print(f"loss {loss:.6g} vs {loss_o:.6g}; rel dW1 {rel(d1, d1_o):.3g} dW2 {rel(d2, d2_o):.3g} "
          f"dw3 {rel(d3, d3_o):.3g} dtable {rel(dt, dt_o):.3g}")
    assert abs(loss - loss_o) <= 1e-4 * loss_o
    for g, go in ((d1, d1_o), (d2, d2_o), (d3, d3_o), (dt, dt_o)):
        assert rel(g, go) < 1e-4
        assert np.corrcoef(g.ravel(), go.ravel())[0, 1] > 0.999999
    # rows without a target contribute nothing: the same batch without them
    keep = np.flatnonzero(valid)[:192]
    loss2, g2 = f.gradients(x[keep], t[keep])
    assert np.isfinite(loss2)


def test_forward_matches_oracle(renderer, device):
    import torch
    f = NeuralField(renderer, levels=16, log2T=12, n_min=16, n_max=128, seed=3)
    cfg = oracle_cfg(f)
    params = random_params(cfg, 4)
    f.setParams(*params)
    b = renderer.modelBounds  # voxel space == world space here
    rng = np.random.default_rng(5)
    n = 1000
    pos = rng.uniform(b[:3], b[3:], (n, 3)).astype(np.float32)
    d_pos = torch.from_numpy(pos).to(device)
    d_out = torch.empty(n, dtype=torch.float32, device=device)
    f.evaluate(d_pos, d_out)
    got = d_out.cpu().numpy()
    vr = renderer.valueRange
    x01 = np.clip((pos - b[:3]) / (b[3:] - b[:3]), 0, 1)
    y, _ = onn.forward(cfg, params, x01)
    want = vr[0] + y * (vr[1] - vr[0])
    print(f"forward rel err {rel(got, want):.3g}")
    assert rel(got - vr[0], want - vr[0]) < 2e-3
    assert np.corrcoef(got, want)[0, 1] > 0.99999


def test_adam_step_matches_oracle(renderer):
    f = NeuralField(renderer, levels=16, log2T=10, n_min=16, n_max=128, seed=9)
    cfg = oracle_cfg(f)
    p0 = random_params(cfg, 6)
    f.setParams(*p0)
    rng = np.random.default_rng(7)
    x = rng.uniform(0, 1, (128, 3)).astype(np.float32)
    t = rng.uniform(0, 1, 128).astype(np.float32)
    _, grads = f.gradients(x, t)
    f.applyAdam(1e-2)
    p1 = f.params()
    for p, g, q in zip(p0, grads, p1):
        pp, m, v = p.copy(), np.zeros_like(p), np.zeros_like(p)
        onn.adam(pp, g, m, v, 1, 1e-2)
        assert np.allclose(q, pp, rtol=1e-5, atol=1e-7)


@pytest.fixture(scope="module")
def trained(renderer):
    f = NeuralField(renderer, levels=16, log2T=16, n_min=16, n_max=128, seed=1)
    nvalid = f.makePool(1 << 20, 0.7, seed=7)
    assert nvalid > 0.9 * (1 << 20)
    loss = f.train(3000, batch=1 << 14, lr=1e-2, lr_end=1e-4)
    return f, loss


def test_training_reduces_loss(trained):
    f, loss = trained
    print(f"loss first {loss[:5].mean():.3g}, last {loss[-20:].mean():.3g}", flush=True)
    assert np.all(np.isfinite(loss))
    assert loss[-20:].mean() < 0.05 * loss[:5].mean()


def test_validate_and_render_with_network(trained, renderer, device):
    """Per-macrocell validation, then the network answers the element stage
    of the accepted macrocells; the image stays within the seed noise of the
    exact sampler's."""
    import torch
    from oracle import oracle as ora
    f, _ = trained
    vr = renderer.valueRange
    v = f.validate(points_per_mc=64, bound=1e-3 * float(vr[1] - vr[0]))
    print(f"validation: {v}", flush=True)
    assert v["neural_mcs"] > 0 and v["max_error"] <= v["bound"]
    flags = f.flags(32 ** 3)
    assert flags.sum() == v["neural_mcs"]
    # the majorants the renderer uses with the network: the TF maximum over
    # the exact value range widened by the bound in validated macrocells
    # (network values are clamped into that range), the exact one elsewhere
    st = renderer.state()
    rgba = scenes.tf128()
    xr = np.array(st.tfDomain, np.float32)
    ranges = renderer.sampler.value_ranges()
    wid = ranges.copy()
    sel = (flags == 1) & (ranges[:, 0] <= ranges[:, 1])
    b32 = np.float32(v["bound"])
    wid[sel, 0] = ranges[sel, 0] - b32
    wid[sel, 1] = ranges[sel, 1] + b32
    renderer.setNeural(f)
    got = renderer.sampler.max_opacities()
    renderer.setNeural(None)
    assert np.array_equal(got, ora.map_max_opacity(wid, rgba, xr))
    assert np.array_equal(renderer.sampler.max_opacities(), ora.map_max_opacity(ranges, rgba, xr))
    W, H = renderer.size
    fb = torch.zeros((H, W), dtype=torch.int32, device=device)
    K = 8
    renderer.setNeural(None)
    renderer.accumID = 0
    for _ in range(K):
        renderer.render(fb)
    torch.cuda.synchronize()
    s1 = renderer.accum().astype(np.float64)
    for _ in range(K):
        renderer.render(fb)
    torch.cuda.synchronize()
    s2 = renderer.accum().astype(np.float64)
    A, B = s1 / K, (s2 - s1) / K
    renderer.setNeural(f)
    renderer.accumID = 0
    for _ in range(K - 1):
        renderer.render(fb)
    c = renderer.render_counted(fb)
    torch.cuda.synchronize()
    N = renderer.accum().astype(np.float64) / K
    renderer.setNeural(None)
    noise = np.sqrt(np.mean((B - A)[..., :3] ** 2))
    err = np.sqrt(np.mean((N - A)[..., :3] ** 2))
    print(f"counters {c}; RMSE neural vs exact {err:.4g}, exact seed noise {noise:.4g}", flush=True)
    assert c["neural"] > 0
    assert np.all(np.isfinite(N))
    assert err <= noise


def test_batch_with_network_matches_single_view(trained, renderer, device):
    """The batched launch runs the network in the element stage of the
    flagged macrocells exactly as the single-view renderer does, with TF
    majorants widened per transfer function; removing the network restores
    the exact batch images."""
    import torch
    from owlexastitcher_amd import RenderBatch
    f, _ = trained
    vr = renderer.valueRange
    f.validate(points_per_mc=64, bound=1e-3 * float(vr[1] - vr[0]))
    W, H = renderer.size
    st = renderer.state()
    cam = (np.array(st.camOrg), np.array(st.camDir00), np.array(st.camDu), np.array(st.camDv))
    b = RenderBatch(renderer)
    t = b.addTransferFunction(scenes.tf128(), st.tfDomain[0], st.tfDomain[1], st.opacityScale)
    b.addRequest(W, H, 1, t, cam, 0)
    b.addRequest(33, 21, 2, t, cam, 0)

    def single(neural):
        renderer.setNeural(f if neural else None)
        renderer.setSpp(1)
        renderer.accumID = 0
        fb = torch.zeros((H, W), dtype=torch.int32, device=device)
        for _ in range(2):
            renderer.render(fb)
        torch.cuda.synchronize()
        return renderer.accum()

    want_net = single(True)  # leaves the network on the renderer
    b.resetAccum()
    c = b.render_counted()
    b.render()
    torch.cuda.synchronize()
    print(f"batch counters with the network: {c}", flush=True)
    assert c["neural"] > 0
    assert np.array_equal(b.accum(0), want_net)
    want_exact = single(False)
    b.resetAccum()
    c = b.render_counted()
    b.render()
    torch.cuda.synchronize()
    assert c["neural"] == 0
    assert np.array_equal(b.accum(0), want_exact)
