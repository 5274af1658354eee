"""GPU parity of the ExaBricks comparison sampler against the oracle.

- point location: the reference's 2-brick data at hand-computed points;
  random points and points on region faces give values bit-identical to the
  oracle's region blend (and so to the brute-force blend over all bricks),
  and misses exactly where no region that is not culled holds the point;
- per-region max opacities equal the oracle's mapping of the region value
  ranges;
- frames of the path tracer, direct lighting and the ExaBricks ray marcher:
  accumulation and frame buffers bit for bit against the oracle, whose region
  walk is a brute-force closest hit over all regions.
"""
import numpy as np
import pytest

from oracle import oracle as ora
from owlexastitcher_amd import AmrScene, ExaBrickModel, ExaBrickSampler, OWLRenderer, camera_look_at
from tests import scenes
from tests.test_bricks import TEST_BRICKS, TEST_SCALARS

pytestmark = pytest.mark.gpu


def _scene(kind):
    if kind == "gauss":
        return AmrScene.gaussians(16, 1, 4, 12, 4, 1)
    return AmrScene.noise(8, 2, 0.45, 0.6, 4, 2.0, 3)


def _locate(sampler, pts, device):
    import torch
    d_pos = torch.from_numpy(np.ascontiguousarray(pts, np.float32)).to(device)
    d_out = torch.full((len(pts), 3), 7, dtype=torch.int32, device=device)
    sampler.locate(d_pos, d_out)
    torch.cuda.synchronize()
    return d_out.cpu().numpy()


def test_locate_two_brick_data(device):
    print("case: brick locate, 2-brick data", flush=True)
    m = ExaBrickModel.from_arrays(TEST_BRICKS, TEST_SCALARS)
    s = ExaBrickSampler(m)
    pts = np.array([[1, 1, 1], [3, 1, 1], [3.25, 1, 1], [2, 1, 1], [10, 1, 1]], np.float32)
    out = _locate(s, pts, device)
    assert out[:4, 0].tolist() == [0] * 4 and out[:4, 1].tolist() == [-1] * 4
    assert out[:4, 2].view(np.float32).tolist() == [0.5, 0.5, 0.75, 0.25]
    assert out[4].tolist() == [0, 0, 0]


@pytest.mark.parametrize("kind,block", [("gauss", 8), ("noise", 4)])
def test_locate_matches_oracle(device, kind, block):
    print(f"case: brick locate {kind} block {block}", flush=True)
    m = ExaBrickModel.from_amr(_scene(kind), block)
    bricks, scalars, _, _ = m.arrays()
    sc = ora.BrickScene(bricks, scalars)
    s = ExaBrickSampler(m)
    cb, _ = sc.bounds()
    rng = np.random.default_rng(17)
    p = rng.uniform(cb[:3] - 2, cb[3:] + 2, size=(60000, 3)).astype(np.float32)
    d = sc.abrs["domain"]
    k = rng.integers(0, len(d), 20000)
    q = rng.uniform(d[k, :3], d[k, 3:]).astype(np.float32)
    ax = rng.integers(0, 3, 20000)
    q[np.arange(20000), ax] = d[k, ax + 3 * rng.integers(0, 2, 20000)]
    p = np.concatenate([p, q])
    out = _locate(s, p, device)
    ref = sc.locate(p)
    assert np.array_equal(out, ref)
    v, w = ora.brick_blend_brute(bricks, scalars, p)
    hit = out[:, 1] == -1
    assert np.array_equal(hit, w > 0)
    assert np.array_equal(out[hit, 2].view(np.float32), v[hit])
    # with culling: a TF opaque only over the upper half of the values
    rgba = np.zeros((128, 4), np.float32)
    rgba[64:, 3] = 1.0
    vr = sc.value_ranges()
    xr = np.array([vr[:, 0].min(), vr[:, 1].max()], np.float32)
    s.computeMaxOpacities(rgba, xr)
    mo = ora.map_max_opacity(vr, rgba, xr)
    assert np.array_equal(s.max_opacities(), mo)
    assert 0 < (mo == 0).sum() < len(mo)
    sc.set_max_opacity(mo)
    out = _locate(s, p, device)
    assert np.array_equal(out, sc.locate(p))


def make_renderer(model, W, H, opacity=115.0, tf=scenes.TF5):
    r = OWLRenderer(model, (16, 16, 16))
    b = r.modelBounds
    center = 0.5 * (b[:3] + b[3:])
    span = b[3:] - b[:3]
    org, d00, du, dv = camera_look_at(center + np.array([-0.3, 0.7, 1.0]) * span, center, (0, 1, 0), 60.0, W, H)
    r.resize(W, H)
    r.setCamera(org, d00, du, dv)
    r.setColorMap(scenes.tf128(tf))
    vr = r.valueRange
    r.setRange(vr[0], vr[1])
    r.setRelDomain(0.0, 100.0)
    r.setOpacityScale(opacity)
    return r


def render_both(model, r, frames, device):
    import torch
    bricks, scalars, _, _ = model.arrays()
    sc = ora.BrickScene(bricks, scalars)
    rgba = scenes.tf128()
    st = r.state()
    xr = np.array(st.tfDomain, np.float32)
    mo = ora.map_max_opacity(sc.value_ranges(), rgba, xr)
    assert np.array_equal(mo, r.sampler.max_opacities())
    sc.set_max_opacity(mo)
    fb = torch.zeros((st.H, st.W), dtype=torch.int32, device=device)
    acc_o = None
    unused = np.zeros(1, np.float32)
    for _ in range(frames):
        st = r.state()
        p = ora.render_params_from_state(st, unused, rgba)
        acc_o, fb_o = ora.render_bricks_f32(p, sc, acc_o)
        r.render(fb)
    torch.cuda.synchronize()
    return r.accum(), acc_o, fb.cpu().numpy().view(np.uint32), fb_o


def _report(a, b):
    d = np.abs(a.astype(np.float64) - b.astype(np.float64))
    npx = np.count_nonzero(np.any(a != b, axis=-1))
    return f"differing pixels {npx}/{a.shape[0] * a.shape[1]}, max abs diff {d.max():.3g}"


@pytest.mark.parametrize("kind", ["gauss", "noise"])
@pytest.mark.parametrize("integrator,light", [(0, "dir"), (1, "point"), (2, "none"), (0, "none")])
def test_render_matches_oracle(device, kind, integrator, light):
    print(f"case: brick render {kind} integrator={integrator} light={light} 48x40", flush=True)
    model = ExaBrickModel.from_amr(_scene(kind), 8)
    r = make_renderer(model, 48, 40)
    r.setType(integrator)
    r.setMaxBounces(4)
    if integrator == 2:
        r.setDt(0.5)
    b = r.modelBounds
    if light == "dir":
        r.setDirectionalLight(0, (0.3, 1.0, 0.5), 2.0, True)
    elif light == "point":
        r.setLightSource(0, b[3:] + 10.0, 60.0 * 500.0, True)
    acc, acc_o, fb, fb_o = render_both(model, r, 2, device)
    print(_report(acc, acc_o))
    assert np.array_equal(acc, acc_o), _report(acc, acc_o)
    assert np.array_equal(fb, fb_o)
    bgmask = np.all(acc == acc[:, :1], axis=(1, 2))
    assert not bgmask.all()


def test_render_counted_and_state_errors(device):
    print("case: brick render counters", flush=True)
    model = ExaBrickModel.from_amr(_scene("gauss"), 8)
    r = make_renderer(model, 32, 24)
    r.setMaxBounces(4)
    c = r.render_counted()
    assert c["pixels"] == 32 * 24
    assert c["samples"] > 0 and c["tf"] == c["samples"]
    assert c["bricks"] >= c["samples"] and c["kd_leaves"] >= c["regions"] > 0
    from owlexastitcher_amd import ExaError
    with pytest.raises(ExaError):
        r.setAccel(1)


@pytest.mark.parametrize("integrator", [0, 2])
def test_render_clip_remap_spp(device, integrator):
    """The brick renderer with a voxel-space remap (Model::setVoxelSpaceTransform),
    a clip plane and 2 spp per frame, against the oracle."""
    print(f"case: brick render integrator={integrator} remap+clip+2spp 40x32", flush=True)
    model = ExaBrickModel.from_amr(_scene("noise"), 8)
    cb, _ = model.bounds()
    remap_from = cb
    remap_to = np.concatenate([cb[:3] * 0.5 + 3.0, cb[3:] * 0.5 + 3.0]).astype(np.float32)
    r = OWLRenderer(model, (16, 16, 16), remap_from, remap_to)
    W, H = 40, 32
    b = r.modelBounds
    center = 0.5 * (b[:3] + b[3:])
    span = b[3:] - b[:3]
    org, d00, du, dv = camera_look_at(center + np.array([0.6, 0.5, -1.1]) * span, center, (0, 1, 0), 55.0, W, H)
    r.resize(W, H)
    r.setCamera(org, d00, du, dv)
    r.setColorMap(scenes.tf128())
    vr = r.valueRange
    r.setRange(vr[0], vr[1])
    r.setRelDomain(0.0, 100.0)
    r.setOpacityScale(118.0)
    r.setType(integrator)
    r.setMaxBounces(3)
    r.setSpp(2)
    r.setClipPlane(0, True, (0.3, 0.0, 1.0), float(0.3 * center[0] + center[2]))
    r.setDirectionalLight(0, (0.2, 1.0, 0.4), 1.5, True)
    acc, acc_o, fb, fb_o = render_both(model, r, 2, device)
    assert np.array_equal(acc, acc_o), _report(acc, acc_o)
    assert np.array_equal(fb, fb_o)
    assert not np.all(acc == acc[:, :1])


@pytest.mark.parametrize("integrator", [0, 1])
def test_persistent_matches_megakernel(device, integrator):
    """The persistent brick renderer (lanes refilled from a pixel counter) and
    the one-thread-per-pixel kernel give the same frames and counters; a
    size with partial edge tiles, 2 spp, a point and a directional light."""
    import torch
    from owlexastitcher_amd import RENDER_MEGAKERNEL, RENDER_PERSISTENT
    print(f"case: brick persistent vs megakernel, integrator={integrator}, 83x61, 2 spp", flush=True)
    model = ExaBrickModel.from_amr(_scene("noise"), 8)
    r = make_renderer(model, 83, 61)
    r.setType(integrator)
    r.setMaxBounces(6)
    r.setSpp(2)
    r.setDirectionalLight(0, (0.3, 1.0, 0.5), 2.0, True)
    r.setLightSource(1, r.modelBounds[3:] + 5.0, 4.0e4, True)
    fb = torch.zeros((61, 83), dtype=torch.int32, device=device)
    out = {}
    for scheme in (RENDER_PERSISTENT, RENDER_MEGAKERNEL):
        r.setScheme(scheme)
        r.accumID = 0
        r.render(fb)
        c = r.render_counted(fb)
        torch.cuda.synchronize()
        out[scheme] = (r.accum(), fb.cpu().numpy().copy(), c)
    a0, f0, c0 = out[RENDER_PERSISTENT]
    a1, f1, c1 = out[RENDER_MEGAKERNEL]
    assert np.array_equal(a0, a1), _report(a0, a1)
    assert np.array_equal(f0, f1)
    assert c0 == c1 and c0["pixels"] == 83 * 61
    assert not np.all(a0 == a0[:, :1])

