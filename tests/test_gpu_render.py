"""GPU parity of the renderers (megakernel and persistent) against the
single-precision oracle.

With the same per-pixel seed (pixelID, accumID), the same LCG<4> stream and
the same operation order, the accumulation buffer is compared bit for bit.
"""
import numpy as np
import pytest

from oracle import oracle as ora
from owlexastitcher_amd import (ACCEL_BVH, ACCEL_TABLES, RENDER_MEGAKERNEL, RENDER_PERSISTENT, OWLRenderer,
                                camera_look_at)
from tests import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg1():
    return scenes.config1()


def make_renderer(sc, W, H, num_mcs=(32, 32, 32), opacity=115.0, tf=scenes.TF5):
    r = OWLRenderer(sc.model, num_mcs)
    b = r.modelBounds
    center = 0.5 * (b[:3] + b[3:])
    span = b[3:] - b[:3]
    org, d00, du, dv = camera_look_at(center + np.array([-0.3, 0.7, 1.0]) * span, center, (0, 1, 0), 70.0, W, H)
    r.resize(W, H)
    r.setCamera(org, d00, du, dv)
    r.setColorMap(scenes.tf128(tf))
    vr = r.valueRange
    r.setRange(vr[0], vr[1])
    r.setRelDomain(0.0, 100.0)
    r.setOpacityScale(opacity)
    return r


def oracle_model(sc, rgba, xf_range):
    om = ora.Model(sc.verts, sc.idx, sc.gridlets, sc.gscalars)
    gmo = ora.map_max_opacity(om.gridlet_value_ranges(), rgba, xf_range)
    emo = om.element_max_opacity(rgba, xf_range)
    om.set_culling(gmo, emo)
    return om


def render_both(sc, r, frames, device):
    import torch
    st = r.state()
    rgba = scenes.tf128()
    xr = np.array(st.tfDomain, np.float32)
    om = oracle_model(sc, rgba, xr)
    dims = np.array(st.gridDims, np.int32)
    maj = ora.map_max_opacity(om.grid_build(dims, np.array(st.gridBounds, np.float32)), rgba, xr)
    assert np.array_equal(maj, r.sampler.max_opacities())
    fb = torch.zeros((st.H, st.W), dtype=torch.int32, device=device)
    acc_o = None
    for f in range(frames):
        st = r.state()
        p = ora.render_params_from_state(st, maj, rgba)
        acc_o, fb_o = ora.render_f32(p, om, acc_o)
        r.render(fb)
    torch.cuda.synchronize()
    return r.accum(), acc_o, fb.cpu().numpy().view(np.uint32), fb_o


def _report(a, b):
    d = np.abs(a.astype(np.float64) - b.astype(np.float64))
    npx = np.count_nonzero(np.any(a != b, axis=-1))
    return f"differing pixels {npx}/{a.shape[0] * a.shape[1]}, max abs diff {d.max():.3g}"


SCHEMES = [RENDER_PERSISTENT, RENDER_MEGAKERNEL]


@pytest.mark.parametrize("scheme", SCHEMES, ids=["persistent", "megakernel"])
@pytest.mark.parametrize("accel", [ACCEL_TABLES, ACCEL_BVH], ids=["tables", "bvh"])
@pytest.mark.parametrize("integrator,light", [(0, "dir"), (1, "point"), (0, "none")])
def test_render_matches_oracle_f32(cfg1, device, integrator, light, accel, scheme):
    print(f"case: render config1 64x48 integrator={integrator} light={light} accel={accel} scheme={scheme}",
          flush=True)
    r = make_renderer(cfg1, 64, 48)
    r.setAccel(accel)
    r.setScheme(scheme)
    assert r.scheme == scheme
    r.setType(integrator)
    r.setMaxBounces(4)
    b = r.modelBounds
    if light == "dir":
        r.setDirectionalLight(0, (0.3, 1.0, 0.5), 2.0, True)
    elif light == "point":
        r.setLightSource(0, b[3:] + 10.0, 60.0 * 500.0, True)
    acc, acc_o, fb, fb_o = render_both(cfg1, r, 2, device)
    print(_report(acc, acc_o))
    assert np.array_equal(acc, acc_o), _report(acc, acc_o)
    assert np.array_equal(fb, fb_o)
    # the image is not trivial: some pixels scattered in the volume
    bgmask = np.all(acc == acc[:, :1], axis=(1, 2))
    assert not bgmask.all()


@pytest.mark.parametrize("scheme", SCHEMES, ids=["persistent", "megakernel"])
@pytest.mark.parametrize("shade", [1, 2], ids=["gridlets", "teaser"])
def test_shade_modes_match_oracle(cfg1, device, shade, scheme):
    """Shade modes (deviceCode.cu:625-697) recolour samples but keep alpha, so
    the paths (and the alpha channel) are those of the default mode."""
    print(f"case: shade mode {shade}, scheme {scheme}, 64x48", flush=True)
    r = make_renderer(cfg1, 64, 48)
    r.setScheme(scheme)
    r.setMaxBounces(4)
    r.setDirectionalLight(0, (0.3, 1.0, 0.5), 2.0, True)
    acc_d, _, _, _ = render_both(cfg1, r, 1, device)
    r.setShadeMode(shade)
    r.accumID = 0
    acc, acc_o, fb, fb_o = render_both(cfg1, r, 1, device)
    assert np.array_equal(acc, acc_o), _report(acc, acc_o)
    assert np.array_equal(fb, fb_o)
    assert np.array_equal(acc[..., 3], acc_d[..., 3])
    assert not np.array_equal(acc[..., :3], acc_d[..., :3])


@pytest.mark.parametrize("accel", [ACCEL_TABLES, ACCEL_BVH], ids=["tables", "bvh"])
@pytest.mark.parametrize("light,dt", [("none", 1.0), ("point", 0.37), ("none", 2.5)])
def test_ray_marcher_matches_oracle(cfg1, device, accel, light, dt):
    """Ray marcher (deviceCode.cu:1213-1260, 348-398), bit-exact, 2 frames."""
    print(f"case: ray marcher accel={accel} light={light} dt={dt}, 64x48", flush=True)
    r = make_renderer(cfg1, 64, 48)
    r.setAccel(accel)
    r.setType(2)
    r.setDt(dt)
    if light == "point":
        b = r.modelBounds
        r.setLightSource(0, b[3:] + 10.0, 60.0 * 500.0, True)
    acc, acc_o, fb, fb_o = render_both(cfg1, r, 2, device)
    assert np.array_equal(acc, acc_o), _report(acc, acc_o)
    assert np.array_equal(fb, fb_o)
    bgmask = np.all(acc == acc[:, :1], axis=(1, 2))
    assert not bgmask.all()


def test_heat_map(cfg1, device):
    """Heat map (deviceCode.cu:1323-1327): over(heat colour with alpha .5,
    sample sum), so alpha becomes .5 + .5 alpha and the colour stays within
    [.5 c, .5 c + 1]."""
    import torch
    print("case: heat map, 64x48", flush=True)
    r = make_renderer(cfg1, 64, 48)
    r.setMaxBounces(4)
    fb = torch.zeros((48, 64), dtype=torch.int32, device=device)
    r.render(fb)
    torch.cuda.synchronize()
    a0 = r.accum()
    r.setHeatMap(True, 1e-6)
    assert r.accumID == 0
    r.render(fb)
    torch.cuda.synchronize()
    a1 = r.accum()
    assert np.all(np.isfinite(a1))
    assert np.array_equal(a1[..., 3], np.float32(0.5) + np.float32(0.5) * a0[..., 3])
    assert np.all(a1[..., :3] >= np.float32(0.5) * a0[..., :3]) and np.all(a1[..., :3] <= 0.5 * a0[..., :3] + 1.0)
    assert not np.array_equal(a1, a0)
    r.setHeatMap(False)
    r.render(fb)
    torch.cuda.synchronize()
    assert np.array_equal(r.accum(), a0)


@pytest.mark.parametrize("tuning", [(1, 0, 0), (1, 64, 64), (64, 32, 0), (3, 17, 64), (5, 0, 40)])
def test_persistent_tuning_matches_oracle(cfg1, device, tuning):
    """Images do not depend on the persistent renderer's scheduling knobs;
    odd sizes exercise partial 8x8 pixel tiles; 3 spp; clip plane on."""
    print(f"case: persistent tuning {tuning}, 45x29, 3 spp, clip", flush=True)
    r = make_renderer(cfg1, 45, 29)
    r.setScheme(RENDER_PERSISTENT)
    r.setPersistentTuning(*tuning)
    r.setSpp(3)
    r.setMaxBounces(4)
    r.setDirectionalLight(0, (0.3, 1.0, 0.5), 2.0, True)
    b = r.modelBounds
    r.setClipPlane(0, True, (0.0, 0.0, 1.0), float(0.5 * (b[2] + b[5])) + 7.0)
    acc, acc_o, fb, fb_o = render_both(cfg1, r, 1, device)
    assert np.array_equal(acc, acc_o), _report(acc, acc_o)
    assert np.array_equal(fb, fb_o)


def test_persistent_matches_megakernel_noise(device):
    """3-level scene, element-heavy (pyramids, wedges, hexes), 160x120: the
    two launch schemes give identical frames, counters included."""
    import torch
    print("case: persistent vs megakernel, noise scene 160x120", flush=True)
    sc = scenes.noise_small()
    r = make_renderer(sc, 160, 120, num_mcs=(16, 16, 16), opacity=125.0)
    r.setMaxBounces(6)
    r.setDirectionalLight(0, (0.3, 1.0, 0.5), 2.0, True)
    r.setLightSource(1, r.modelBounds[3:] + 5.0, 4.0e4, True)
    fb = torch.zeros((120, 160), dtype=torch.int32, device=device)
    out = {}
    for scheme in SCHEMES:
        r.setScheme(scheme)
        r.accumID = 0
        r.render(fb)
        c1 = r.render_counted(fb)
        torch.cuda.synchronize()
        out[scheme] = (r.accum(), fb.cpu().numpy().copy(), c1)
    a0, f0, c0 = out[RENDER_PERSISTENT]
    a1, f1, c1 = out[RENDER_MEGAKERNEL]
    assert np.array_equal(a0, a1), _report(a0, a1)
    assert np.array_equal(f0, f1)
    assert c0 == c1, (c0, c1)
    assert c0["hex"] > 0 and c0["wedge"] > 0 and c0["pyr"] > 0
    assert c0["pixels"] == 160 * 120


@pytest.mark.parametrize("scheme", SCHEMES, ids=["persistent", "megakernel"])
def test_sub_image_and_selection_match_oracle(cfg1, device, scheme):
    """setSubImage / setSubImageSelection (deviceCode.cu:1285-1296, 1337-1340):
    the frame of a sub-image window with the selection box drawn, bit for bit
    against the oracle."""
    print(f"case: sub-image + selection, scheme {scheme}, 64x48", flush=True)
    r = make_renderer(cfg1, 64, 48)
    r.setScheme(scheme)
    r.setMaxBounces(4)
    r.setDirectionalLight(0, (0.3, 1.0, 0.5), 2.0, True)
    r.setSubImage((0.15, 0.2, 0.65, 0.8), True)
    r.setSubImageSelection((0.1, 0.25, 0.6, 0.75), True)
    acc, acc_o, fb, fb_o = render_both(cfg1, r, 2, device)
    assert np.array_equal(acc, acc_o), _report(acc, acc_o)
    assert np.array_equal(fb, fb_o)
