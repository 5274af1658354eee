"""The timed workload at its full size, bit for bit against the oracle.

The 1080p frame that bench.py times: 3-level fractal-noise AMR (base 128^3 of
width 4, 31.2M leaf cells), its dual mesh (5.9M stitching elements, 106K
gridlets), 64^3 majorant grid, path tracer capped at 4 bounces, one
directional light, the 5-point TF at opacity scale 100.  The majorant grid and
the per-primitive culling come from the oracle's own build and are checked
against the library's; then two frames of each launch scheme are compared with
the oracle's accumulation buffer and RGBA8 frame buffer, every pixel.
"""
import numpy as np
import pytest

from oracle import oracle as ora
from owlexastitcher_amd import (RENDER_MEGAKERNEL, RENDER_PERSISTENT, AmrScene, DualMesh, ExaStitchModel,
                                OWLRenderer, camera_look_at)
from tests import scenes

pytestmark = pytest.mark.gpu

W, H = 1920, 1080


@pytest.fixture(scope="module")
def workload():
    print("case: timed workload build (noise 128^3 base, 1080p)", flush=True)
    amr = AmrScene.noise(128, 2, 0.55, 0.7, 6, 2.0, 1)
    dm = DualMesh(amr)
    model = ExaStitchModel.from_dual_mesh(dm)
    r = OWLRenderer(model, (64, 64, 64))
    b = r.modelBounds
    center = 0.5 * (b[:3] + b[3:])
    span = b[3:] - b[:3]
    r.resize(W, H)
    r.setCamera(*camera_look_at(center + np.array([-0.3, 0.7, 1.0]) * span, center, (0, 1, 0), 70.0, W, H))
    rgba = scenes.tf128()
    r.setColorMap(rgba)
    vr = r.valueRange
    r.setRange(vr[0], vr[1])
    r.setRelDomain(0.0, 100.0)
    r.setOpacityScale(100.0)
    r.setType(0)
    r.setMaxBounces(4)
    r.setSpp(1)
    r.setDirectionalLight(0, (0.3, 1.0, 0.5), 2.0, True)
    r.setPersistentTuning(6, 16, 16)
    idx, verts, gl, gs = dm.arrays()
    assert len(idx) > 5_000_000 and len(gl) > 100_000
    st = r.state()
    xr = np.array(st.tfDomain, np.float32)
    om = ora.Model(verts, idx, gl, gs)
    gmo = ora.map_max_opacity(om.gridlet_value_ranges(), rgba, xr)
    emo = om.element_max_opacity(rgba, xr)
    g_lib, e_lib = r.sampler.prim_max_opacities(len(gl), len(idx))
    assert np.array_equal(gmo, g_lib) and np.array_equal(emo, e_lib)
    om.set_culling(gmo, emo)
    maj = ora.map_max_opacity(om.grid_build(np.array(st.gridDims, np.int32), np.array(st.gridBounds, np.float32)),
                              rgba, xr)
    assert np.array_equal(maj, r.sampler.max_opacities())
    return r, om, maj, rgba


@pytest.mark.parametrize("scheme", [RENDER_PERSISTENT, RENDER_MEGAKERNEL], ids=["persistent", "megakernel"])
def test_timed_frame_matches_oracle(workload, device, scheme):
    import torch
    r, om, maj, rgba = workload
    print(f"case: timed workload 1920x1080, scheme {scheme}, 2 frames", flush=True)
    r.setScheme(scheme)
    r.resetAccum()
    fb = torch.zeros((H, W), dtype=torch.int32, device=device)
    acc_o = None
    for _ in range(2):
        p = ora.render_params_from_state(r.state(), maj, rgba)
        acc_o, fb_o = ora.render_f32(p, om, acc_o, threads=16)
        r.render(fb)
    torch.cuda.synchronize()
    acc = r.accum()
    npx = np.count_nonzero(np.any(acc != acc_o, axis=-1))
    assert npx == 0, f"differing pixels {npx}/{W * H}"
    assert np.array_equal(fb.cpu().numpy().view(np.uint32), fb_o)
    # the view shows the volume, not only the background
    assert np.count_nonzero(np.any(acc != acc[:1, :1], axis=-1)) > W * H // 10
