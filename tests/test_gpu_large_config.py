"""The large configuration at its full size against the oracle.

5-level shock-shell AMR (base 256^3 of width 16 over [0,4096]^3, 383M leaf
cells), its dual mesh (about 63M stitching elements, 16% of the leaf cells;
0.97M gridlets with 612M corner scalars), 64^3 majorant grid, the 3840x2160 main view of the large
workload at 1 spp: the oracle's majorant grid and per-primitive culling
against the library's, then every eighth image row (270 rows, 1.04M pixels)
against the oracle's accumulation and frame buffers, bit for bit.

The scene build alone takes over a minute (about 110 s for the whole file on
one MI355X), so the test carries its own time limit.
"""
import time

import numpy as np
import pytest

from oracle import oracle as ora
from owlexastitcher_amd import AmrScene, DualMesh, ExaStitchModel, OWLRenderer, camera_look_at, tf_rasterize

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

TF_SHOCK = np.array([
    [0.00, 0.1, 0.1, 0.3, 0.00],
    [0.12, 0.2, 0.3, 0.9, 0.02],
    [0.35, 0.1, 0.8, 0.9, 0.10],
    [0.60, 0.9, 0.6, 0.2, 0.40],
    [1.00, 1.0, 0.9, 0.8, 0.90],
], np.float32)
W, H = 3840, 2160
ROWS = [(y, y + 1) for y in range(3, H, 8)]


def test_large_main_view_rows_match_oracle(device):
    import ctypes as C
    from concurrent.futures import ThreadPoolExecutor
    import torch
    print("case: large shock scene (base 256^3, 5 levels), 3840x2160 main view, 1 spp", flush=True)
    t = time.time()
    amr = AmrScene.shock(256, 4, 1200.0, 40.0, 0.74, 0.43, 1)
    ncells = len(amr)
    dm = DualMesh(amr)
    del amr
    idx, verts, gl, gs = dm.arrays()
    model = ExaStitchModel.from_dual_mesh(dm)
    del dm
    print(f"  scene: {ncells} cells, {len(idx)} elements, {len(gl)} gridlets ({time.time() - t:.0f}s)", flush=True)
    assert ncells > 370_000_000 and len(idx) > 0.14 * ncells
    t = time.time()
    r = OWLRenderer(model, (64, 64, 64))
    del model
    b = r.modelBounds
    c = 0.5 * (b[:3] + b[3:])
    rgba = tf_rasterize(TF_SHOCK, 128)
    vr = r.valueRange
    r.setColorMap(rgba)
    r.setRange(vr[0], vr[1])
    r.setOpacityScale(100.0 + np.log(0.05) / np.log(1.1))
    r.setMaxBounces(4)
    r.setDirectionalLight(0, (0.3, 1.0, 0.5), 2.0, True)
    r.resize(W, H)
    r.setCamera(*camera_look_at(c + np.array([-0.5, 0.45, 1.0]) * 1.9 * 1200.0, c, (0, 1, 0), 50.0, W, H))
    fb = torch.zeros((H, W), dtype=torch.int32, device=device)
    st = r.state()
    r.render(fb)
    torch.cuda.synchronize()
    acc = r.accum()
    print(f"  renderer build + frame ({time.time() - t:.0f}s)", flush=True)
    t = time.time()
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
    bins = om.bins(64)
    print(f"  oracle model, culling, majorants ({time.time() - t:.0f}s)", flush=True)
    t = time.time()
    p = ora.render_params_from_state(st, maj, rgba)
    acc_o = np.zeros((H, W, 4), np.float32)
    fb_o = np.zeros((H, W), np.uint32)
    tasks = [(y, y + 1) for y0, y1 in ROWS for y in range(y0, y1)]
    with ThreadPoolExecutor(16) as ex:
        list(ex.map(lambda yy: ora.lib().ora_render_f32(C.byref(p), om.ref, bins, ora.ptr(acc_o), ora.ptr(fb_o),
                                                         yy[0], yy[1]), tasks))
    print(f"  oracle rows: {len(tasks)} x {W} pixels ({time.time() - t:.0f}s)", flush=True)
    rows = np.concatenate([np.arange(y0, y1) for y0, y1 in ROWS])
    npx = np.count_nonzero(np.any(acc[rows] != acc_o[rows], axis=-1))
    assert npx == 0, f"differing pixels {npx}/{len(rows) * W}"
    assert np.array_equal(fb.cpu().numpy().view(np.uint32)[rows], fb_o[rows])
    assert np.count_nonzero(np.any(acc[rows] != acc[:1, :1], axis=-1)) > len(rows) * W // 10
