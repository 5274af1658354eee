"""The timed workload at its full size, bit for bit against the oracle.

The 1080p frame that bench.py times: 3-level fractal-noise AMR (base 128^3 of
width 4, 31.2M leaf cells), its dual mesh (5.9M stitching elements, 106K
gridlets), 64^3 majorant grid, path tracer capped at 4 bounces, one
directional light, the 5-point TF at opacity scale 100.  The majorant grid and
the per-primitive culling come from the oracle's own build and are checked
against the library's; then two frames of each launch scheme are compared with
the oracle's accumulation buffer and RGBA8 frame buffer, every pixel.  The
locate, batched-requests and ExaBricks workloads follow at their full sizes.
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
def amr():
    return AmrScene.noise(128, 2, 0.55, 0.7, 6, 2.0, 1)


@pytest.fixture(scope="module")
def workload(amr):
    print("case: timed workload build (noise 128^3 base, 1080p)", flush=True)
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
    r.setPersistentTuning(4, 16, 24)
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


@pytest.mark.parametrize("scheme,stage", [(RENDER_PERSISTENT, 0), (RENDER_MEGAKERNEL, 0), (RENDER_PERSISTENT, 1),
                                          (RENDER_PERSISTENT, 2)],
                         ids=["persistent", "megakernel", "persistent-wgqueue", "persistent-rounds"])
def test_timed_frame_matches_oracle(workload, device, scheme, stage):
    import torch
    r, om, maj, rgba = workload
    print(f"case: timed workload 1920x1080, scheme {scheme}, element stage {stage}, 2 frames", flush=True)
    r.setScheme(scheme)
    r.setLaunchTuning(stage, 4, 0, 256)
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
    r.setLaunchTuning(0, 1, 0, 256)


def test_locate_workload_matches_oracle(workload, device):
    """The sampler-alone workload: 2^24 points uniformly random over the model
    box (the point set of the locate benchmark, seed 1), located and
    interpolated under the TF's culling, against the oracle's binned search."""
    from concurrent.futures import ThreadPoolExecutor
    from tests.test_gpu_sampler import VALUE_RTOL, _compare, _locate_gpu
    r, om, _, _ = workload
    n = 1 << 24
    print(f"case: locate workload, {n} uniform points, seed 1", flush=True)
    b = r.modelBounds
    pos = np.random.default_rng(1).uniform(b[:3], b[3:], (n, 3)).astype(np.float32)
    ids, vals = _locate_gpu(r.sampler, pos, device)
    chunk = 1 << 20
    om.bins()  # built once, before the threads share it
    with ThreadPoolExecutor(16) as ex:
        ref = np.concatenate(list(ex.map(lambda i: om.locate_binned(pos[i:i + chunk]), range(0, n, chunk))))
    mism, rel, ndiff = _compare(ids, vals, ref)
    assert mism == 0 and rel <= VALUE_RTOL, (mism, rel, ndiff)
    hit = ref[:, 0] >= 0
    assert hit.sum() > n // 2 and np.count_nonzero(hit & (ref[:, 1] < 0)) > 10000


def test_batch_workload_matches_single_view(workload, device):
    """The batched-requests workload: 64 requests of sizes drawn from
    {256, 512, 768, 1024}^2 with seeded orbit cameras and 4 random 5-8 point
    TFs, 16 spp, one persistent launch; four of the requests (the first of
    each size) against the single-view renderer with the same settings, and
    every fourth row of each against the oracle, bit for bit."""
    import ctypes as C
    from concurrent.futures import ThreadPoolExecutor
    import torch
    from owlexastitcher_amd import RenderBatch, tf_rasterize
    r, om, _, rgba0 = workload
    st0 = r.state()
    xr0 = np.array(st0.tfDomain, np.float32)
    r_culling = (ora.map_max_opacity(om.gridlet_value_ranges(), rgba0, xr0), om.element_max_opacity(rgba0, xr0))
    print("case: batch workload, 64 requests, 16 spp, 1 frame", flush=True)
    rng = np.random.default_rng(1)
    vr = r.valueRange
    tfs = []
    for _ in range(4):
        npt = int(rng.integers(5, 9))
        pts = np.zeros((npt, 5), np.float32)
        pts[:, 0] = np.concatenate([[0.0], np.sort(rng.uniform(0.0, 1.0, npt - 2)), [1.0]])
        pts[:, 1:4] = rng.uniform(0.0, 1.0, (npt, 3))
        pts[:, 4] = rng.uniform(0.0, 1.0, npt)
        # the opacity multiplier U[0.5, 4] as a slider value, os = 1.1^(s - 100)
        slider = 100.0 + float(np.log(rng.uniform(0.5, 4.0)) / np.log(1.1))
        r.setOpacityScale(slider)
        tfs.append((tf_rasterize(pts, 128), slider, r.state().opacityScale))
    bnd = r.modelBounds
    c = 0.5 * (bnd[:3] + bnd[3:])
    rad = 1.2 * float(np.linalg.norm(bnd[3:] - bnd[:3]))
    reqs = []
    for _ in range(64):
        size = int(rng.choice([256, 512, 768, 1024]))
        az, el = float(rng.uniform(0, 2 * np.pi)), float(rng.uniform(-0.8, 0.8))
        eye = c + rad * np.array([np.cos(el) * np.cos(az), np.sin(el), np.cos(el) * np.sin(az)])
        reqs.append((size, int(rng.integers(0, 4)), camera_look_at(eye, c, (0, 1, 0), 45.0, size, size)))
    r.setRange(vr[0], vr[1])
    r.setRelDomain(0.0, 100.0)
    bt = RenderBatch(r)
    ti = [bt.addTransferFunction(rgba, vr[0], vr[1], os) for rgba, _, os in tfs]
    for size, tf, cam in reqs:
        bt.addRequest(size, size, 16, ti[tf], cam, 0)
    bt.resetAccum()
    bt.render()
    torch.cuda.synchronize()
    picks = [next(i for i, q in enumerate(reqs) if q[0] == s) for s in (256, 512, 768, 1024)]
    r.setScheme(RENDER_PERSISTENT)
    r.setSpp(16)
    r.setType(0)
    om.bins()
    ranges = None
    try:
        for i in picks:
            size, tf, cam = reqs[i]
            r.resize(size, size)
            r.setCamera(*cam)
            r.setColorMap(tfs[tf][0])
            r.setRange(vr[0], vr[1])
            r.setOpacityScale(tfs[tf][1])
            r.resetAccum()
            st = r.state()
            fb = torch.zeros((size, size), dtype=torch.int32, device=device)
            r.render(fb)
            torch.cuda.synchronize()
            acc_s, acc_b = r.accum(), bt.accum(i)
            assert np.array_equal(acc_b, acc_s), f"request {i}: {np.count_nonzero(np.any(acc_b != acc_s, -1))} px"
            assert np.array_equal(bt.frame(i), fb.cpu().numpy().view(np.uint32))
            assert not np.all(acc_b == acc_b[:1, :1])
            # the batched frame against the oracle itself, every fourth row,
            # under this request's TF culling and majorants
            t_rgba = tfs[tf][0]
            xr = np.array(st.tfDomain, np.float32)
            om.set_culling(ora.map_max_opacity(om.gridlet_value_ranges(), t_rgba, xr),
                           om.element_max_opacity(t_rgba, xr))
            if ranges is None:
                ranges = om.grid_build(np.array(st.gridDims, np.int32), np.array(st.gridBounds, np.float32))
            maj = ora.map_max_opacity(ranges, t_rgba, xr)
            p = ora.render_params_from_state(st, maj, t_rgba)
            acc_o = np.zeros((size, size, 4), np.float32)
            fb_o = np.zeros((size, size), np.uint32)
            rows = list(range(i % 4, size, 4))
            with ThreadPoolExecutor(16) as ex:
                list(ex.map(lambda y: ora.lib().ora_render_f32(C.byref(p), om.ref, om.bins(), ora.ptr(acc_o),
                                                                ora.ptr(fb_o), y, y + 1), rows))
            npx = np.count_nonzero(np.any(acc_b[rows] != acc_o[rows], axis=-1))
            print(f"  request {i} ({size}^2, TF {tf}): {npx} of {len(rows) * size} oracle pixels differ", flush=True)
            assert npx == 0
            assert np.array_equal(bt.frame(i)[rows], fb_o[rows])
    finally:
        om.set_culling(*r_culling)


def test_bricks_workload_full_size(amr, workload, device):
    """The ExaBricks workload at its full size (the same leaf cells as 503K
    bricks of the brick maker, 23.6M regions): point location at 4096 random
    points against the brute-force blend over every brick (hits exactly where
    some brick's basis functions have weight, values bit-identical), and the
    1080p frame of the persistent kernel against the one-thread-per-pixel
    kernel, bit for bit."""
    import torch
    from owlexastitcher_amd import ExaBrickModel, ExaBrickSampler
    print("case: bricks workload full size, locate 4096 points + 1080p frame, both schemes", flush=True)
    r, _, _, rgba = workload
    bm = ExaBrickModel.from_amr(amr, 16)
    assert bm.counts()["abrs"] > 20_000_000
    bricks, scalars, _, _ = bm.arrays()
    s = ExaBrickSampler(bm)
    cb, _ = bm.bounds()
    p = np.random.default_rng(3).uniform(cb[:3] - 4, cb[3:] + 4, (4096, 3)).astype(np.float32)
    d_pos = torch.from_numpy(p).to(device)
    d_out = torch.full((len(p), 3), 7, dtype=torch.int32, device=device)
    s.locate(d_pos, d_out)
    torch.cuda.synchronize()
    out = d_out.cpu().numpy()
    v, w = ora.brick_blend_brute(bricks, scalars, p)
    hit = out[:, 1] == -1
    assert np.array_equal(hit, w > 0) and hit.sum() > 1000
    assert np.array_equal(out[hit, 2].view(np.float32), v[hit])
    rb = OWLRenderer(bm, (64, 64, 64))
    st = r.state()
    rb.resize(W, H)
    rb.setCamera(st.camOrg, st.camDir00, st.camDu, st.camDv)
    rb.setColorMap(rgba)
    vr = rb.valueRange
    rb.setRange(vr[0], vr[1])
    rb.setRelDomain(0.0, 100.0)
    rb.setOpacityScale(100.0)
    rb.setType(0)
    rb.setMaxBounces(4)
    rb.setDirectionalLight(0, (0.3, 1.0, 0.5), 2.0, True)
    fb = torch.zeros((H, W), dtype=torch.int32, device=device)
    frames = []
    for scheme in (RENDER_PERSISTENT, RENDER_MEGAKERNEL):
        rb.setScheme(scheme)
        rb.accumID = 0
        rb.render(fb)
        torch.cuda.synchronize()
        frames.append((rb.accum(), fb.cpu().numpy().copy()))
    (a0, f0), (a1, f1) = frames
    assert np.count_nonzero(np.any(a0 != a1, axis=-1)) == 0
    assert np.array_equal(f0, f1)
    assert np.count_nonzero(np.any(a0 != a0[:1, :1], axis=-1)) > W * H // 10



def test_learned_sampler_full_size(workload, device):
    """The learned-sampler configuration at its full size on the timed scene:
    hash grid of 16 levels x 2^19 entries x 2 features (bf16, float32 master
    weights), MLP 32-64-64-1, a pool of 2^24 samples drawn by the exact
    sampler (70% in stitching elements), 20 000 Adam steps of 2^18 samples
    (LR 1e-2 decaying to 1e-4); per-macrocell validation at 1e-3 of the
    value range; then 1080p frames with the network answering the element
    stage of the validated macrocells against frames of the exact sampler
    with the same seeds.  Reports what the network gains (frame time, image
    error); asserts the training, the validation bound, the widened majorants
    and the image error."""
    import torch
    from owlexastitcher_amd import NeuralField
    r, om, maj, rgba = workload
    print("case: learned sampler, full size (16 x 2^19 x 2 grid, 2^24 pool, 20000 x 2^18 steps)", flush=True)
    # the timed frame's settings (earlier tests of this file change the view)
    b = r.modelBounds
    center = 0.5 * (b[:3] + b[3:])
    r.resize(W, H)
    r.setCamera(*camera_look_at(center + np.array([-0.3, 0.7, 1.0]) * (b[3:] - b[:3]), center, (0, 1, 0), 70.0, W,
                                H))
    r.setColorMap(rgba)
    vr = r.valueRange
    r.setRange(vr[0], vr[1])
    r.setOpacityScale(100.0)
    r.setType(0)
    r.setSpp(1)
    r.setScheme(RENDER_PERSISTENT)
    r.setLaunchTuning(0, 1, 0, 256)
    f = NeuralField(r, levels=16, log2T=19, n_min=16, n_max=512, seed=1)
    assert f.makePool(1 << 24, 0.7, seed=7) > 0.9 * (1 << 24)
    loss = f.train(20000, batch=1 << 18, lr=1e-2, lr_end=1e-4)
    print(f"  loss {loss[:5].mean():.3g} -> {loss[-50:].mean():.3g}", flush=True)
    assert np.all(np.isfinite(loss)) and loss[-50:].mean() < 1e-3 * loss[:5].mean()
    v = f.validate(points_per_mc=64, bound=1e-3 * float(vr[1] - vr[0]))
    print(f"  validation {v}", flush=True)
    assert v["neural_mcs"] > 0 and v["max_error"] <= v["bound"]
    K = 8
    fb = torch.zeros((H, W), dtype=torch.int32, device=device)
    stream = torch.cuda.current_stream()

    def frames(k):
        r.accumID = 0
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(k):
            r.render(fb, stream)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / k, r.accum().astype(np.float64)

    r.setNeural(None)
    ms_exact, a1 = frames(K)
    for _ in range(K):
        r.render(fb, stream)
    torch.cuda.synchronize()
    a2 = r.accum().astype(np.float64)
    r.setNeural(f)
    ms_net, n1 = frames(K)
    r.accumID = 0
    c = r.render_counted(fb, stream)
    r.setNeural(None)
    A, B, N = a1 / K, (a2 - a1) / K, n1 / K
    noise = float(np.sqrt(np.mean((B - A)[..., :3] ** 2)))
    err = float(np.sqrt(np.mean((N - A)[..., :3] ** 2)))
    print(f"  exact {ms_exact:.2f} ms/frame, with the network {ms_net:.2f} ms/frame "
          f"(x{ms_exact / ms_net:.3f}); image RMSE vs exact {err:.3g} (exact seed noise {noise:.3g}); "
          f"network lookups {c['neural']}, exact element lookups {c['pyr'] + c['wedge'] + c['hex']}", flush=True)
    assert c["neural"] > 0
    assert np.all(np.isfinite(N)) and err <= noise
