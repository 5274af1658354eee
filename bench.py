"""Benchmark: path-traced samples/s at 1920x1080 on one MI355X.

Workload (the medium GPU configuration): 3-level AMR, refinement ratio 2,
base 128^3 cells of width 4 over [0,512]^3, refined where 6-octave fractal
value noise exceeds 0.55 and 0.7 (about 31M leaf cells), scalar = noise x
radial falloff; dual mesh with gridlets of <= 8^3 cells; 64^3 majorant grid;
1920x1080, 1 spp per frame, Woodcock-tracked path tracing capped at 4
bounces with one directional light.  One step = one frame.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "path-traced samples/sec @1080p on 1 GPU; neural vs exact sampler at fixed RMSE"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0  # dense bf16 MFMA

# algorithmic bytes per unit of work (leaf records only, traversal nodes are
# overhead): gridlet hit = 32 B record + 8 corners x 4 B; element hit = 32 B
# index row + 16 B per vertex; a cached gridlet hit re-reads the same 64 B;
# 4 B per macrocell visited; 36 B per pixel (accum read + write + fb write)
BYTES = {"gridlet": 64, "tet": 96, "pyr": 112, "wedge": 128, "hex": 160, "cached": 64, "dda_cells": 4,
         "pixels": 36, "neural": 512}

TF5 = np.array([
    [0.00, 0.0, 0.0, 0.3, 0.00],
    [0.20, 0.1, 0.4, 0.9, 0.05],
    [0.45, 0.2, 0.9, 0.5, 0.30],
    [0.70, 0.9, 0.8, 0.1, 0.60],
    [1.00, 1.0, 0.2, 0.1, 1.00],
], np.float32)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_scene(args):
    from owlexastitcher_amd import AmrScene, DualMesh, ExaStitchModel
    t = time.time()
    amr = AmrScene.noise(args.base_res, 2, 0.55, 0.7, 6, 2.0, args.seed)
    log(f"[bench] AMR: {len(amr)} leaf cells ({time.time() - t:.1f}s)")
    t = time.time()
    dm = DualMesh(amr)
    counts = dm.counts()
    log(f"[bench] dual mesh: {counts} ({time.time() - t:.1f}s)")
    model = ExaStitchModel.from_dual_mesh(dm)
    return amr, dm, model, counts


def setup_renderer(model, args):
    from owlexastitcher_amd import OWLRenderer, camera_look_at, tf_rasterize
    t = time.time()
    r = OWLRenderer(model, (args.mcs, args.mcs, args.mcs))
    log(f"[bench] renderer + sampler build: {time.time() - t:.1f}s, {r.sampler.info()}")
    b = r.modelBounds
    center = 0.5 * (b[:3] + b[3:])
    span = b[3:] - b[:3]
    org, d00, du, dv = camera_look_at(center + np.array([-0.3, 0.7, 1.0]) * span, center, (0, 1, 0), 70.0,
                                      args.width, args.height)
    r.resize(args.width, args.height)
    r.setCamera(org, d00, du, dv)
    rgba = tf_rasterize(TF5, 128)
    r.setColorMap(rgba)
    vr = r.valueRange
    r.setRange(vr[0], vr[1])
    r.setRelDomain(0.0, 100.0)
    r.setOpacityScale(100.0)
    r.setType(args.integrator)
    r.setMaxBounces(args.bounces)
    r.setSpp(args.spp)
    r.setDirectionalLight(0, (0.3, 1.0, 0.5), 2.0, True)
    from owlexastitcher_amd import RENDER_MEGAKERNEL, RENDER_PERSISTENT
    r.setScheme(RENDER_PERSISTENT if args.scheme == "persistent" else RENDER_MEGAKERNEL)
    r.setPersistentTuning(args.ctrl_steps, args.elem_batch, args.rare_batch)
    r.setLaunchTuning(args.element_stage, args.queue_rounds, args.queue_depth, args.block_size)
    if args.accel == "bvh":
        from owlexastitcher_amd import ACCEL_BVH
        t = time.time()
        r.setAccel(ACCEL_BVH)
        log(f"[bench] BVH accelerator built ({time.time() - t:.1f}s)")
    return r, rgba


def algorithmic_bytes(c):
    return sum(BYTES[k] * c[k] for k in BYTES)


def host_threads():
    """host threads for the CPU baseline: the CPUs this process may run on,
    at most 16 (the host share of one GPU)"""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(r, rgba, dm, args, target_s, dev):
    """The oracle (single-precision restatement of the reference's sampling
    and integration code) on this host's cores: whole frames of the same
    workload, consecutive accumIDs, until about target_s seconds of work."""
    import ctypes as C
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as ora
    st = r.state()
    idx, verts, gl, gs = dm.arrays()
    gmo, emo = r.sampler.prim_max_opacities(len(gl), len(idx))
    om = ora.Model(verts, idx, gl, gs, gmo, emo)
    t = time.time()
    bins = om.bins(64)
    log(f"[bench] cpu baseline: oracle bins built ({time.time() - t:.1f}s)")
    maj = r.sampler.max_opacities()
    threads = host_threads()
    accum = np.zeros((st.H, st.W, 4), np.float32)
    fb = np.zeros((st.H, st.W), np.uint32)
    frames, elapsed = 0, 0.0
    with ThreadPoolExecutor(threads) as ex:
        while elapsed < target_s and frames < 1000:
            st.accumID = frames
            p = ora.render_params_from_state(st, maj, rgba)
            t = time.time()
            list(ex.map(lambda y: ora.lib().ora_render_f32(C.byref(p), om.ref, bins, ora.ptr(accum), ora.ptr(fb),
                                                           y, y + 1), range(st.H)))
            elapsed += time.time() - t
            frames += 1
    samples = frames * st.H * st.W * st.spp
    out = {"value": samples / elapsed, "unit": "samples/s", "cores": threads, "kind": "port",
           "sample": f"{frames} whole {st.W}x{st.H} frames ({st.spp} spp, accumIDs 0..{frames - 1}) of the same "
                     f"workload, oracle f32 restatement, {threads} host threads (one image row per task), "
                     f"{elapsed:.1f}s", "cpu_model": cpu_model_name()}
    try:
        out["point_location"] = cpu_locate_baseline(r, om, maj, rgba, args, dev, threads)
    except Exception as e:  # reported, not required
        log(f"[bench] cpu locate baseline failed: {e!r}")
    return out


def cpu_model_name():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _time_locate(om, pos, threads):
    """wall seconds of the oracle's point location (its binned search over
    gridlets and elements, values interpolated) over pos on `threads` host
    threads, and the samples it returned"""
    from concurrent.futures import ThreadPoolExecutor
    chunk = max(1, -(-len(pos) // (threads * 8)))
    parts = [pos[i:i + chunk] for i in range(0, len(pos), chunk)]
    om.bins()
    t = time.perf_counter()
    if threads == 1:
        out = [om.locate_binned(p) for p in parts]
    else:
        with ThreadPoolExecutor(threads) as ex:
            out = list(ex.map(om.locate_binned, parts))
    return time.perf_counter() - t, np.concatenate(out)


def cpu_locate_baseline(r, om, maj, rgba, args, dev, threads):
    """Point location on the host (the CPU sampling path: the oracle's
    restatement of Gridlet.h / Plane.h / UElems.h behind uniform bins) on
    the query sets the survey plans: (i) 1M uniform points in [0,128]^3 on
    the 2-level Gaussian scene, (ii) the queries of the first 1080p frame of
    the timed workload (recorded by the oracle, whose frame equals the GPU's
    bit for bit), replayed through point location alone; all threads and one
    thread (one thread on a bounded prefix).  The GPU locates the same
    replayed positions, and its samples must equal the host's."""
    import torch
    from oracle import oracle as ora
    from owlexastitcher_amd import AmrScene, DualMesh
    res = {"cpu_model": cpu_model_name(), "threads": threads, "host_cpus": os.cpu_count()}
    # (i) config-1 scene, 1M uniform points
    amr1 = AmrScene.gaussians(64, 1, 32, 96, 8, 1)
    idx, verts, gl, gs = DualMesh(amr1).arrays()
    om1 = ora.Model(verts, idx, gl, gs)
    pos1 = np.random.default_rng(7).uniform(0.0, 128.0, (1 << 20, 3)).astype(np.float32)
    t_all, _ = _time_locate(om1, pos1, threads)
    t_one, _ = _time_locate(om1, pos1, 1)
    res["uniform_1M_gaussian_scene"] = {
        "queries": len(pos1), "queries_per_s_all_threads": len(pos1) / t_all,
        "queries_per_s_1_thread": len(pos1) / t_one, "ns_per_query_1_thread": t_one / len(pos1) * 1e9}
    log(f"[bench] cpu locate, 1M uniform points (2-level Gaussian scene): "
        f"{len(pos1) / t_all / 1e6:.1f}M/s on {threads} threads, {len(pos1) / t_one / 1e6:.2f}M/s on 1")
    # (ii) the first frame's queries
    st = r.state()
    st.accumID = 0
    p = ora.render_params_from_state(st, maj, rgba)
    t = time.perf_counter()
    q = ora.render_queries_f32(p, om, 1 << 26, threads=threads)
    log(f"[bench] cpu locate: first frame's {len(q)} queries recorded ({time.perf_counter() - t:.1f}s)")
    t_all, ref = _time_locate(om, q, threads)
    n1 = min(len(q), 1 << 22)
    t_one, _ = _time_locate(om, q[:n1], 1)
    d_pos = torch.from_numpy(q).to(dev)
    d_out = torch.empty((len(q), 3), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    r.sampler.locate(d_pos, d_out, stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(5):
        r.sampler.locate(d_pos, d_out, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    gpu_s = e0.elapsed_time(e1) / 5 * 1e-3
    same = bool(np.array_equal(d_out.cpu().numpy(), ref))
    res["first_frame_replay"] = {
        "queries": len(q), "queries_per_s_all_threads": len(q) / t_all, "queries_per_s_1_thread": n1 / t_one,
        "ns_per_query_1_thread": t_one / n1 * 1e9, "one_thread_sample": f"first {n1} queries",
        "gpu_queries_per_s_same_positions": len(q) / gpu_s, "gpu_samples_equal_cpu": same}
    log(f"[bench] cpu locate, first-frame replay: {len(q) / t_all / 1e6:.1f}M/s on {threads} threads, "
        f"{n1 / t_one / 1e6:.2f}M/s on 1; GPU {len(q) / gpu_s / 1e9:.2f}G/s on the same positions; "
        f"samples equal: {same}")
    return res


def copy_calibration(nbytes, dev):
    """One plain device copy of nbytes (torch's copy kernel): the reference
    point that scales rocprofv3's FETCH_SIZE / WRITE_SIZE to bytes."""
    import torch
    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    log(f"[bench] copy calibration: {nbytes} bytes read + {nbytes} bytes written")
    del a, b


def read_traffic(args):
    """HBM bytes per frame of the render kernel from the committed rocprofv3
    summary for this workload (FETCH_SIZE + WRITE_SIZE, copy-calibrated)."""
    path = os.path.join(ROOT, "profiles", f"kRender{'Persistent' if args.scheme == 'persistent' else ''}"
                                          f"_noise3_{args.width}x{args.height}_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_frame"):
            return float(line.split("=")[1])
    return None


def random_tf(rng, n_points):
    """Random piecewise-linear RGBA transfer function with n_points control
    points (x ascending from 0 to 1), rasterised to 128 texels."""
    from owlexastitcher_amd import tf_rasterize
    xs = np.sort(rng.uniform(0.0, 1.0, n_points - 2))
    pts = np.zeros((n_points, 5), np.float32)
    pts[:, 0] = np.concatenate([[0.0], xs, [1.0]])
    pts[:, 1:4] = rng.uniform(0.0, 1.0, (n_points, 3))
    pts[:, 4] = rng.uniform(0.0, 1.0, n_points)
    return tf_rasterize(pts, 128)


def batch_workload(r, args, dev):
    """Batched requests: 64 views, sizes drawn from {256,512,768,1024}^2,
    seeded orbit cameras, 4 random 5-8 point TFs with opacity scale
    U[0.5,4], 16 spp each, in one persistent launch per frame."""
    import torch
    from owlexastitcher_amd import RENDER_PERSISTENT, RenderBatch, camera_look_at
    rng = np.random.default_rng(args.seed)
    vr = r.valueRange
    # the single 1080p view's work per path sample and frame time, for the
    # per-unit comparison below
    stream = torch.cuda.current_stream()
    fb1 = torch.zeros((args.height, args.width), dtype=torch.int32, device=dev)
    for _ in range(3):
        r.render(fb1, stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(10):
        r.render(fb1, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    single_frame_ms = e0.elapsed_time(e1) / 10
    single_counts = r.render_counted()
    del fb1
    tfs = [(random_tf(rng, int(rng.integers(5, 9))), float(rng.uniform(0.5, 4.0))) for _ in range(4)]
    b = r.modelBounds
    c = 0.5 * (b[:3] + b[3:])
    rad = 1.2 * float(np.linalg.norm(b[3:] - b[:3]))
    reqs = []
    for _ in range(args.requests):
        size = int(rng.choice([256, 512, 768, 1024]))
        az, el = float(rng.uniform(0, 2 * np.pi)), float(rng.uniform(-0.8, 0.8))
        eye = c + rad * np.array([np.cos(el) * np.cos(az), np.sin(el), np.cos(el) * np.sin(az)])
        reqs.append((size, int(rng.integers(0, 4)), camera_look_at(eye, c, (0, 1, 0), 45.0, size, size)))

    def make(spp):
        bt = RenderBatch(r)
        ti = [bt.addTransferFunction(rgba, vr[0], vr[1], op) for rgba, op in tfs]
        for size, tf, cam in reqs:
            bt.addRequest(size, size, spp, ti[tf], cam, args.integrator)
        return bt

    t = time.time()
    tune = make(1)
    # seven axes: ctrlSteps, elemBatch, rareBatch, element stage (in the wave /
    # workgroup queue in LDS / device queue rounds), queue rounds, queue depth,
    # workgroup size
    best, best_ms = tune.autotuneLaunch((8, 16), (16, 32), (0, 8), (0, 1, 2), (2,), (0, 1 << 18), (128, 256))
    log(f"[bench] autotune on the 1-spp batch: best {best} ({best_ms:.1f} ms/frame), {time.time() - t:.1f}s")
    del tune
    # the batch's work per path sample (one counted 1-spp frame, same tuning)
    bc = make(1)
    bc.setTuning(*best[:3])
    bc.setLaunchTuning(*best[3:])
    batch_counts = bc.render_counted()
    del bc
    bt = make(args.batch_spp)
    bt.setTuning(*best[:3])
    bt.setLaunchTuning(*best[3:])
    for _ in range(args.warmup):
        bt.render(stream)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for i in range(args.steps):
        bt.render(stream)
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    frame_ms = [ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps)]
    pixels = sum(q[0] * q[0] for q in reqs)
    samples = pixels * args.batch_spp
    # the same requests one at a time through the single-view renderer
    r.setScheme(RENDER_PERSISTENT)
    r.setSpp(args.batch_spp)
    r.setType(args.integrator)
    single_ms = 0.0
    fb = None
    for size, tf, cam in reqs:
        r.resize(size, size)
        r.setCamera(*cam)
        r.setColorMap(tfs[tf][0])
        r.setRange(vr[0], vr[1])
        fb = torch.zeros((size, size), dtype=torch.int32, device=dev)
        r.render(fb, stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        r.render(fb, stream)
        e1.record(stream)
        torch.cuda.synchronize()
        single_ms += e0.elapsed_time(e1)
    log(f"[bench] batch frame {np.mean(frame_ms):.1f} ms vs {single_ms:.1f} ms for the requests one at a time")

    def per_sample(cnt, ms, spp):
        n = cnt["pixels"] * spp
        q = {k: cnt[k] / cnt["pixels"] for k in ("dda_cells", "tf", "gridlet", "cached", "pyr", "wedge", "hex")}
        q["ns_per_path_sample"] = ms * 1e6 / n
        q["ns_per_tf_fetch"] = ms * 1e6 / (cnt["tf"] * spp)
        q["ns_per_dda_step"] = ms * 1e6 / (cnt["dda_cells"] * spp)
        return q
    work = {"batch": per_sample(batch_counts, float(np.mean(frame_ms)), args.batch_spp),
            "single_1080p": per_sample(single_counts, single_frame_ms, 1)}
    log(f"[bench] work per path sample: batch {work['batch']}")
    log(f"[bench] work per path sample: 1080p view {work['single_1080p']}")
    return {
        "metric": METRIC, "value": samples * args.steps / wall, "unit": "samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"{args.requests} batched requests, sizes U{{256,512,768,1024}}^2, seeded orbit "
                               f"cameras, 4 random 5-8 point TFs (opacity U[0.5,4]), {args.batch_spp} spp, one "
                               f"persistent launch per frame, same 3-level scene",
                   "pixels": pixels, "autotuned": list(best), "frame_ms": float(np.mean(frame_ms)),
                   "single_view_sum_ms": single_ms, "work_per_path_sample": work},
        "roofline": None, "cpu_baseline": None,
    }


def neural_workload(r, args, dev):
    """Learned sampler: hash grid (16 levels x 2^19 x 2) + MLP 32-64-64-1,
    trained with Adam on samples of the exact sampler (70% in stitching
    elements, 30% in gridlets), validated per macrocell against a bound of
    1e-3 of the value range; then the same frames rendered with the exact
    sampler and with the network answering the element stage."""
    import torch
    from owlexastitcher_amd import NeuralField
    vr = r.valueRange
    f = NeuralField(r, levels=16, log2T=19, n_min=16, n_max=512, seed=args.seed)
    t = time.time()
    nvalid = f.makePool(args.pool, 0.7, seed=args.seed + 6)
    log(f"[bench] pool: {nvalid} of {args.pool} samples found by the exact sampler ({time.time() - t:.1f}s)")
    torch.cuda.synchronize()
    t = time.time()
    loss = f.train(args.train_steps, batch=args.train_batch, lr=1e-2, lr_end=args.lr_end)
    train_s = time.time() - t
    log(f"[bench] trained {args.train_steps} steps x {args.train_batch} in {train_s:.1f}s; "
        f"loss {loss[0]:.3g} -> {np.mean(loss[-50:]):.3g}")
    t = time.time()
    v = f.validate(points_per_mc=args.val_points, bound=args.val_bound * float(vr[1] - vr[0]))
    log(f"[bench] validation ({time.time() - t:.1f}s): {v}")
    stream = torch.cuda.current_stream()
    fb = torch.zeros((args.height, args.width), dtype=torch.int32, device=dev)

    def frames(k):
        r.accumID = 0
        for _ in range(args.warmup):
            r.render(fb, stream)
        r.accumID = 0
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(k):
            r.render(fb, stream)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / k, r.accum().astype(np.float64)

    K = args.steps
    r.setNeural(None)
    ms_exact, s_exact = frames(K)
    for _ in range(K):
        r.render(fb, stream)
    torch.cuda.synchronize()
    s_exact2 = r.accum().astype(np.float64)
    c_exact = None
    r.setNeural(f)
    ms_net, s_net = frames(K)
    r.accumID = 0
    c_net = r.render_counted(fb, stream)
    r.setNeural(None)
    r.accumID = 0
    c_exact = r.render_counted(fb, stream)
    A, B, N = s_exact / K, (s_exact2 - s_exact) / K, s_net / K
    rmse = float(np.sqrt(np.mean((N - A)[..., :3] ** 2)))
    noise = float(np.sqrt(np.mean((B - A)[..., :3] ** 2)))
    samples = args.width * args.height * args.spp
    # MFMA rate of the network kernels: training (forward + backward of a
    # batch, 3 x 12 416 FLOP per sample, the count rocprofv3's
    # SQ_INSTS_VALU_MFMA_MOPS_BF16 x 512 reproduces) and inference over 2^22
    # points (12 416 FLOP each), against the dense bf16 peak
    flop_sample = 2 * (32 * 64 + 64 * 64 + 64 * 1)
    train_tflops = 3 * flop_sample * args.train_batch / (train_s / args.train_steps) / 1e12
    b = r.modelBounds
    n_eval = 1 << 22
    d_pos = torch.from_numpy(np.random.default_rng(args.seed).uniform(b[:3], b[3:], (n_eval, 3)).astype(np.float32)).to(dev)
    d_val = torch.empty(n_eval, dtype=torch.float32, device=dev)
    f.evaluate(d_pos, d_val, stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(5):
        f.evaluate(d_pos, d_val, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    eval_ms = e0.elapsed_time(e1) / 5
    eval_tflops = flop_sample * n_eval / (eval_ms * 1e-3) / 1e12
    log(f"[bench] network MFMA rate: training {train_tflops:.2f} TFLOP/s ({100 * train_tflops / BF16_PEAK_TFLOPS:.2f}% "
        f"of the bf16 peak), inference {eval_tflops:.1f} TFLOP/s ({100 * eval_tflops / BF16_PEAK_TFLOPS:.2f}%)")
    log(f"[bench] exact {ms_exact:.2f} ms/frame, neural {ms_net:.2f} ms/frame; image RMSE {rmse:.3g} "
        f"(seed noise {noise:.3g}); element lookups exact {sum(c_exact[k] for k in ('tet', 'pyr', 'wedge', 'hex'))}, "
        f"with network: {c_net['neural']} by the network + "
        f"{sum(c_net[k] for k in ('tet', 'pyr', 'wedge', 'hex'))} exact")
    return {
        "metric": METRIC, "value": samples / (ms_net * 1e-3), "unit": "samples/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": ms_net, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 (network) / fp32", "data": "synthetic",
        "config": {"workload": "1080p frame of the 3-level noise scene with the neural field answering the "
                               "element stage of validated macrocells (hash grid 16 x 2^19 x 2, MLP 32-64-64-1)",
                   "exact_samples_per_s": samples / (ms_exact * 1e-3), "exact_ms": ms_exact, "neural_ms": ms_net,
                   "speedup_vs_exact": ms_exact / ms_net, "image_rmse_vs_exact_same_seeds": rmse,
                   "exact_seed_noise_rmse": noise, "validation": v, "train_steps": args.train_steps,
                   "train_batch": args.train_batch, "train_seconds": train_s, "pool": args.pool,
                   "lr": [1e-2, args.lr_end],
                   "final_loss": float(np.mean(loss[-50:])), "counters_neural": c_net,
                   "counters_exact": c_exact, "train_step_ms": train_s / args.train_steps * 1e3,
                   "inference_ms_per_2e22_points": eval_ms, "inference_tflops": eval_tflops,
                   "inference_mfma_frac": eval_tflops / BF16_PEAK_TFLOPS},
        "roofline": {"bound": "mfma", "achieved": train_tflops, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": train_tflops / BF16_PEAK_TFLOPS, "traffic": None, "kernel": "kNeuralTrain"},
        "cpu_baseline": None,
    }


TF_SHOCK = np.array([
    [0.00, 0.1, 0.1, 0.3, 0.00],
    [0.12, 0.2, 0.3, 0.9, 0.02],
    [0.35, 0.1, 0.8, 0.9, 0.10],
    [0.60, 0.9, 0.6, 0.2, 0.40],
    [1.00, 1.0, 0.9, 0.8, 0.90],
], np.float32)


def large_workload(args, dev):
    """The large configuration: 5-level AMR, base 256^3 cells of width 16 over
    [0,4096]^3, refined along an expanding spherical shock shell (radius 1200,
    thickness 40) with turbulence in its wake (about 380M leaf cells, 16% of
    them in stitching elements); the
    neural field retrained on it; batched frames of one 3840x2160 view plus 16
    side views at 1024^2, 64 spp each, in one persistent launch, with the exact
    element stage and with the network.  --large-scale k divides the base
    resolution and the radius by k (a smaller copy of the same scene)."""
    import torch
    from owlexastitcher_amd import (AmrScene, DualMesh, ExaStitchModel, NeuralField, OWLRenderer, RenderBatch,
                                    camera_look_at, tf_rasterize)
    k = args.large_scale
    t = time.time()
    amr = AmrScene.shock(256 // k, 4, 1200.0 / k, 40.0, 0.74, 0.43, args.seed)
    ncells = len(amr)
    log(f"[bench] shock AMR: {ncells} leaf cells ({time.time() - t:.1f}s)")
    t = time.time()
    dm = DualMesh(amr)
    counts = dm.counts()
    del amr
    log(f"[bench] dual mesh: {counts} ({time.time() - t:.1f}s)")
    t = time.time()
    model = ExaStitchModel.from_dual_mesh(dm)
    del dm
    r = OWLRenderer(model, (args.mcs, args.mcs, args.mcs))
    del model
    free, total = torch.cuda.mem_get_info()
    log(f"[bench] renderer + sampler build: {time.time() - t:.1f}s, {r.sampler.info()}, "
        f"device memory in use {(total - free) / 2**30:.1f} GiB")
    b = r.modelBounds
    c = 0.5 * (b[:3] + b[3:])
    R = 1200.0 / k
    vr = r.valueRange
    rgba = tf_rasterize(TF_SHOCK, 128)
    os_ = 0.05
    r.setColorMap(rgba)
    r.setRange(vr[0], vr[1])
    r.setOpacityScale(100.0 + np.log(os_) / np.log(1.1))
    r.setMaxBounces(args.bounces)
    r.setDirectionalLight(0, (0.3, 1.0, 0.5), 2.0, True)
    W, H = 3840, 2160
    r.resize(W, H)
    main_cam = camera_look_at(c + np.array([-0.5, 0.45, 1.0]) * 1.9 * R, c, (0, 1, 0), 50.0, W, H)
    r.setCamera(*main_cam)
    rng = np.random.default_rng(args.seed)
    views = [(W, H, main_cam)]
    for _ in range(16):
        az, el = float(rng.uniform(0, 2 * np.pi)), float(rng.uniform(-0.7, 0.7))
        eye = c + 2.2 * R * np.array([np.cos(el) * np.cos(az), np.sin(el), np.cos(el) * np.sin(az)])
        views.append((1024, 1024, camera_look_at(eye, c, (0, 1, 0), 45.0, 1024, 1024)))
    pixels = sum(w * h for w, h, _ in views)
    st = r.state()

    def make(spp):
        bt = RenderBatch(r)
        ti = bt.addTransferFunction(rgba, st.tfDomain[0], st.tfDomain[1], st.opacityScale)
        for w, h, cam in views:
            bt.addRequest(w, h, spp, ti, cam, args.integrator)
        return bt

    t = time.time()
    tune = make(1)
    best, best_ms = tune.autotune((8, 16, 32), (8, 16, 32), (0, 8, 16))
    del tune
    log(f"[bench] autotune on the 1-spp batch: best {best} ({best_ms:.1f} ms/frame), {time.time() - t:.1f}s")
    stream = torch.cuda.current_stream()

    def timed_frame(bt):
        bt.resetAccum()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        bt.render(stream)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    bt = make(args.large_spp)
    bt.setTuning(*best)
    ms_exact = timed_frame(bt)
    img_exact = bt.accum(0).astype(np.float64)
    log(f"[bench] exact batch frame ({pixels} pixels x {args.large_spp} spp): {ms_exact:.0f} ms, "
        f"{pixels * args.large_spp / (ms_exact * 1e-3) / 1e6:.0f}M samples/s")
    small = make(1)
    small.setTuning(*best)
    c_exact = small.render_counted(stream)
    # neural field retrained on this scene
    f = NeuralField(r, levels=16, log2T=19, n_min=16, n_max=args.large_nmax, seed=args.seed)
    t = time.time()
    nvalid = f.makePool(args.pool, 0.7, seed=args.seed + 6)
    log(f"[bench] pool: {nvalid} of {args.pool} samples ({time.time() - t:.1f}s)")
    t = time.time()
    loss = f.train(args.train_steps, batch=args.train_batch, lr=1e-2, lr_end=args.lr_end)
    train_s = time.time() - t
    log(f"[bench] trained {args.train_steps} steps x {args.train_batch} in {train_s:.1f}s; "
        f"loss {loss[0]:.3g} -> {np.mean(loss[-50:]):.3g}")
    t = time.time()
    v = f.validate(points_per_mc=args.val_points, bound=args.val_bound * float(vr[1] - vr[0]))
    log(f"[bench] validation ({time.time() - t:.1f}s): {v}")
    r.setNeural(f)
    ms_net = timed_frame(bt)
    img_net = bt.accum(0).astype(np.float64)
    c_net = small.render_counted(stream)
    r.setNeural(None)
    rmse = float(np.sqrt(np.mean((img_net - img_exact)[..., :3] ** 2))) / args.large_spp
    free, total = torch.cuda.mem_get_info()
    log(f"[bench] neural batch frame: {ms_net:.0f} ms; image RMSE vs exact (same seeds) {rmse:.3g}; "
        f"network lookups {c_net['neural']}; device memory in use {(total - free) / 2**30:.1f} GiB")
    samples = pixels * args.large_spp
    return {
        "metric": METRIC, "value": samples / (ms_exact * 1e-3), "unit": "samples/s", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": ms_exact, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 (exact) / bf16 network", "data": "synthetic",
        "config": {"workload": f"5-level shock-shell AMR ({ncells} leaf cells, base {256 // k}^3 x width 16, "
                               f"radius {1200.0 / k:g}, thickness 40), dual mesh {counts['elements']} stitching "
                               f"elements + {counts['gridlets']} gridlets, {args.mcs}^3 majorant grid; batched "
                               f"3840x2160 + 16 x 1024^2 views, {args.large_spp} spp each, one persistent launch",
                   "stitching_share_of_cells": counts["elements"] / ncells, "pixels": pixels,
                   "autotuned": list(best), "exact_ms": ms_exact, "neural_ms": ms_net,
                   "neural_samples_per_s": samples / (ms_net * 1e-3), "image_rmse_neural_vs_exact": rmse,
                   "validation": v, "train_steps": args.train_steps, "train_seconds": train_s,
                   "counters_exact_1spp": c_exact, "counters_neural_1spp": c_net,
                   "device_gib_in_use": (total - free) / 2**30},
        "roofline": None, "cpu_baseline": None,
    }


def locate_workload(r, args, dev):
    """The sampler alone (ExaStitchSampler sample(), ExaStitchSampler.h:128-140):
    point location + interpolation of 2^24 points per launch, uniformly random
    over the model box (every query an independent gather) and the same points
    in Morton order (neighbouring lanes hit neighbouring records, as samples
    along coherent rays do).  Algorithmic bytes: the leaf record per hit
    (BYTES) + 12 B position in + 12 B sample out."""
    import torch
    n = args.locate_points
    b = r.modelBounds
    rng = np.random.default_rng(args.seed)
    pos = rng.uniform(b[:3], b[3:], (n, 3)).astype(np.float32)
    # 10-bit Morton code of the position in the box
    q = np.clip(((pos - b[:3]) / (b[3:] - b[:3]) * 1023.0).astype(np.uint64), 0, 1023)
    code = np.zeros(n, np.uint64)
    for bit in range(10):
        for k in range(3):
            code |= ((q[:, k] >> np.uint64(bit)) & np.uint64(1)) << np.uint64(3 * bit + k)
    orders = {"uniform": pos, "morton": pos[np.argsort(code, kind="stable")]}
    out = torch.empty((n, 3), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    res = {}
    for name, p in orders.items():
        d_pos = torch.from_numpy(np.ascontiguousarray(p)).to(dev)
        for _ in range(args.warmup):
            r.sampler.locate(d_pos, out, stream)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        torch.cuda.synchronize()
        ev[0].record(stream)
        for _ in range(args.steps):
            r.sampler.locate(d_pos, out, stream)
        ev[1].record(stream)
        torch.cuda.synchronize()
        ms = ev[0].elapsed_time(ev[1]) / args.steps
        c = r.sampler.locate_counted(d_pos, out, stream)
        alg = sum(BYTES[k] * c[k] for k in ("gridlet", "tet", "pyr", "wedge", "hex")) + 24 * n
        gbs = alg / (ms * 1e-3) / 1e9
        res[name] = {"ms_per_launch": ms, "queries_per_s": n / (ms * 1e-3), "algorithmic_bytes": alg,
                     "achieved_GBps": gbs, "frac": gbs / HBM_PEAK_GBS, "counters": c}
        log(f"[bench] locate {name}: {ms:.3f} ms per {n} queries = {n / (ms * 1e-3) / 1e9:.2f}G queries/s, "
            f"{gbs:.0f} GB/s algorithmic ({100 * gbs / HBM_PEAK_GBS:.1f}% of HBM peak); {c}")
    best = max(res.values(), key=lambda v: v["queries_per_s"])
    return {
        "metric": METRIC, "value": best["queries_per_s"], "unit": "point queries/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": best["ms_per_launch"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"point location + interpolation of {n} points per launch in the 3-level noise "
                               f"scene (tables accelerator)", "orders": res},
        "roofline": {"bound": "hbm", "achieved": best["achieved_GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": best["frac"], "traffic": None, "kernel": "kLocate"},
        "cpu_baseline": None,
    }


def bricks_workload(amr, r, args, dev):
    """The ExaBricks comparison sampler (sampler/ExaBrickSampler.h, ABR mode)
    on the same AMR leaf cells, view, transfer function, light and bounce cap
    as the 1080p frame: bricks from the brick maker (16^3-cell blocks),
    regions with per-region majorants walked on their kd-tree, every sample
    a blend over the region's bricks.  The exaStitch frame of the same run is
    timed beside it.  Algorithmic bytes: 64 B per brick blended (32 B record
    + 8 corner scalars), 36 B per region entered (box + majorant), 36 B per
    pixel."""
    import torch
    from owlexastitcher_amd import ExaBrickModel, OWLRenderer
    t = time.time()
    bm = ExaBrickModel.from_amr(amr, 16)
    bc = bm.counts()
    log(f"[bench] brick model: {bc} ({time.time() - t:.1f}s)")
    t = time.time()
    rb = OWLRenderer(bm, (args.mcs, args.mcs, args.mcs))
    log(f"[bench] brick renderer: {time.time() - t:.1f}s")
    st = r.state()
    rb.resize(args.width, args.height)
    rb.setCamera(st.camOrg, st.camDir00, st.camDu, st.camDv)
    from owlexastitcher_amd import tf_rasterize
    rb.setColorMap(tf_rasterize(TF5, 128))
    vr = rb.valueRange
    rb.setRange(vr[0], vr[1])
    rb.setRelDomain(0.0, 100.0)
    rb.setOpacityScale(100.0)
    rb.setType(args.integrator)
    rb.setMaxBounces(args.bounces)
    rb.setSpp(args.spp)
    rb.setDirectionalLight(0, (0.3, 1.0, 0.5), 2.0, True)
    stream = torch.cuda.current_stream()
    fb = torch.zeros((args.height, args.width), dtype=torch.int32, device=dev)

    def timed(renderer):
        for _ in range(args.warmup):
            renderer.render(fb, stream)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev[0].record(stream)
        for _ in range(args.steps):
            renderer.render(fb, stream)
        ev[1].record(stream)
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) / args.steps, (time.perf_counter() - t0) / args.steps * 1e3

    from owlexastitcher_amd import RENDER_MEGAKERNEL, RENDER_PERSISTENT
    rb.setScheme(RENDER_MEGAKERNEL)
    ms_mega, _ = timed(rb)
    rb.setScheme(RENDER_PERSISTENT)
    ms_b, wall_b = timed(rb)
    ms_s, wall_s = timed(r)
    cnt = rb.render_counted(fb, stream)
    alg = 64 * cnt["bricks"] + 36 * cnt["regions"] + 36 * cnt["pixels"]
    achieved = alg / (ms_b * 1e-3) / 1e9
    spf = args.width * args.height * args.spp
    log(f"[bench] bricks: {ms_b:.2f} ms/frame persistent ({spf / (ms_b * 1e-3) / 1e6:.0f}M samples/s), "
        f"{ms_mega:.2f} ms/frame one thread per pixel, exaStitch {ms_s:.2f} ms/frame; counters {cnt}; "
        f"{achieved:.0f} GB/s algorithmic")
    return {
        "metric": METRIC, "value": spf / (wall_b * 1e-3), "unit": "samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall_b, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"ExaBricks comparison sampler on the same {len(amr)} leaf cells "
                               f"({bc['bricks']} bricks, {bc['abrs']} regions), {args.width}x{args.height}, "
                               f"{args.spp} spp/frame, path tracer <= {args.bounces} bounces, 1 directional light",
                   "brick_model": bc, "frame_ms": ms_b, "megakernel_frame_ms": ms_mega, "exastitch_frame_ms": ms_s,
                   "exastitch_samples_per_s": spf / (wall_s * 1e-3),
                   "volume_samples_per_s": cnt["samples"] / (ms_b * 1e-3), "counters": cnt},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "kRenderBricksPersistent"},
        "cpu_baseline": None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=1)
    ap.add_argument("--bounces", type=int, default=4)
    ap.add_argument("--integrator", type=int, default=0, help="0 path tracer, 1 single scatter")
    ap.add_argument("--base-res", type=int, default=128)
    ap.add_argument("--mcs", type=int, default=64)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--accel", choices=["tables", "bvh"], default="tables")
    ap.add_argument("--scheme", choices=["persistent", "megakernel"], default="persistent")
    ap.add_argument("--ctrl-steps", type=int, default=4)
    ap.add_argument("--elem-batch", type=int, default=16)
    ap.add_argument("--rare-batch", type=int, default=24)
    ap.add_argument("--element-stage", type=int, default=0, choices=[0, 1, 2],
                    help="0: in the waiting lanes' wave; 1: workgroup queue in LDS; 2: device queue between launches")
    ap.add_argument("--queue-rounds", type=int, default=1, help="element stage 2: queue rounds per frame")
    ap.add_argument("--queue-depth", type=int, default=0, help="entries per element-stage queue (0: one per pixel)")
    ap.add_argument("--block-size", type=int, default=256, choices=[64, 128, 256])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--locate-points", type=int, default=1 << 24)
    ap.add_argument("--workload", choices=["frame", "batch", "neural", "large", "locate", "bricks"], default="frame",
                    help="frame: the 1080p frame (default); batch: batched requests in one launch; "
                         "neural: exact vs neural-field element stage; large: the 400M-cell shock scene, "
                         "batched 4K + 16 side views with and without the network; locate: the sampler alone; "
                         "bricks: the ExaBricks comparison sampler beside the exaStitch frame")
    ap.add_argument("--large-scale", type=int, default=1, help="large workload: base resolution and radius / k")
    ap.add_argument("--large-spp", type=int, default=64)
    ap.add_argument("--large-nmax", type=int, default=2048, help="large workload: finest hash-grid resolution")
    ap.add_argument("--pool", type=int, default=1 << 24)
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--train-batch", type=int, default=1 << 18)
    ap.add_argument("--val-points", type=int, default=64)
    ap.add_argument("--lr-end", type=float, default=1e-4, help="learning rate at the last step (exponential decay)")
    ap.add_argument("--val-bound", type=float, default=1e-3, help="validation bound, fraction of the value range")
    ap.add_argument("--requests", type=int, default=64)
    ap.add_argument("--batch-spp", type=int, default=16)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="run one plain device copy of BYTES first (for rocprofv3 counter scaling)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (the renderer shards in image space "
              "across GPUs by design, but no multi-GPU path is built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from owlexastitcher_amd import lib
    lib()
    dev = torch.device("cuda:0")
    if args.calibrate_copy:
        copy_calibration(args.calibrate_copy, dev)
    if args.workload == "large":
        print(json.dumps(large_workload(args, dev)), flush=True)
        return
    amr, dm, model, counts = build_scene(args)
    r, rgba = setup_renderer(model, args)
    if args.workload == "batch":
        print(json.dumps(batch_workload(r, args, dev)), flush=True)
        return
    if args.workload == "neural":
        print(json.dumps(neural_workload(r, args, dev)), flush=True)
        return
    if args.workload == "locate":
        print(json.dumps(locate_workload(r, args, dev)), flush=True)
        return
    if args.workload == "bricks":
        print(json.dumps(bricks_workload(amr, r, args, dev)), flush=True)
        return
    stream = torch.cuda.current_stream()
    fb = torch.zeros((args.height, args.width), dtype=torch.int32, device=dev)

    log(f"[bench] warmup {args.warmup} frames")
    for _ in range(args.warmup):
        r.render(fb, stream)
    torch.cuda.synchronize()
    first_timed = r.accumID
    log(f"[bench] timing {args.steps} frames")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev[0].record(stream)
    for i in range(args.steps):
        r.render(fb, stream)
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    frame_ms = [ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps)]
    kernel_ms = float(np.mean(frame_ms))
    log(f"[bench] frames: mean {kernel_ms:.2f} ms, min {min(frame_ms):.2f}, max {max(frame_ms):.2f}; wall {wall:.2f}s")

    # instrumented frame with the seeds of the first timed frame
    r.accumID = first_timed
    cnt = r.render_counted(fb, stream)
    alg = algorithmic_bytes(cnt)
    log(f"[bench] counters (frame accumID={first_timed}): {cnt}; algorithmic bytes {alg / 1e6:.1f} MB")
    achieved = alg / (kernel_ms * 1e-3) / 1e9

    samples_per_frame = args.width * args.height * args.spp
    value = samples_per_frame * args.steps / wall
    queries = sum(cnt[k] for k in ("gridlet", "tet", "pyr", "wedge", "hex", "miss", "cached"))

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(r, rgba, dm, args, args.cpu_seconds, dev)
            log(f"[bench] cpu baseline: {cpu}")
        except Exception as e:  # the baseline is reported, not required
            log(f"[bench] cpu baseline failed: {e!r}")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": (f"3-level fractal-noise AMR ({len(amr)} leaf cells, base {args.base_res}^3 x width 4), "
                         f"dual mesh {counts['elements']} stitching elements + {counts['gridlets']} gridlets, "
                         f"{args.mcs}^3 majorant grid, {args.width}x{args.height}, {args.spp} spp/frame, "
                         + ("path tracer <= %d bounces" % args.bounces if args.integrator == 0
                            else "single scatter") + ", 1 directional light"),
            "accel": args.accel,
            "scheme": args.scheme,
            "persistent_tuning": [args.ctrl_steps, args.elem_batch, args.rare_batch],
            "launch_tuning": {"element_stage": args.element_stage, "queue_rounds": args.queue_rounds,
                              "queue_depth": args.queue_depth, "block_size": args.block_size},
            "frame_ms": kernel_ms,
            "volume_samples_per_s": queries / (kernel_ms * 1e-3),
            "counters": cnt,
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(args),
                     "kernel": "kRenderPersistent" if args.scheme == "persistent" else "kRender"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
