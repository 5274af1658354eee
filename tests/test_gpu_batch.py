"""Batched requests in one persistent launch: every request's image equals
the oracle's image for its settings and the single-view renderer's, bit for
bit, under any scheduling knobs, eagerly and through a replayed HIP graph."""
import numpy as np
import pytest

from oracle import oracle as ora
from owlexastitcher_amd import OWLRenderer, RenderBatch, camera_look_at
from tests import scenes
from tests.test_gpu_render import oracle_model

pytestmark = pytest.mark.gpu

TF_B = np.array([
    [0.00, 0.9, 0.9, 0.9, 0.00],
    [0.35, 0.2, 0.5, 1.0, 0.40],
    [0.60, 1.0, 0.4, 0.2, 0.10],
    [1.00, 0.3, 1.0, 0.3, 0.90],
], np.float32)

# (width, height, spp, integrator, tf, orbit azimuth, elevation, opacity slider)
REQUESTS = [
    (37, 23, 1, 0, 0, 0.3, 0.5, 115.0),
    (64, 48, 2, 0, 1, 2.1, 0.2, 110.0),
    (50, 61, 1, 1, 0, 4.0, -0.3, 120.0),
    (16, 16, 3, 0, 1, 5.5, 0.8, 105.0),
    (80, 33, 1, 0, 0, 1.2, -0.6, 125.0),
    (9, 70, 2, 1, 1, 3.3, 0.1, 100.0),
]


@pytest.fixture(scope="module")
def cfg1():
    return scenes.config1()


def orbit_camera(r, W, H, az, el):
    b = r.modelBounds
    c = 0.5 * (b[:3] + b[3:])
    rad = 1.3 * float(np.linalg.norm(b[3:] - b[:3]))
    eye = c + rad * np.array([np.cos(el) * np.cos(az), np.sin(el), np.cos(el) * np.sin(az)])
    return camera_look_at(eye, c, (0, 1, 0), 50.0, W, H)


def configure_single(r, q, tfs):
    """The single-view renderer set to request q; returns (tfDomain, opacity)."""
    W, H, spp, integ, tf, az, el, slider = q
    r.resize(W, H)
    r.setCamera(*orbit_camera(r, W, H, az, el))
    r.setColorMap(tfs[tf])
    vr = r.valueRange
    r.setRange(vr[0], vr[1])
    r.setRelDomain(0.0, 100.0)
    r.setOpacityScale(slider)
    r.setSpp(spp)
    r.setType(integ)
    st = r.state()
    return (st.tfDomain[0], st.tfDomain[1]), st.opacityScale


@pytest.fixture(scope="module")
def batch_setup(cfg1, device):
    r = OWLRenderer(cfg1.model, (32, 32, 32))
    r.setMaxBounces(4)
    r.setDirectionalLight(0, (0.3, 1.0, 0.5), 2.0, True)
    tfs = [scenes.tf128(scenes.TF5), scenes.tf128(TF_B)]
    # the TF domain / opacity each request gets from the single-view path
    per_req = [configure_single(r, q, tfs) for q in REQUESTS]
    b = RenderBatch(r)
    tf_index = {}
    for q, (dom, op) in zip(REQUESTS, per_req):
        key = (q[4], dom, op)
        if key not in tf_index:
            tf_index[key] = b.addTransferFunction(tfs[q[4]], dom[0], dom[1], op)
        W, H, spp, integ, tf, az, el, _ = q
        b.addRequest(W, H, spp, tf_index[key], orbit_camera(r, W, H, az, el), integ)
    return r, b, tfs


def single_images(r, tfs, frames):
    import torch
    out = []
    for q in REQUESTS:
        configure_single(r, q, tfs)
        r.accumID = 0
        fb = torch.zeros((q[1], q[0]), dtype=torch.int32, device="cuda:0")
        for _ in range(frames):
            r.render(fb)
        torch.cuda.synchronize()
        out.append((r.accum(), fb.cpu().numpy().view(np.uint32)))
    return out


def oracle_images(sc, r, tfs, frames):
    """each request rendered by the oracle with the settings the single-view
    renderer takes for it (TF culling and majorants of its TF)"""
    out = []
    ranges = None
    for q in REQUESTS:
        configure_single(r, q, tfs)
        r.accumID = 0
        st = r.state()
        rgba = tfs[q[4]]
        xr = np.array(st.tfDomain, np.float32)
        om = oracle_model(sc, rgba, xr)
        if ranges is None:
            ranges = om.grid_build(np.array(st.gridDims, np.int32), np.array(st.gridBounds, np.float32))
        maj = ora.map_max_opacity(ranges, rgba, xr)
        acc = None
        for f in range(frames):
            p = ora.render_params_from_state(st, maj, rgba)
            p.accumID = f
            acc, fb = ora.render_f32(p, om, acc)
        out.append((acc, fb))
    return out


def test_batch_matches_oracle(cfg1, batch_setup):
    import torch
    print("case: batch of 6 requests against the oracle, 2 frames", flush=True)
    r, b, tfs = batch_setup
    b.resetAccum()
    for _ in range(2):
        b.render()
    torch.cuda.synchronize()
    for i, (acc_o, fb_o) in enumerate(oracle_images(cfg1, r, tfs, 2)):
        acc_b = b.accum(i)
        assert np.array_equal(acc_b, acc_o), f"request {i}: {np.count_nonzero(np.any(acc_b != acc_o, -1))} px differ"
        assert np.array_equal(b.frame(i), fb_o)


def test_batch_matches_single_view(batch_setup):
    import torch
    print("case: batch of 6 requests (2 TFs, odd sizes, spp 1-3, both integrators), 2 frames", flush=True)
    r, b, tfs = batch_setup
    b.resetAccum()
    for _ in range(2):
        b.render()
    torch.cuda.synchronize()
    ref = single_images(r, tfs, 2)
    for i, (acc_s, fb_s) in enumerate(ref):
        assert b.accumID(i) == 2
        acc_b = b.accum(i)
        assert np.array_equal(acc_b, acc_s), f"request {i}: {np.count_nonzero(np.any(acc_b != acc_s, -1))} px differ"
        assert np.array_equal(b.frame(i), fb_s)
        assert not np.all(acc_b == acc_b[:1, :1])  # something was rendered


@pytest.mark.parametrize("tuning", [(1, 0, 0), (64, 64, 64), (3, 24, 40)])
def test_batch_tuning_does_not_change_images(batch_setup, tuning):
    import torch
    print(f"case: batch tuning {tuning}", flush=True)
    r, b, tfs = batch_setup
    b.setTuning(*tuning)
    b.resetAccum()
    b.render()
    torch.cuda.synchronize()
    for i, (acc_s, _) in enumerate(single_images(r, tfs, 1)):
        assert np.array_equal(b.accum(i), acc_s), f"request {i}"
    b.setTuning(16, 16, 8)


def test_batch_graph_replay(batch_setup):
    """Three replays of a captured frame, with eager work and synchronisation
    between them, give the three eager frames."""
    import torch
    print("case: batch graph capture + 3 replays", flush=True)
    r, b, tfs = batch_setup
    b.resetAccum()
    for _ in range(3):
        b.render()
    torch.cuda.synchronize()
    eager = [b.accum(i) for i in range(len(b))]
    b.resetAccum()
    s = torch.cuda.Stream()
    b.captureGraph(s)
    x = torch.ones(1 << 20, device="cuda:0")
    for k in range(3):
        b.renderGraph(s)
        s.synchronize()
        x = x * 1.0001 + 1.0  # eager work between replays
        torch.cuda.synchronize()
    assert all(b.accumID(i) == 3 for i in range(len(b)))
    for i in range(len(b)):
        assert np.array_equal(b.accum(i), eager[i]), f"request {i}"


def test_batch_autotune(batch_setup):
    import torch
    print("case: batch autotune over 2x2x2 knobs", flush=True)
    r, b, tfs = batch_setup
    best, ms = b.autotune((4, 16), (0, 16), (0, 8))
    assert best[0] in (4, 16) and best[1] in (0, 16) and best[2] in (0, 8) and ms > 0
    assert b.tuning == best
    assert all(b.accumID(i) == 0 for i in range(len(b)))
    b.render()
    torch.cuda.synchronize()
    for i, (acc_s, _) in enumerate(single_images(r, tfs, 1)):
        assert np.array_equal(b.accum(i), acc_s), f"request {i}"


def test_batch_counters_cover_all_pixels(batch_setup):
    r, b, tfs = batch_setup
    b.resetAccum()
    c = b.render_counted()
    assert c["pixels"] == sum(q[0] * q[1] for q in REQUESTS)
    assert c["gridlet"] > 0 and c["dda_cells"] > 0
