"""The element-stage queue of the persistent renderer: lanes whose gridlet
stage misses leave their path in a queue, a separate launch runs the element
stage over the queue with every lane busy, and the next launch resumes the
paths (queue rounds per frame, queue depth = entries per queue, workgroup
size).  Every lane keeps its own random stream and arithmetic order, so the
frames equal the oracle's bit for bit whatever the launch shape, including a
queue too small for the paths that wait (the excess runs the renderer's own
element pass) and the last round, whose lanes run the element stage
themselves."""
import numpy as np
import pytest

from owlexastitcher_amd import RENDER_PERSISTENT
from tests import scenes
from tests.test_gpu_render import _report, make_renderer, render_both

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg1():
    return scenes.config1()


@pytest.mark.parametrize("rounds,depth,block", [(1, 0, 256), (3, 0, 128), (8, 0, 64), (4, 37, 256), (32, 0, 256)])
@pytest.mark.parametrize("integrator,light", [(0, "dir"), (1, "point"), (0, "none")])
def test_queue_frames_match_oracle(cfg1, device, rounds, depth, block, integrator, light):
    print(f"case: element-stage queue rounds={rounds} depth={depth} block={block} integrator={integrator} "
          f"light={light}, config1 64x48, 2 frames", flush=True)
    r = make_renderer(cfg1, 64, 48)
    r.setScheme(RENDER_PERSISTENT)
    r.setType(integrator)
    r.setMaxBounces(4)
    b = r.modelBounds
    if light == "dir":
        r.setDirectionalLight(0, (0.3, 1.0, 0.5), 2.0, True)
    elif light == "point":
        r.setLightSource(0, b[3:] + 10.0, 60.0 * 500.0, True)
    r.setLaunchTuning(rounds, depth, block)
    assert r.tuning[3:] == (rounds, depth, block)
    acc, acc_o, fb, fb_o = render_both(cfg1, r, 2, device)
    assert np.array_equal(acc, acc_o), _report(acc, acc_o)
    assert np.array_equal(fb, fb_o)
    # the same work as without the queue
    r.accumID = 0
    c_q = r.render_counted()
    r.setLaunchTuning(0, 0, 256)
    r.accumID = 0
    c_0 = r.render_counted()
    assert c_q == c_0, (c_q, c_0)
    if integrator == 0:  # the path tracer's paths reach the stitching elements
        assert c_q["pyr"] + c_q["wedge"] + c_q["hex"] > 0


@pytest.mark.parametrize("rounds,depth,block", [(3, 0, 256), (2, 50, 64)])
def test_queue_batch_matches_oracle(cfg1, device, rounds, depth, block):
    """the batched launch with the queue: every request against the oracle"""
    import torch
    from tests.test_gpu_batch import REQUESTS, configure_single, oracle_images, orbit_camera, TF_B
    from owlexastitcher_amd import OWLRenderer, RenderBatch
    print(f"case: batch of {len(REQUESTS)} requests with the element-stage queue rounds={rounds} depth={depth} "
          f"block={block}, 2 frames", flush=True)
    r = OWLRenderer(cfg1.model, (32, 32, 32))
    r.setMaxBounces(4)
    r.setDirectionalLight(0, (0.3, 1.0, 0.5), 2.0, True)
    tfs = [scenes.tf128(scenes.TF5), scenes.tf128(TF_B)]
    per_req = [configure_single(r, q, tfs) for q in REQUESTS]
    b = RenderBatch(r)
    tf_index = {}
    for q, (dom, op) in zip(REQUESTS, per_req):
        key = (q[4], dom, op)
        if key not in tf_index:
            tf_index[key] = b.addTransferFunction(tfs[q[4]], dom[0], dom[1], op)
        W, H, spp, integ, tf, az, el, _ = q
        b.addRequest(W, H, spp, tf_index[key], orbit_camera(r, W, H, az, el), integ)
    b.setLaunchTuning(rounds, depth, block)
    assert b.launchTuning == (rounds, depth, block)
    b.resetAccum()
    for _ in range(2):
        b.render()
    torch.cuda.synchronize()
    for i, (acc_o, fb_o) in enumerate(oracle_images(cfg1, r, tfs, 2)):
        acc_b = b.accum(i)
        assert np.array_equal(acc_b, acc_o), f"request {i}: {np.count_nonzero(np.any(acc_b != acc_o, -1))} px differ"
        assert np.array_equal(b.frame(i), fb_o)


def test_autotune_launch_axes(cfg1, device):
    """the six-axis autotuner keeps a candidate of every axis and leaves the
    images unchanged"""
    import torch
    from tests.test_gpu_batch import REQUESTS, configure_single, orbit_camera, TF_B, single_images
    from owlexastitcher_amd import OWLRenderer, RenderBatch
    print("case: six-axis batch autotune", flush=True)
    r = OWLRenderer(cfg1.model, (32, 32, 32))
    r.setMaxBounces(4)
    r.setDirectionalLight(0, (0.3, 1.0, 0.5), 2.0, True)
    tfs = [scenes.tf128(scenes.TF5), scenes.tf128(TF_B)]
    per_req = [configure_single(r, q, tfs) for q in REQUESTS]
    b = RenderBatch(r)
    for q, (dom, op) in zip(REQUESTS, per_req):
        t = b.addTransferFunction(tfs[q[4]], dom[0], dom[1], op)
        W, H, spp, integ, tf, az, el, _ = q
        b.addRequest(W, H, spp, t, orbit_camera(r, W, H, az, el), integ)
    best, ms = b.autotuneLaunch((4, 8), (16,), (8,), (0, 2), (0, 64), (128, 256))
    assert best[0] in (4, 8) and best[3] in (0, 2) and best[4] in (0, 64) and best[5] in (128, 256) and ms > 0
    assert b.tuning == best[:3] and b.launchTuning == best[3:]
    b.render()
    torch.cuda.synchronize()
    for i, (acc_s, _) in enumerate(single_images(r, tfs, 1)):
        assert np.array_equal(b.accum(i), acc_s), f"request {i}"
