"""Where the persistent renderer runs the element stage: in the waiting
lanes' wave, through the workgroup's queue in LDS (any wave answers up to 64
posted queries at once), or through a device queue between launches (lanes
leave their path in the queue, a separate launch runs the element stage over
the queue with every lane busy, and the next launch resumes the paths; queue
rounds per frame, queue depth = entries per queue); and the workgroup size.  Every lane keeps its own random stream and arithmetic order, so the
frames equal the oracle's bit for bit whatever the launch shape, including a
queue too small for the paths that wait (the excess runs the renderer's own
element pass) and the last round, whose lanes run the element stage
themselves."""
import numpy as np
import pytest

from owlexastitcher_amd import ELEMENT_IN_WAVE, ELEMENT_QUEUE_ROUNDS, ELEMENT_WORKGROUP_QUEUE, RENDER_PERSISTENT
from tests import scenes
from tests.test_gpu_render import _report, make_renderer, render_both

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg1():
    return scenes.config1()


SHAPES = [(ELEMENT_QUEUE_ROUNDS, 1, 0, 256, 16), (ELEMENT_QUEUE_ROUNDS, 3, 0, 128, 16),
          (ELEMENT_QUEUE_ROUNDS, 8, 0, 64, 16), (ELEMENT_QUEUE_ROUNDS, 4, 37, 256, 16),
          (ELEMENT_QUEUE_ROUNDS, 32, 0, 256, 16), (ELEMENT_WORKGROUP_QUEUE, 1, 0, 256, 16),
          (ELEMENT_WORKGROUP_QUEUE, 1, 0, 64, 1), (ELEMENT_WORKGROUP_QUEUE, 1, 0, 128, 64)]


@pytest.mark.parametrize("stage,rounds,depth,block,batch", SHAPES)
@pytest.mark.parametrize("integrator,light", [(0, "dir"), (1, "point"), (0, "none")])
def test_queue_frames_match_oracle(cfg1, device, stage, rounds, depth, block, batch, integrator, light):
    print(f"case: element stage {stage} rounds={rounds} depth={depth} block={block} batch={batch} "
          f"integrator={integrator} light={light}, config1 64x48, 2 frames", flush=True)
    r = make_renderer(cfg1, 64, 48)
    r.setScheme(RENDER_PERSISTENT)
    r.setType(integrator)
    r.setMaxBounces(4)
    b = r.modelBounds
    if light == "dir":
        r.setDirectionalLight(0, (0.3, 1.0, 0.5), 2.0, True)
    elif light == "point":
        r.setLightSource(0, b[3:] + 10.0, 60.0 * 500.0, True)
    r.setPersistentTuning(6, batch, 16)
    r.setLaunchTuning(stage, rounds, depth, block)
    assert r.tuning[3:] == (stage, rounds, depth, block)
    acc, acc_o, fb, fb_o = render_both(cfg1, r, 2, device)
    assert np.array_equal(acc, acc_o), _report(acc, acc_o)
    assert np.array_equal(fb, fb_o)
    # the same work as without the queue
    r.accumID = 0
    c_q = r.render_counted()
    r.setLaunchTuning(ELEMENT_IN_WAVE, 1, 0, 256)
    r.accumID = 0
    c_0 = r.render_counted()
    assert c_q == c_0, (c_q, c_0)
    if integrator == 0:  # the path tracer's paths reach the stitching elements
        assert c_q["pyr"] + c_q["wedge"] + c_q["hex"] > 0


@pytest.mark.parametrize("stage,rounds,depth,block", [(ELEMENT_QUEUE_ROUNDS, 3, 0, 256), (ELEMENT_QUEUE_ROUNDS, 2, 50, 64),
                                                  (ELEMENT_WORKGROUP_QUEUE, 1, 0, 128)])
def test_queue_batch_matches_oracle(cfg1, device, stage, rounds, depth, block):
    """the batched launch with the queue: every request against the oracle"""
    import torch
    from tests.test_gpu_batch import REQUESTS, configure_single, oracle_images, orbit_camera, TF_B
    from owlexastitcher_amd import OWLRenderer, RenderBatch
    print(f"case: batch of {len(REQUESTS)} requests, element stage {stage} rounds={rounds} depth={depth} "
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
    b.setLaunchTuning(stage, rounds, depth, block)
    assert b.launchTuning == (stage, rounds, depth, block)
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
    best, ms = b.autotuneLaunch((4, 8), (16,), (8,), (0, 1, 2), (2,), (0, 64), (128, 256))
    assert best[0] in (4, 8) and best[3] in (0, 1, 2) and best[5] in (0, 64) and best[6] in (128, 256) and ms > 0
    assert b.tuning == best[:3] and b.launchTuning == best[3:]
    b.render()
    torch.cuda.synchronize()
    for i, (acc_s, _) in enumerate(single_images(r, tfs, 1)):
        assert np.array_equal(b.accum(i), acc_s), f"request {i}"
