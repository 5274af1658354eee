"""The headless harness end to end on the GPU: model files written by the
dual-mesh pipeline, the reference's .xf file, a few frames; the log, fps and
PNG files follow headless.cpp:97-186, and the renderer carries the .xf
settings (headless.cpp:1497-1510)."""
import os

import numpy as np
import pytest

from owlexastitcher_amd import AmrScene, headless as hl

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("headless")
    a = AmrScene.noise(16, 2, 0.55, 0.7, 6, 2.2, 5)
    paths = {k: str(d / f"m.{k}") for k in ("exum", "grids", "scalars", "cells", "bricks")}
    a.write_dual_mesh(paths["exum"], None, paths["grids"])
    a.save(paths["cells"], paths["scalars"])
    a.write_bricks(paths["bricks"], None, 8)
    return d, paths


def test_headless_run_with_model_files(files, device):
    from PIL import Image
    d, p = files
    out, fps = str(d / "run1"), str(d / "fps.log")
    print("case: headless 64x48, 13 frames, model files + cloud.xf", flush=True)
    s = hl.run([p["exum"], "-grids", p["grids"], "-scalars", p["scalars"], "-xf", os.path.join(GOLDEN, "cloud.xf"),
                "--size", "64", "48", "-o", out, "-fps", fps, "--frames", "13", "--num-mcs", "16", "16", "16"])
    assert s["timed_frames"] == 3 and s["fps"] > 0
    log = open(out + ".log").read()
    assert "FRAME_ID;SEC.\n" in log and "avg. secs over 3 time" in log and log.rstrip().endswith("FPS")
    frames = [line for line in log.splitlines() if ";" in line and line.split(";")[0].isdigit()]
    assert [int(f.split(";")[0]) for f in frames] == list(range(1, 14))
    line = open(fps).read().strip().splitlines()[-1].split(",")
    assert line[1] == out and float(line[2]) == pytest.approx(s["fps"], rel=1e-3)
    img = np.asarray(Image.open(out + ".png").convert("RGBA"))
    assert img.shape == (48, 64, 4) and np.all(img[..., 3] == 255)
    r = s["renderer"]
    st = r.state()
    xf = hl.xf_load(os.path.join(GOLDEN, "cloud.xf"))
    lo, hi = hl.xf_range(r.valueRange, xf["relDomain"])
    # setRange(xfRange) then setRelDomain(relDomain): the renderer lerps the
    # range by the relative domain once more (OWLRenderer.cpp:647-652)
    rel = np.asarray(xf["relDomain"], np.float64) / 100.0
    want = lo + rel * (hi - lo)
    assert np.allclose(list(st.tfDomain), want, rtol=1e-5, atol=1e-7), (list(st.tfDomain), want)
    assert st.integrator == 0 and st.W == 64 and st.H == 48
    # the screenshot is frame 10's frame buffer, rows flipped
    fb = s["viewer"].fb.cpu().numpy().view(np.uint32)
    assert len(np.unique(fb)) > 10  # not a blank image


def test_headless_run_from_cells(files, device):
    d, p = files
    out = str(d / "run2")
    print("case: headless from .cells, ray marcher, 32x32, 3 frames", flush=True)
    s = hl.run(["-cells", p["cells"], "-scalars", p["scalars"], "--size", "32", "32", "-o", out, "-fps",
                str(d / "fps2.log"), "--frames", "3", "--screenshot-frame", "1", "-rt", "2", "-dt", "0.5", "-sm", "1",
                "--light", "0", "0", "0", "1", "--clip-plane", "0", "0", "1", "30"])
    st = s["renderer"].state()
    assert st.integrator == 2 and st.shadeMode == 1 and st.dt == 0.5 and st.clipEnabled == 1
    assert st.lightOn[0] == 1 and os.path.exists(out + ".png") and s["timed_frames"] == 2


def test_headless_run_with_bricks(files, device):
    d, p = files
    out = str(d / "run3")
    print("case: headless -bricks, direct lighting, 32x32, 3 frames", flush=True)
    s = hl.run(["-bricks", p["bricks"], "-scalars", p["scalars"], "--size", "32", "32", "-o", out, "-fps",
                str(d / "fps3.log"), "--frames", "3", "--screenshot-frame", "1", "-rt", "1",
                "-xf", os.path.join(GOLDEN, "cloud.xf")])
    r = s["renderer"]
    assert r.bricks and r.state().integrator == 1
    assert os.path.exists(out + ".png") and s["timed_frames"] == 2
    fb = s["viewer"].fb.cpu().numpy().view(np.uint32)
    assert len(np.unique(fb)) > 10
