"""The C++ host side on the GPU: the C++ headless harness (exa_headless, on
include/exa/OWLRenderer.h) and the Python harness give the same screenshot,
pixel for pixel, for the same command line; its log and fps files follow
headless.cpp:97-186; a program using the C++ classes directly renders and
builds a sampler."""
import os
import subprocess

import numpy as np
import pytest

from owlexastitcher_amd import AmrScene, headless as hl
from tests.test_cpp_host import HEADLESS, PROGRAM, ROOT

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("cpp_headless")
    a = AmrScene.noise(16, 2, 0.55, 0.7, 6, 2.2, 5)
    paths = {k: str(d / f"m.{k}") for k in ("exum", "grids", "scalars", "cells", "bricks")}
    a.write_dual_mesh(paths["exum"], None, paths["grids"])
    a.save(paths["cells"], paths["scalars"])
    a.write_bricks(paths["bricks"], None, 8)
    return d, paths


def _png(path):
    from PIL import Image
    return np.asarray(Image.open(path).convert("RGBA"))


CASES = {
    "files_xf_camera": lambda p: [p["exum"], "-grids", p["grids"], "-scalars", p["scalars"], "-xf",
                                  os.path.join(GOLDEN, "cloud.xf"), "--camera", "-20", "70", "90", "32", "32", "32",
                                  "0", "1", "0", "-fovy", "55", "--num-mcs", "16", "16", "16"],
    "default_camera_dl_light": lambda p: [p["exum"], "-grids", p["grids"], "-scalars", p["scalars"], "-rt", "1",
                                          "--light", "80", "90", "100", "5000", "--num-mcs", "8", "8", "8"],
    "cells_marcher_clip": lambda p: ["-cells", p["cells"], "-scalars", p["scalars"], "-rt", "2", "-dt", "0.5",
                                     "-sm", "1", "--clip-plane", "0", "0", "1", "30", "-spp", "2"],
    "bricks_xf": lambda p: ["-bricks", p["bricks"], "-scalars", p["scalars"], "-xf",
                            os.path.join(GOLDEN, "cloud.xf")],
    "subimg_range_remap": lambda p: [p["exum"], "-grids", p["grids"], "-scalars", p["scalars"], "--subimg", "0.2",
                                     "0.1", "0.7", "0.8", "--range", "0.1", "0.9", "-spp", "3", "--remap-from", "0",
                                     "0", "0", "64", "64", "64", "--remap-to", "0", "0", "0", "32", "32", "32"],
}


@pytest.mark.parametrize("case", list(CASES))
def test_cpp_and_python_harness_same_image(files, device, case):
    d, p = files
    print(f"case: C++ vs Python headless, {case}, 64x48, 12 frames", flush=True)
    args = CASES[case](p) + ["--size", "64", "48", "--frames", "12"]
    out_c, out_p = str(d / f"{case}_cpp"), str(d / f"{case}_py")
    r = subprocess.run([HEADLESS] + args + ["-o", out_c, "-fps", str(d / "fps_cpp.log")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    s = hl.run(args + ["-o", out_p, "-fps", str(d / "fps_py.log")])
    img_c, img_p = _png(out_c + ".png"), _png(out_p + ".png")
    assert img_c.shape == (48, 64, 4)
    assert np.array_equal(img_c, img_p), f"{np.count_nonzero(np.any(img_c != img_p, -1))} pixels differ"
    assert len(np.unique(img_c.reshape(-1, 4), axis=0)) > 10
    log = open(out_c + ".log").read()
    assert "FRAME_ID;SEC.\n" in log and "avg. secs over 2 time" in log and log.rstrip().endswith("FPS")
    frames = [line for line in log.splitlines() if ";" in line and line.split(";")[0].isdigit()]
    assert [int(f.split(";")[0]) for f in frames] == list(range(1, 13))
    last = open(d / "fps_cpp.log").read().strip().splitlines()[-1].split(",")
    assert last[1] == out_c and float(last[2]) > 0 and s["timed_frames"] == 2


def test_cpp_classes_render_and_build(files, tmp_path, device):
    print("case: program on the C++ classes, 4x4 frame + sampler build", flush=True)
    d, p = files
    src = tmp_path / "use_header.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "use_header"
    lib_dir = os.path.join(ROOT, "owlexastitcher_amd")
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe), "-L",
                    lib_dir, "-lexastitch", f"-Wl,-rpath,{lib_dir}"], check=True, timeout=300)
    r = subprocess.run([str(exe), p["exum"], p["scalars"]], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["1", "16", "3", "1", "1", "0",
                                "ExaBrickSampler", "0", "0", "ExaStitchSampler", "1", "1", "64", "1",
                                "ExaStitchSampler", "ExaBrickSampler", "1"]
