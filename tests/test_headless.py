"""Headless harness host logic (viewer.cpp:119-158 defaults, 755-885 flags;
headless.cpp:285-298 xfRange): CPU only, no rendering."""
import io
import os

import numpy as np
import pytest

from owlexastitcher_amd import headless as hl
from owlexastitcher_amd import xf_load

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_defaults_follow_the_reference_viewer():
    c = hl.parse(["m.exum"])
    assert c.windowSize == (1024, 1024) and c.numMCs == (128, 128, 128) and c.camera["fovy"] == 70.0
    assert c.outFileName == "Witcher3.png" and c.fpsFileName == "fps.log" and c.spp == 1
    assert c.rendererType == 0 and c.shadeMode == 0 and not c.light["on"]
    assert c.clipPlanes == [{"enabled": False, "N": (0.0, 0.0, 1.0), "d": 1000.0}]
    assert c.frames == 510 and c.screenshotFrame == 10


def test_flags():
    c = hl.parse(["-grids", "g.grids", "-scalars", "s.scalars", "m.exum", "-xf", "t.xf", "--camera", "1", "2", "3",
                  "4", "5", "6", "0", "1", "0", "--fov", "40", "-win", "64", "48", "-o", "out", "-fps", "f.log",
                  "--spp", "4", "--heat-map", "0.5", "-dt", "0.25", "--remap-from", "0", "0", "0", "2", "2", "2",
                  "--remap-to", "1", "1", "1", "3", "3", "3", "--clip-plane", "0", "1", "0", "7", "-rt", "2",
                  "-sm", "1", "--num-mcs", "8", "16", "32", "--light", "1", "2", "3", "9", "--range", "0.1", "0.9",
                  "--subimg", "0", "0", "0.5", "0.5", "--frames", "20", "--screenshot-frame", "3"])
    assert (c.inFileName, c.gridsFileName, c.scalarFileName, c.xfFileName) == ("m.exum", "g.grids", "s.scalars",
                                                                               "t.xf")
    assert c.camera == {"vp": (1.0, 2.0, 3.0), "vi": (4.0, 5.0, 6.0), "vu": (0.0, 1.0, 0.0), "fovy": 40.0}
    assert c.windowSize == (64, 48) and (c.outFileName, c.fpsFileName) == ("out", "f.log")
    assert c.spp == 4 and c.heatMapEnabled and c.heatMapScale == 0.5 and c.dt == 0.25
    assert c.remap_from == (0, 0, 0, 2, 2, 2) and c.remap_to == (1, 1, 1, 3, 3, 3)
    assert c.clipPlanes[0] == {"enabled": True, "N": (0.0, 1.0, 0.0), "d": 7.0}
    assert (c.rendererType, c.shadeMode, c.numMCs) == (2, 1, (8, 16, 32))
    assert c.light == {"pos": (1.0, 2.0, 3.0), "intensity": 9.0, "on": True}
    assert c.valueRange == (0.1, 0.9) and c.subImage == (0, 0, 0.5, 0.5)
    assert (c.frames, c.screenshotFrame) == (20, 3)
    assert hl.parse(["-cells", "a.cells", "-scalar", "a.scalars"]).amrCellFileName == "a.cells"
    assert hl.parse(["-bricks", "a.bricks", "-scalars", "a.scalars"]).exaBrickFileName == "a.bricks"


@pytest.mark.parametrize("argv,msg", [
    (["m.exum", "--bogus"], "unknown cmdline arg"),
    (["m.exum", "-bm", "b"], "BigMesh"),
    (["m.exum", "-kdtree", "k"], "k-d tree"),
    (["m.exum", "--clip-plane", "0", "0", "1", "1", "--clip-plane", "0", "0", "1", "2"], "at most 1"),
    (["m.exum", "--camera", "1", "2"], "missing value"),
    ([], "no model"),
])
def test_usage_errors(argv, msg):
    with pytest.raises(hl.UsageError, match=msg):
        hl.parse(argv)


def test_main_reports_usage(capsys):
    assert hl.main(["--bogus"]) == 1
    assert "unknown cmdline arg" in capsys.readouterr().err


def test_xf_range_on_reference_files():
    """headless.cpp:1503-1507: absDomain := the model's value range, then
    lerped by relDomain / 100."""
    xf = xf_load(os.path.join(GOLDEN, "meteor-46k.xf"))
    assert np.allclose(xf["relDomain"], [0.0, 50.0])
    lo, hi = hl.xf_range([2.0, 6.0], xf["relDomain"])
    assert lo == 2.0 and hi == pytest.approx(4.0, abs=1e-6)
    lo, hi = hl.xf_range([6.0, 2.0], [100.0, 0.0])  # both intervals are ordered first
    assert (lo, hi) == (2.0, 6.0)


def test_png_writer_decodes():
    from PIL import Image
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (7, 5, 4), dtype=np.uint8)
    buf = io.BytesIO()
    path = os.path.join(os.environ.get("TMPDIR", "/tmp"), "exa_png_test.png")
    hl.write_png(path, img)
    with open(path, "rb") as f:
        buf.write(f.read())
    os.remove(path)
    buf.seek(0)
    got = np.asarray(Image.open(buf).convert("RGBA"))
    assert np.array_equal(got, img)


def test_default_color_map():
    cm = hl.default_color_map()
    assert cm.shape == (128, 4) and cm.dtype == np.float32
    assert np.all(np.diff(cm[:, 3]) > 0) and cm[0, 3] == 0.0 and cm[-1, 3] == 1.0
    assert np.allclose(cm[0, :3], np.array([59, 76, 192]) / 255.0)
