"""The C-ABI library on the CPU: it loads, exports every entry point that
include/exa/exa.h declares, validates its inputs (error codes + message), and
its host-side builders (scene generators, model bounds, camera, TF helpers)
behave as the reference's host code.  No GPU compute is called here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import oracle as ora
from owlexastitcher_amd import (AmrScene, ExaError, ExaStitchModel, camera_look_at, lib, tf_rasterize, xf_load)
from owlexastitcher_amd._capi import EXA_ERR_INVALID, EXPORTED, StitchModelDesc
from tests import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "exa", "exa.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(exa_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = lib()
    names = header_functions()
    assert len(names) > 60
    for n in names:
        assert hasattr(L, n), n
    # the Python binding covers the whole header
    assert set(names) == set(EXPORTED)


def test_version_and_error_message():
    assert lib().exa_version() >= 1
    bad = np.array([0, 1, -1, 2, -1, -1, -1, -1], np.int32)  # -1 before a valid index
    verts = np.zeros((3, 4), np.float32)
    with pytest.raises(ExaError) as e:
        ExaStitchModel.from_arrays(verts, bad, np.zeros((0, 8), np.int32), np.zeros(0, np.float32))
    assert e.value.code == EXA_ERR_INVALID
    assert "padding" in str(e.value)


@pytest.mark.parametrize("row,msg", [
    ([0, 1, 2, -1, -1, -1, -1, -1], "3 vertices"),
    ([0, 1, 2, 3, 4, 5, 6, -1], "7 vertices"),
    ([0, 1, 2, 9, -1, -1, -1, -1], "out of range"),
])
def test_model_rejects_bad_element_rows(row, msg):
    verts = np.zeros((8, 4), np.float32)
    with pytest.raises(ExaError, match=msg):
        ExaStitchModel.from_arrays(verts, np.array(row, np.int32), np.zeros((0, 8), np.int32),
                                   np.zeros(0, np.float32))


def test_model_rejects_bad_gridlets():
    g = np.array([[0, 0, 0, 0, 2, 2, 2, 0]], np.int32)
    with pytest.raises(ExaError, match="scalars out of range"):
        ExaStitchModel.from_arrays(np.zeros((0, 4)), np.zeros((0, 8), np.int32), g, np.zeros(26, np.float32))
    g[0, 4] = 0
    with pytest.raises(ExaError, match="invalid level/dims"):
        ExaStitchModel.from_arrays(np.zeros((0, 4)), np.zeros((0, 8), np.int32), g, np.zeros(27, np.float32))


def test_null_arguments_are_errors():
    L = lib()
    assert L.exa_model_create(None, None) == EXA_ERR_INVALID
    assert L.exa_renderer_set_spp(None, 1) == EXA_ERR_INVALID
    assert L.exa_sampler_get_grid_info(None, None, None) == EXA_ERR_INVALID
    assert b"NULL" in L.exa_get_last_error() or len(L.exa_get_last_error()) > 0


def test_model_bounds_match_oracle():
    """cellBounds / valueRange as ExaStitchModel::load computes them
    (ExaStitchModel.cpp:78-99, 224-245), NaN scalars ignored."""
    sc = scenes.small()
    om = ora.Model(sc.verts, sc.idx, sc.gridlets, sc.gscalars)
    cb, vr = om.bounds()
    assert np.array_equal(sc.cell_bounds, cb)
    assert np.array_equal(sc.value_range, vr)
    assert np.isnan(sc.gscalars).any()  # the scene has NaN holes
    m2 = ExaStitchModel.from_arrays(sc.verts, sc.idx, sc.gridlets, sc.gscalars)
    cb2, vr2 = m2.bounds()
    assert np.array_equal(cb2, cb) and np.array_equal(vr2, vr)
    assert m2.counts()["elements"] == len(sc.idx)


def test_gaussian_scene_matches_its_definition():
    amr = AmrScene.gaussians(64, 1, 32, 96, 8, 1)
    cells, scal = amr.arrays()
    assert len(cells) == 64 ** 3 - 32 ** 3 + 64 ** 3
    assert set(np.unique(cells[:, 3])) == {0, 1}
    fine = cells[cells[:, 3] == 0, :3]
    assert fine.min() == 32 and fine.max() == 95
    P = amr.params().reshape(8, 5)
    assert (P[:, :3] >= 0).all() and (P[:, :3] <= 128).all()
    assert (P[:, 3] >= 6).all() and (P[:, 3] <= 20).all()
    assert (P[:, 4] >= 0.2).all() and (P[:, 4] <= 1).all()
    rng = np.random.default_rng(0)
    for i in rng.integers(0, len(cells), 200):
        c = cells[i, :3] + 0.5 * (1 << cells[i, 3])
        v = np.sum(P[:, 4] * np.exp(-np.sum((c - P[:, :3]) ** 2, 1) / (2 * P[:, 3] ** 2)))
        assert abs(scal[i] - v) <= 1e-6 * max(v, 1e-30) + 1e-12


def test_noise_scene_is_balanced():
    amr = AmrScene.noise(16, 2, 0.55, 0.7, 6, 2.0, 5)
    cells, scal = amr.arrays()
    assert np.sum((1 << cells[:, 3]).astype(np.int64) ** 3) == 64 ** 3
    assert (scal >= 0).all() and (scal <= 1).all()
    # vertex 2:1 balance: no level-0 cell touches a level-2 cell
    lvl = np.full((64, 64, 64), -1, np.int8)
    for (x, y, z, L) in cells:
        w = 1 << L
        lvl[x:x + w, y:y + w, z:z + w] = L
    for d in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1), (1, -1, 0)]:
        a = lvl[max(0, d[0]):64 + min(0, d[0]), max(0, d[1]):64 + min(0, d[1]), max(0, d[2]):64 + min(0, d[2])]
        b = lvl[max(0, -d[0]):64 - max(0, d[0]), max(0, -d[1]):64 - max(0, d[1]), max(0, -d[2]):64 - max(0, d[2])]
        assert np.abs(a.astype(int) - b.astype(int)).max() <= 1


def test_camera_look_at_centre_ray():
    W, H = 64, 48
    frm, at = np.array([10.0, 20.0, 30.0]), np.array([1.0, 2.0, 3.0])
    org, d00, du, dv = camera_look_at(frm, at, (0, 1, 0), 60.0, W, H)
    assert np.allclose(org, frm)
    centre = d00 + W / 2 * du + H / 2 * dv
    t = at - frm
    assert np.allclose(centre / np.linalg.norm(centre), t / np.linalg.norm(t), atol=1e-5)
    # vertical extent: fovy
    top = d00 + W / 2 * du + H * dv
    bot = d00 + W / 2 * du
    ang = np.degrees(np.arccos(np.dot(top, bot) / np.linalg.norm(top) / np.linalg.norm(bot)))
    assert abs(ang - 60.0) < 1e-3
    assert abs(np.linalg.norm(du) * W / (np.linalg.norm(dv) * H) - W / H) < 1e-5


def test_tf_rasterize_hits_control_points():
    rgba = tf_rasterize(scenes.TF5, 21)  # texel i sits at x = i/20: every control point is a texel
    for p in scenes.TF5:
        i = int(round(p[0] * 20))
        assert np.allclose(rgba[i], p[1:], atol=1e-6)
    assert np.all(np.diff(rgba[:, 3]) >= -1e-7)  # alpha of TF5 is monotone


def test_xf_loader_on_reference_data():
    """headless.cpp:258-283 .xf layout, on the reference's own TF file."""
    xf = xf_load(os.path.join(ROOT, "tests", "golden", "cloud.xf"))
    assert xf["colorMap"].shape == (128, 4)
    assert xf["opacityScale"] == 100.0
    assert np.allclose(xf["absDomain"], [0, 1])
    with pytest.raises(ExaError, match="not a .xf"):
        xf_load(os.path.join(ROOT, "include", "exa", "exa.h"))


def test_model_desc_layout():
    assert C.sizeof(StitchModelDesc) == 64
