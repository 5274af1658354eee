"""On-disk formats of the ExaStitch pipeline (host code, CPU only).

Each format is checked against an independent numpy reading of its byte
layout (makeDual.cpp:932-940 .cells; ExaStitchModel.cpp:44-54 .scalars;
makeDual.cpp:445-449 / makeGrids.cpp:54-66 .cubes; makeGrids.cpp:240-246 and
ExaStitchModel.cpp:191-213 .grids; headless.cpp:258-283 .xf), and the model
loader (ExaStitchModel.cpp:24-257) against hand-built files whose expected
model follows from the reference's rules: vertex value = scalars[vertexTag],
unreferenced vertices compacted in order when a grids file is given, a
gridlet scalar ID outside the scalars is a NaN hole, NaN ignored by the value
range.  The .xf files under tests/golden are the reference's own TF files.
"""
import os
import struct

import numpy as np
import pytest

from owlexastitcher_amd import AmrScene, DualMesh, ExaError, ExaStitchModel, make_grids, xf_load, xf_save

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_cells_scalars_roundtrip(tmp_path):
    a = AmrScene.noise(16, 2, 0.55, 0.7, 6, 2.2, 5)
    cells, scalars = a.arrays()
    c, s = str(tmp_path / "x.cells"), str(tmp_path / "x.scalars")
    a.save(c, s)
    assert np.array_equal(np.fromfile(c, np.int32).reshape(-1, 4), cells)
    assert np.array_equal(bits(np.fromfile(s, np.float32)), bits(scalars))
    b = AmrScene.load(c, s)
    c2, s2 = b.arrays()
    assert np.array_equal(c2, cells) and np.array_equal(bits(s2), bits(scalars))
    # the same scene in a dual mesh either way
    assert DualMesh(a).counts() == DualMesh(b).counts()


def test_cells_errors(tmp_path):
    c, s = tmp_path / "bad.cells", tmp_path / "bad.scalars"
    c.write_bytes(b"\0" * 20)
    s.write_bytes(b"\0" * 8)
    with pytest.raises(ExaError, match="multiple of 16"):
        AmrScene.load(str(c), str(s))
    np.array([[0, 0, 0, 0], [1, 0, 0, 0]], np.int32).tofile(str(c))
    np.array([1.0], np.float32).tofile(str(s))
    with pytest.raises(ExaError, match="fewer scalars"):
        AmrScene.load(str(c), str(s))
    with pytest.raises(ExaError, match="cannot open"):
        AmrScene.load(str(tmp_path / "missing.cells"), str(s))


def parse_grids(path, scalars):
    """Independent reading of a .grids file: per gridlet lower[3], level,
    dims[3], then (dims+1)^3 int32 scalar IDs."""
    raw = np.fromfile(path, np.int32)
    gl, gs, off = [], [], 0
    while off < len(raw):
        lower, level, dims = raw[off:off + 3], raw[off + 3], raw[off + 4:off + 7]
        ns = int(np.prod(dims + 1))
        ids = raw[off + 7:off + 7 + ns]
        vals = np.full(ns, np.nan, np.float32)
        ok = (ids >= 0) & (ids < len(scalars))
        vals[ok] = scalars[ids[ok]]
        gl.append([*lower, level, *dims, sum(len(x) for x in gs)])
        gs.append(vals)
        off += 7 + ns
    return np.array(gl, np.int32).reshape(-1, 8), (np.concatenate(gs) if gs else np.zeros(0, np.float32))


def parse_element_mesh(path):
    raw = open(path, "rb").read()
    assert raw[:8] == b"EXAELEM1"
    nv, nt, npy, nw, nh, flags = struct.unpack_from("<6Q", raw, 8)
    off = 56
    xyz = np.frombuffer(raw, np.float32, 3 * nv, off).reshape(-1, 3)
    off += 12 * nv
    tag = vals = None
    if flags & 1:
        tag = np.frombuffer(raw, np.int32, nv, off)
        off += 4 * nv
    if flags & 2:
        vals = np.frombuffer(raw, np.float32, nv, off)
        off += 4 * nv
    elems = []
    for n, k in ((nt, 4), (npy, 5), (nw, 6), (nh, 8)):
        elems.append(np.frombuffer(raw, np.int32, n * k, off).reshape(-1, k))
        off += 4 * n * k
    assert off == len(raw)
    return xyz, tag, vals, elems


def test_dual_mesh_files_load_as_in_memory_model(tmp_path):
    a = AmrScene.noise(16, 2, 0.55, 0.7, 6, 2.2, 5)
    cells, scalars = a.arrays()
    em, pre, gr, sc = (str(tmp_path / n) for n in ("m.exum", "m", "m.grids", "m.scalars"))
    counts = a.write_dual_mesh(em, pre, gr)
    a.save(None, sc)
    mem = ExaStitchModel.from_dual_mesh(DualMesh(a))
    assert counts == DualMesh(a).counts()
    disk = ExaStitchModel.load(em, gr, sc)
    for x, y in zip(disk.arrays(), mem.arrays()):
        assert x.shape == y.shape
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))  # bit-exact, NaN holes included
    for x, y in zip(disk.bounds(), mem.bounds()):
        assert np.array_equal(x, y)
    v, idx, g, gs = mem.arrays()
    # independent reading of the files
    xyz, tag, vals, elems = parse_element_mesh(em)
    assert np.array_equal(xyz, v[:, :3]) and np.array_equal(bits(vals), bits(v[:, 3]))
    assert np.array_equal(bits(scalars[tag]), bits(v[:, 3]))  # vertexTag = the vertex's cell
    rows = np.concatenate([np.pad(e, ((0, 0), (0, 8 - e.shape[1])), constant_values=-1) for e in elems])
    assert np.array_equal(rows, idx)
    g2, gs2 = parse_grids(gr, scalars)
    assert np.array_equal(g2, g) and np.array_equal(bits(gs2), bits(gs))
    # vertex positions are cell centres of the cells they are tagged with
    w = (1 << cells[tag, 3]).astype(np.float32)
    assert np.array_equal(v[:, :3], cells[tag, :3] + 0.5 * w[:, None])


def test_cubes_files_and_make_grids(tmp_path):
    a = AmrScene.noise(16, 2, 0.55, 0.7, 6, 2.2, 5)
    cells, scalars = a.arrays()
    pre, gr = str(tmp_path / "m"), str(tmp_path / "direct.grids")
    counts = a.write_dual_mesh(None, pre, gr)
    files = sorted((f for f in os.listdir(tmp_path) if f.endswith(".cubes")), key=lambda f: int(f[2:-6]))
    assert files, "no cubes written"
    total = 0
    for f in files:
        L = int(f[2:-6])
        rec = np.fromfile(str(tmp_path / f), np.dtype([("lower", "<f4", 3), ("level", "<i4"), ("ids", "<i4", 8)]))
        total += len(rec)
        assert np.all(rec["level"] == L)
        w = float(1 << L)
        cid = (rec["lower"] - 0.5 * w) / w
        assert np.array_equal(cid, np.floor(cid))  # lower = cell centre of the lowest corner
        # the 8 corners are same-level cells around the dual cube, lowest first
        corner = cells[rec["ids"]]
        assert np.all(corner[:, :, 3] == L)
        assert np.array_equal(corner[:, 0, :3], (cid * w).astype(np.int32))
    assert total == counts["cubes"]
    via = str(tmp_path / "via.grids")
    assert make_grids([str(tmp_path / f) for f in files], via) == 0
    assert open(via, "rb").read() == open(gr, "rb").read()
    with pytest.raises(ExaError, match="not a cubes file"):
        make_grids([str(tmp_path / "direct.grids")], via)


def write_elem(path, xyz, tets=(), tag=None, vals=None, hexes=()):
    tets = np.asarray(tets, np.int32).reshape(-1, 4)
    hexes = np.asarray(hexes, np.int32).reshape(-1, 8)
    flags = (1 if tag is not None else 0) | (2 if vals is not None else 0)
    with open(path, "wb") as f:
        f.write(b"EXAELEM1")
        f.write(struct.pack("<6Q", len(xyz), len(tets), 0, 0, len(hexes), flags))
        f.write(np.asarray(xyz, np.float32).tobytes())
        if tag is not None:
            f.write(np.asarray(tag, np.int32).tobytes())
        if vals is not None:
            f.write(np.asarray(vals, np.float32).tobytes())
        f.write(tets.tobytes())
        f.write(hexes.tobytes())


def write_grids(path, recs):
    with open(path, "wb") as f:
        for lower, level, dims, ids in recs:
            f.write(np.array([*lower, level, *dims], np.int32).tobytes())
            f.write(np.asarray(ids, np.int32).tobytes())


def test_model_loader_rules(tmp_path):
    sc, em, gr = str(tmp_path / "s.scalars"), str(tmp_path / "m.exum"), str(tmp_path / "m.grids")
    np.array([10, 20, 30, 40], np.float32).tofile(sc)
    xyz = [[0, 0, 0], [4, 0, 0], [0, 4, 0], [0, 0, 4], [9, 9, 9], [4, 4, 4]]
    write_elem(em, xyz, tets=[[1, 2, 3, 5]], tag=[3, 0, 1, 2, 0, 1], vals=[7, 7, 7, 7, 7, 7])
    ids = [0, 1, 2, 3, -1, 4, 2 ** 31 - 1, 3]
    write_grids(gr, [((0, 0, 0), 1, (1, 1, 1), ids)])
    # grids given: vertices 0 and 4 are dropped, order kept
    v, idx, g, gs = ExaStitchModel.load(em, gr, sc).arrays()
    assert np.array_equal(v, np.array([[4, 0, 0, 10], [0, 4, 0, 20], [0, 0, 4, 30], [4, 4, 4, 20]], np.float32))
    assert np.array_equal(idx, [[0, 1, 2, 3, -1, -1, -1, -1]])
    assert np.array_equal(g, [[0, 0, 0, 1, 1, 1, 1, 0]])
    want = np.array([10, 20, 30, 40, np.nan, np.nan, np.nan, 40], np.float32)
    assert np.array_equal(bits(gs), bits(want))
    cb, vr = ExaStitchModel.load(em, gr, sc).bounds()
    assert np.array_equal(vr, [10, 40])  # NaN holes do not enter the range
    assert np.array_equal(cb, [0, 0, 0, 4, 4, 4])  # elements [0,4]^3 and the gridlet box [1,3]^3
    # no grids: no compaction
    v, idx, _, _ = ExaStitchModel.load(em, None, sc).arrays()
    assert len(v) == 6 and np.array_equal(idx[0, :4], [1, 2, 3, 5])
    assert np.array_equal(v[:, 3], [40, 10, 20, 30, 10, 20])
    # no scalars: the stored per-vertex values
    v, _, _, _ = ExaStitchModel.load(em, None, None).arrays()
    assert np.all(v[:, 3] == 7)
    # gridlets only, no scalars file: every gridlet scalar is a hole
    _, _, _, gs = ExaStitchModel.load(None, gr, None).arrays()
    assert np.all(np.isnan(gs))


def test_model_loader_errors(tmp_path):
    sc, em, gr = str(tmp_path / "s.scalars"), str(tmp_path / "m.exum"), str(tmp_path / "m.grids")
    np.array([1, 2], np.float32).tofile(sc)
    write_elem(em, [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], tets=[[0, 1, 2, 3]], tag=[0, 1, 2, 1])
    with pytest.raises(ExaError, match="vertex tag 2"):
        ExaStitchModel.load(em, None, sc)
    write_elem(em, [[0, 0, 0]] * 4, tets=[[0, 1, 2, 4]])
    with pytest.raises(ExaError, match="out of range"):
        ExaStitchModel.load(em, None, None)
    with open(em, "wb") as f:
        f.write(b"NOTAMESH" + b"\0" * 48)
    with pytest.raises(ExaError, match="not an element-mesh file"):
        ExaStitchModel.load(em, None, None)
    # an element count whose size wraps the 64-bit byte sum to the file length
    # (2^60 tets x 16 B = 2^64) is refused, not read as zero elements
    with open(em, "wb") as f:
        f.write(b"EXAELEM1" + struct.pack("<6Q", 1, 2 ** 60, 0, 0, 0, 0) + bytes(12))
    with pytest.raises(ExaError, match="exceeds 2\\^31-1"):
        ExaStitchModel.load(em, None, None)
    write_grids(gr, [((0, 0, 0), 0, (1, 1, 1), [0] * 8)])
    with open(gr, "ab") as f:
        f.write(np.array([0, 0, 0, 0, 1, 1, 1, 0, 0], np.int32).tobytes())  # second record cut short
    with pytest.raises(ExaError, match="truncated"):
        ExaStitchModel.load(None, gr, sc)
    with pytest.raises(ExaError, match="no elements and no gridlets"):
        ExaStitchModel.load(None, None, sc)


@pytest.mark.parametrize("name", ["cloud.xf", "meteor-46k.xf"])
def test_xf_reference_files(tmp_path, name):
    path = os.path.join(GOLDEN, name)
    raw = open(path, "rb").read()
    magic, = struct.unpack_from("<Q", raw, 0)
    assert magic == 0x1235ABC000
    params = struct.unpack_from("<5f", raw, 8)
    n, = struct.unpack_from("<i", raw, 28)
    cm = np.frombuffer(raw, np.float32, 4 * n, 32).reshape(n, 4)
    xf = xf_load(path)
    assert xf["opacityScale"] == params[0]
    assert np.array_equal(xf["absDomain"], params[1:3]) and np.array_equal(xf["relDomain"], params[3:5])
    assert np.array_equal(xf["colorMap"], cm) and len(cm) == 128
    out = str(tmp_path / name)
    xf_save(out, xf)
    assert open(out, "rb").read() == raw
