"""ExaBricks comparison sampler on the CPU: the oracle pinned against known
answers and invariants, and the library's host side (brick maker, region
build, file formats) checked against the oracle.

- the reference's 2-brick test data (tools/makeTestData.cpp): blended values
  at hand-computed points, and the active brick regions its build yields,
  derived by hand from ABRs.cpp:88-201 (split order, IDs, leaf lists);
- region invariants on brick models of AMR scenes: the regions tile the union
  of brick domains, each leaf list is exactly the set of bricks overlapping
  the region, every listed brick covers the whole region, value ranges and
  finest widths from their definitions;
- tent weights: partition of unity and reproduction of affine fields inside a
  brick; the region blend equals the brute-force blend over all bricks;
- own expf / powf within a few ulp of double precision;
- .bricks round trip and the reference's file layout;
- renders: the ExaBricks ray marcher composites a constant medium to
  1 - (1-a)^L over the chord length L; direct lighting in an opaque constant
  medium shows the TF colour.
"""
import os

import numpy as np
import pytest

from oracle import oracle as ora
from owlexastitcher_amd import AmrScene, ExaBrickModel

# tools/makeTestData.cpp:8-39
TEST_BRICKS = np.array([((0, 0, 0), (1, 1, 1), 1, 0), ((2, 0, 0), (2, 2, 2), 0, 1)], ora.BRICK_DTYPE)
TEST_SCALARS = np.array([0.5, 0, 1, 0, 1, 0, 1, 0, 1], np.float32)


def test_two_brick_blend_known_values():
    pts = np.array([[1, 1, 1], [3, 1, 1], [3.25, 1, 1], [2, 1, 1], [10, 1, 1]], np.float32)
    v, w = ora.brick_blend_brute(TEST_BRICKS, TEST_SCALARS, pts)
    # (1,1,1): brick 0's only cell, weight 1; (3,1,1): brick 1 at the centre
    # of its 8 cells (brick 0's upper domain face: weight 0); (3.25,1,1):
    # x-weights 1/4, 3/4 between the 0 and 1 columns; (2,1,1): half of brick
    # 0's cell (0.5) and half of brick 1's x=0 column (0); outside: no weight
    assert np.array_equal(v[:4], np.array([0.5, 0.5, 0.75, 0.25], np.float32))
    assert np.array_equal(w[:4], np.ones(4, np.float32))
    assert w[4] == 0 and np.isnan(v[4])


def _abr_table(abrs, leaf):
    return [(tuple(a["domain"].tolist()), tuple(leaf[a["leafListBegin"]:a["leafListBegin"] + a["leafListSize"]]))
            for a in abrs]


# ABRs.cpp:88-201 worked by hand on the 2-brick data (domains [-1,3]^3 and
# [1.5,4.5]x[-.5,2.5]^2): split x=1.5; upper side first: y=2.5 (widest axis
# y, first of the equal spans y,z), ...
TWO_BRICK_ABRS = [
    ((1.5, 2.5, -1.0, 3.0, 3.0, 3.0), (0,)),
    ((1.5, -1.0, 2.5, 3.0, 2.5, 3.0), (0,)),
    ((3.0, -0.5, -0.5, 4.5, 2.5, 2.5), (1,)),
    ((1.5, -0.5, -0.5, 3.0, 2.5, 2.5), (0, 1)),
    ((1.5, -0.5, -1.0, 3.0, 2.5, -0.5), (0,)),
    ((1.5, -1.0, -1.0, 3.0, -0.5, 2.5), (0,)),
    ((-1.0, -1.0, -1.0, 1.5, 3.0, 3.0), (0,)),
]


def test_two_brick_regions_by_hand():
    abrs, leaf = ora.abrs_build(TEST_BRICKS, TEST_SCALARS)
    assert _abr_table(abrs, leaf) == TWO_BRICK_ABRS
    assert abrs["finestLevelCellWidth"].tolist() == [2, 2, 1, 1, 2, 2, 2]
    vr = abrs["valueRange"]
    assert vr[3].tolist() == [0, 1] and vr[2].tolist() == [0, 1] and vr[6].tolist() == [0.5, 0.5]
    # the library's build
    m = ExaBrickModel.from_arrays(TEST_BRICKS, TEST_SCALARS)
    b, s, a2, l2 = m.arrays()
    assert np.array_equal(a2.view(np.uint8), abrs.view(np.uint8)) and np.array_equal(l2, leaf)
    cb, vr = m.bounds()
    assert cb.tolist() == [0, 0, 0, 4, 2, 2] and vr.tolist() == [0, 1]


def _scene(kind):
    if kind == "gauss":
        return AmrScene.gaussians(16, 1, 4, 12, 4, 1)
    return AmrScene.noise(8, 2, 0.45, 0.6, 4, 2.0, 3)


def _domains(bricks):
    cw = (1 << bricks["level"].astype(np.int64)).astype(np.float32)[:, None]
    lo = bricks["lower"].astype(np.float32) - np.float32(0.5) * cw
    hi = bricks["lower"].astype(np.float32) + (bricks["size"].astype(np.float32) + np.float32(0.5)) * cw
    return lo, hi


@pytest.mark.parametrize("kind,block", [("gauss", 4), ("gauss", 16), ("noise", 8)])
def test_region_invariants(kind, block):
    m = ExaBrickModel.from_amr(_scene(kind), block)
    bricks, scalars, abrs, leaf = m.arrays()
    oa, ol = ora.abrs_build(bricks, scalars)
    assert np.array_equal(abrs.view(np.uint8), oa.view(np.uint8)) and np.array_equal(leaf, ol)
    blo, bhi = _domains(bricks)
    d = abrs["domain"].astype(np.float64)
    assert np.all(d[:, :3] < d[:, 3:])
    check_range = set(np.random.default_rng(0).choice(len(abrs), min(len(abrs), 150), replace=False).tolist())
    # exactly the bricks overlapping each region with positive volume, each
    # covering all of it, ascending
    begin, size = abrs["leafListBegin"], abrs["leafListSize"]
    assert np.array_equal(begin, np.concatenate([[0], np.cumsum(size)[:-1]]))
    for c0 in range(0, len(abrs), 1024):
        dl = abrs["domain"][c0:c0 + 1024, :3][:, None]
        dh = abrs["domain"][c0:c0 + 1024, 3:][:, None]
        ov = np.all(np.minimum(bhi[None], dh) > np.maximum(blo[None], dl), axis=2)
        r_, b_ = np.nonzero(ov)
        assert np.array_equal(b_, leaf[begin[c0]:begin[c0] + size[c0:c0 + 1024].sum()])
        assert np.array_equal(np.bincount(r_, minlength=len(dl)), size[c0:c0 + 1024])
        assert np.all(blo[b_] <= dl[r_, 0]) and np.all(bhi[b_] >= dh[r_, 0])
    for r, a in enumerate(abrs):
        lo, hi = a["domain"][:3], a["domain"][3:]
        ids = leaf[a["leafListBegin"]:a["leafListBegin"] + a["leafListSize"]]
        assert a["finestLevelCellWidth"] == float(1 << int(bricks["level"][ids].min()))
        if r not in check_range:
            continue
        # computeValueRange (ABRs.cpp:204-270) from its definition
        vals = []
        for i in ids:
            b = bricks[i]
            cw = np.float32(1 << int(b["level"]))
            g = [np.float32(b["lower"][k]) + (np.arange(b["size"][k], dtype=np.float32) + np.float32(.5)) * cw
                 for k in range(3)]
            ok = [((g[k] - cw) <= hi[k]) & ((g[k] + cw) >= lo[k]) for k in range(3)]
            sel = ok[2][:, None, None] & ok[1][None, :, None] & ok[0][None, None, :]
            cells = scalars[b["begin"]:b["begin"] + int(np.prod(b["size"]))].reshape(b["size"][::-1])
            vals.append(cells[sel])
        vals = np.concatenate(vals)
        assert a["valueRange"].tolist() == [vals.min(), vals.max()]
    # the regions tile the union of brick domains: random points lie in
    # exactly one region iff they lie in some brick domain
    rng = np.random.default_rng(5)
    root_lo, root_hi = blo.min(0), bhi.max(0)
    p = rng.uniform(root_lo - 1, root_hi + 1, size=(4000, 3)).astype(np.float32)
    inside = np.any(np.all((p[:, None] >= blo[None]) & (p[:, None] <= bhi[None]), axis=2), axis=1)
    cnt = np.sum(np.all((p[:, None] >= d[None, :, :3]) & (p[:, None] <= d[None, :, 3:]), axis=2), axis=1)
    assert np.array_equal(cnt, inside.astype(int))
    assert inside.sum() > 1000


def test_kd_nodes_describe_the_regions():
    m = ExaBrickModel.from_amr(_scene("gauss"), 8)
    _, _, abrs, _ = m.arrays()
    split, info = m.kd_nodes()
    leaves = info[(info & 3) == 3] >> 2
    # every region appears in exactly one leaf, in region order
    ids = leaves[leaves > 0] - 1
    assert np.array_equal(ids, np.arange(len(abrs)))
    # children follow their parent; a leaf's region lies on the recorded
    # sides of the planes on its path
    def walk(n, lo, hi):
        if (info[n] & 3) == 3:
            r = (info[n] >> 2) - 1
            if r >= 0:
                d = abrs["domain"][r]
                assert np.all(d[:3] >= lo) and np.all(d[3:] <= hi)
            return
        dim, low = info[n] & 3, info[n] >> 2
        assert low > n + 1
        hi2 = hi.copy()
        hi2[dim] = split[n]
        lo2 = lo.copy()
        lo2[dim] = split[n]
        walk(n + 1, lo2, hi)
        walk(low, lo, hi2)
    walk(0, np.full(3, -np.inf), np.full(3, np.inf))


def test_tent_weights_partition_of_unity_and_affine_fields():
    rng = np.random.default_rng(2)
    b = np.array([((4, 8, 0), (5, 3, 4), 1, 0)], ora.BRICK_DTYPE)
    n = 5 * 3 * 4
    z, y, x = np.meshgrid(np.arange(4), np.arange(3), np.arange(5), indexing="ij")
    cx = 4 + (x + .5) * 2
    cy = 8 + (y + .5) * 2
    cz = 0 + (z + .5) * 2
    field = (0.25 * cx - 0.5 * cy + 0.125 * cz + 3).reshape(-1).astype(np.float32)
    # points whose tent weights all fall inside the brick (between the
    # first and last cell centres)
    p = np.stack([rng.uniform(5, 13, 500), rng.uniform(9, 13, 500), rng.uniform(1, 7, 500)], 1).astype(np.float32)
    v, w = ora.brick_blend_brute(b, np.ones(n, np.float32), p)
    assert np.allclose(w, 1, atol=2e-7) and np.allclose(v, 1, atol=2e-7)
    v, w = ora.brick_blend_brute(b, field, p)
    exact = 0.25 * p[:, 0].astype(np.float64) - 0.5 * p[:, 1] + 0.125 * p[:, 2] + 3
    assert np.allclose(v, exact, atol=1e-5)


@pytest.mark.parametrize("kind", ["gauss", "noise"])
def test_region_blend_equals_brute_force(kind):
    m = ExaBrickModel.from_amr(_scene(kind), 8)
    bricks, scalars, _, _ = m.arrays()
    sc = ora.BrickScene(bricks, scalars)
    cb, _ = sc.bounds()
    rng = np.random.default_rng(9)
    p = rng.uniform(cb[:3] - 2, cb[3:] + 2, size=(20000, 3)).astype(np.float32)
    # plus points on region faces (ties between regions)
    d = sc.abrs["domain"]
    k = rng.integers(0, len(d), 3000)
    q = rng.uniform(d[k, :3], d[k, 3:]).astype(np.float32)
    q[np.arange(3000), rng.integers(0, 3, 3000)] = d[k, rng.integers(0, 6, 3000) % 3]
    p = np.concatenate([p, q])
    got = sc.locate(p)
    v, w = ora.brick_blend_brute(bricks, scalars, p)
    hit = w > 0
    assert np.array_equal(got[:, 1] == -1, hit)
    assert np.array_equal(got[hit, 2].view(np.float32), v[hit])
    assert np.all(got[~hit] == 0)


def test_culled_regions_miss():
    m = ExaBrickModel.from_amr(_scene("gauss"), 8)
    bricks, scalars, _, _ = m.arrays()
    sc = ora.BrickScene(bricks, scalars)
    cut = np.median(sc.value_ranges()[:, 1])
    rgba = np.zeros((64, 4), np.float32)
    rgba[32:, 3] = 1
    mo = ora.map_max_opacity(sc.value_ranges(), rgba, np.array([0, 2 * cut], np.float32))
    assert 0 < (mo == 0).sum() < len(mo)
    sc.set_max_opacity(mo)
    rng = np.random.default_rng(4)
    cb, _ = sc.bounds()
    p = rng.uniform(cb[:3], cb[3:], size=(5000, 3)).astype(np.float32)
    got = sc.locate(p)
    d = sc.abrs["domain"]
    inr = np.all((p[:, None] >= d[None, :, :3]) & (p[:, None] <= d[None, :, 3:]), axis=2)
    live = np.any(inr & (mo[None] > 0), axis=1)
    assert np.array_equal(got[:, 1] == -1, live)


def test_own_exp_and_pow():
    xs = np.concatenate([np.linspace(-87, 88, 20001), [-1e-30, 0.0, 1e-30, 1.0, -1.0]]).astype(np.float32)
    got = np.array([ora.expf(x) for x in xs], np.float64)
    ref = np.exp(xs.astype(np.float64))
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(got - ref) / ulp) <= 2.0
    assert ora.expf(-100.0) == 0.0 and ora.expf(100.0) == np.inf
    rng = np.random.default_rng(1)
    bases = rng.uniform(0, 1, 4000).astype(np.float32)
    exps = rng.uniform(0, 16, 4000).astype(np.float32)
    got = np.array([ora.powf(a, e) for a, e in zip(bases, exps)])
    ref = bases.astype(np.float64) ** exps.astype(np.float64)
    rel = np.abs(got - ref) / np.maximum(ref, 1e-37)
    big = ref > 1e-30
    # exp(y log x): the log's error scaled by |y log x| (< 88) plus the exp's
    assert np.max(rel[big] / np.maximum(1.0, np.abs(exps[big] * np.log(bases[big])))) < 8 * 2.0 ** -23
    assert ora.powf(0.0, 0.0) == 1.0 and ora.powf(0.0, 2.0) == 0.0 and ora.powf(1.0, 7.0) == 1.0
    assert ora.powf(0.3, 0.0) == 1.0


def test_brick_maker_covers_every_cell_once():
    amr = _scene("noise")
    cells, scal = amr.arrays()
    for block in (1, 5, 16):
        m = ExaBrickModel.from_amr(amr, block)
        bricks, scalars, _, _ = m.arrays()
        assert len(scalars) == len(cells)
        # every brick cell is an input cell of the brick's level at the
        # expected position, with its scalar
        key = {(int(c[0]), int(c[1]), int(c[2]), int(c[3])): i for i, c in enumerate(cells)}
        seen = np.zeros(len(cells), bool)
        for b in bricks:
            assert np.all(b["size"] >= 1) and np.all(b["size"] <= block)
            w = 1 << int(b["level"])
            k = int(b["begin"])
            for z in range(b["size"][2]):
                for y in range(b["size"][1]):
                    for x in range(b["size"][0]):
                        c = (int(b["lower"][0]) + x * w, int(b["lower"][1]) + y * w, int(b["lower"][2]) + z * w,
                             int(b["level"]))
                        i = key[c]
                        assert not seen[i]
                        seen[i] = True
                        assert scalars[k] == scal[i]
                        k += 1
        assert seen.all()


def test_brick_file_round_trip_and_reference_layout(tmp_path):
    amr = _scene("gauss")
    bp, sp = str(tmp_path / "m.bricks"), str(tmp_path / "m.scalars")
    amr.write_bricks(bp, sp, 8)
    a = ExaBrickModel.from_amr(amr, 8).arrays()
    b = ExaBrickModel.load(bp, sp).arrays()
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
    # tools/makeTestData.cpp:23-38: size, lower, level, cell IDs per brick;
    # raw float scalars
    raw = bytearray()
    ids = list(range(9))
    for br in TEST_BRICKS:
        raw += np.asarray(br["size"], np.int32).tobytes() + np.asarray(br["lower"], np.int32).tobytes()
        raw += np.int32(br["level"]).tobytes()
        n = int(np.prod(br["size"]))
        raw += np.asarray(ids[br["begin"]:br["begin"] + n], np.int32).tobytes()
    (tmp_path / "exa-test-data.bricks").write_bytes(bytes(raw))
    (tmp_path / "exa-test-data.scalar").write_bytes(TEST_SCALARS.tobytes())
    m = ExaBrickModel.load(str(tmp_path / "exa-test-data.bricks"), str(tmp_path / "exa-test-data.scalar"))
    bb, ss, aa, ll = m.arrays()
    assert np.array_equal(bb.view(np.uint8), TEST_BRICKS.view(np.uint8))
    assert np.array_equal(ss, TEST_SCALARS)
    assert _abr_table(aa, ll) == TWO_BRICK_ABRS
    # an ID beyond the scalar file is an error (ExaBrickModel.cpp:89-90)
    (tmp_path / "short.scalar").write_bytes(TEST_SCALARS[:5].tobytes())
    with pytest.raises(Exception):
        ExaBrickModel.load(str(tmp_path / "exa-test-data.bricks"), str(tmp_path / "short.scalar"))


def test_invalid_bricks_rejected():
    """Bricks the host refuses before any region is built: empty sizes,
    levels past 30, and extents beyond the int32 coordinate range (the
    device computes lower + size << level in int32)."""
    from owlexastitcher_amd import ExaError
    s = np.zeros(8, np.float32)
    for lower, size, level in [((0, 0, 0), (0, 2, 2), 0), ((0, 0, 0), (2, 2, 2), 31),
                               ((2**31 - 8, 0, 0), (2, 2, 2), 2), ((0, 0, 0), (2, 2, 2), 30)]:
        b = np.array([(lower, size, level, 0)], ora.BRICK_DTYPE)
        with pytest.raises(ExaError):
            ExaBrickModel.from_arrays(b, s)
    # files whose brick claims more than 2^31-1 cells fail before allocating:
    # the device reads offsets and in-brick cell indices as int32
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "huge.bricks")
        for size in [(65536, 65536, 2), (32768, 32768, 2), (2**31 - 1, 2, 1)]:
            with open(path, "wb") as f:
                f.write(np.array(list(size) + [0, 0, 0, 0], np.int32).tobytes() + bytes(64))
            with pytest.raises(ExaError):
                ExaBrickModel.load(path, path)
    # arrays whose last brick ends past 2^31-1 scalars are refused before the
    # scalars are read
    b = np.array([((0, 0, 0), (2, 2, 2), 0, 2**31 - 4)], ora.BRICK_DTYPE)
    with pytest.raises(ExaError):
        ExaBrickModel.from_arrays(b, s)
    # the largest extent that still fits: upper face at 2^31 - 4
    b = np.array([((2**31 - 8, 0, 0), (2, 2, 2), 1, 0)], ora.BRICK_DTYPE)
    assert ExaBrickModel.from_arrays(b, s).counts()["bricks"] == 1


# ---- renders -----------------------------------------------------------------
W_IMG, H_IMG = 40, 32


def _const_scene():
    """one level-1 brick of 6x5x4 cells with scalar 0.5 everywhere"""
    b = np.array([((2, 4, 6), (6, 5, 4), 1, 0)], ora.BRICK_DTYPE)
    return ora.BrickScene(b, np.full(120, 0.5, np.float32))


def _params(sc, rgba, integrator, os_=1.0, dt=1.0, offset=(-9.0, -4.5, -12.5)):
    cb, _ = sc.bounds()
    c = 0.5 * (cb[:3] + cb[3:]).astype(np.float64)
    eye = c + np.array(offset)
    fwd = (c - eye) / np.linalg.norm(c - eye)
    right = np.cross(fwd, [0.0, 1.0, 0.0])
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    h = 2 * np.tan(np.radians(16.0))
    du = right * h * (W_IMG / H_IMG) / W_IMG
    dv = -up * h / H_IMG
    d00 = fwd - 0.5 * W_IMG * du - 0.5 * H_IMG * dv
    p = ora.OraRenderParams()
    p.W, p.H, p.spp, p.accumID = W_IMG, H_IMG, 1, 0
    p.integrator, p.maxBounces, p.shadeMode, p.dt = integrator, 4, 0, dt
    for name, v in (("camOrg", eye), ("camDir00", d00), ("camDu", du), ("camDv", dv)):
        for k in range(3):
            getattr(p, name)[k] = float(v[k])
    ident = [1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0]
    for k in range(12):
        p.voxelXfm[k] = p.voxelXfmInv[k] = p.lightXfm[k] = ident[k]
    for k in range(6):
        p.worldBounds[k] = float(cb[k])
    xr = np.array([0.0, 1.0], np.float32)
    sc.set_max_opacity(ora.map_max_opacity(sc.value_ranges(), rgba, xr))
    p.tfN = len(rgba)
    p.tfDomain[0], p.tfDomain[1] = 0.0, 1.0
    p.opacityScale = os_
    p._keep = (rgba,)
    p.tfRGBA = ora.ptr(rgba).value
    return p, cb


def _rays(p, sx, sy):
    d = np.array(p.camDir00)[None] + sx[:, None] * np.array(p.camDu) + sy[:, None] * np.array(p.camDv)
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def _chords(p, cb, d):
    org = np.array(p.camOrg, np.float64)
    t0 = (cb[:3] - org) / d
    t1 = (cb[3:] - org) / d
    tn = np.maximum(np.max(np.minimum(t0, t1), axis=1), 0)
    tf = np.min(np.maximum(t0, t1), axis=1)
    return np.maximum(tf - tn, 0)


def test_brick_ray_marcher_closed_form():
    sc = _const_scene()
    a, col = 0.03, np.array([0.2, 0.5, 0.8])
    rgba = np.zeros((64, 4), np.float32)
    rgba[:, :3] = col
    rgba[:, 3] = a
    p, cb = _params(sc, rgba, 2, dt=0.5)
    acc, _ = ora.render_bricks_f32(p, sc)
    ys, xs = np.mgrid[0:H_IMG, 0:W_IMG]
    jit = np.array([ora.lcg_stream(int(x) + W_IMG * int(y), 0, 2) for y, x in zip(ys.ravel(), xs.ravel())])
    d = _rays(p, xs.ravel() + jit[:, 0].astype(np.float64), ys.ravel() + jit[:, 1].astype(np.float64))
    L = _chords(p, cb, d)
    t = ys.ravel() / H_IMG
    bg = np.stack([(1 - t) + 0.5 * t, (1 - t) + 0.7 * t, np.ones_like(t)], 1)
    alpha = 1 - (1 - a) ** L
    assert alpha.max() < 0.9 and (L > 0).sum() > 0.3 * L.size
    want = col[None] * alpha[:, None] + (1 - alpha[:, None]) * bg
    got = acc.reshape(-1, 4).astype(np.float64)
    assert np.abs(got[:, :3] - want).max() < 5e-5
    assert np.abs(got[:, 3] - 1).max() < 1e-6


def test_brick_direct_lighting_opaque_constant_medium():
    sc = _const_scene()
    col = np.array([0.3, 0.6, 0.9], np.float32)
    rgba = np.zeros((64, 4), np.float32)
    rgba[:, :3] = col
    rgba[:, 3] = 1
    p, cb = _params(sc, rgba, 1, os_=1e4)
    acc, _ = ora.render_bricks_f32(p, sc)
    ys, xs = np.mgrid[0:H_IMG, 0:W_IMG]
    jit = np.array([ora.lcg_stream(int(x) + W_IMG * int(y), 0, 2) for y, x in zip(ys.ravel(), xs.ravel())])
    d = _rays(p, xs.ravel() + jit[:, 0].astype(np.float64), ys.ravel() + jit[:, 1].astype(np.float64))
    L = _chords(p, cb, d)
    got = acc.reshape(-1, 4)
    thick = L > 0.05
    assert thick.sum() > 0.3 * L.size
    assert np.array_equal(got[thick, :3], np.broadcast_to(col, (thick.sum(), 3)))
    t = ys.ravel()[L == 0] / H_IMG
    bg = np.stack([(1 - t) + 0.5 * t, (1 - t) + 0.7 * t, np.ones_like(t)], 1)
    assert np.allclose(got[L == 0, :3], bg, atol=1e-6)
