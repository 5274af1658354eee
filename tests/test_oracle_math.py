"""CPU pinning of the oracle's numeric building blocks.

Each check is chosen so that a plausible restatement mistake (dropped term,
wrong sign or index, transposed operand) fails: closed forms (affine-field
reproduction, border halving of the texture filter), invariants (DDA
intervals tile the segment, majorant ranges contain every sample), and
independent restatements of published methods (TEA + LCG stream).
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as ora
from tests import scenes

L = ora.lib


def f32(*a):
    return np.ascontiguousarray(np.array(a, np.float32).reshape(-1))


# --------------------------------------------------------------------------
# math functions
# --------------------------------------------------------------------------
def ulp_err(got, ref):
    ref32 = np.float32(ref)
    spacing = np.spacing(np.abs(ref32)).astype(np.float64)
    return abs(float(got) - ref) / spacing


def test_logf_error_bound():
    rng = np.random.default_rng(1)
    # the device only evaluates log(1 - xi), xi = k / 2^24
    xs = np.concatenate([1.0 - np.arange(0, 1 << 24, 997, dtype=np.float64) / 2 ** 24,
                         rng.uniform(1e-6, 1.0, 20000), rng.uniform(1e-30, 1.0, 5000), rng.uniform(1.0, 1e30, 5000),
                         [1.0, 0.5, 2.0 ** -24, 0.7071067811865476, 1.4142135]])
    worst = 0.0
    for x in xs.astype(np.float32):
        got = L().ora_logf(float(x))
        ref = math.log(float(x))
        if ref == 0.0:
            assert got == 0.0
            continue
        worst = max(worst, ulp_err(got, ref))
    # the fitted polynomial measures 0.82 ulp at worst over these points
    assert worst <= 1.0, worst


def test_sincosf_error_bound():
    xs = np.linspace(0.0, 2 * np.pi, 20001).astype(np.float32)
    s = C.c_float()
    c = C.c_float()
    worst = 0.0
    for x in xs:
        L().ora_sincosf(float(x), C.byref(s), C.byref(c))
        worst = max(worst, abs(s.value - math.sin(float(x))), abs(c.value - math.cos(float(x))))
    assert worst < 3e-7, worst


# --------------------------------------------------------------------------
# RNG: TEA seeding (Zafar et al. 2010, 4 rounds) + Numerical Recipes LCG
# --------------------------------------------------------------------------
def tea_lcg_numpy(v0, v1, n):
    M = 0xFFFFFFFF
    s0 = 0
    for _ in range(4):
        s0 = (s0 + 0x9E3779B9) & M
        v0 = (v0 + ((((v1 << 4) & M) + 0xA341316C) ^ ((v1 + s0) & M) ^ ((v1 >> 5) + 0xC8013EA4))) & M
        v1 = (v1 + ((((v0 << 4) & M) + 0xAD90777D) ^ ((v0 + s0) & M) ^ ((v0 >> 5) + 0x7E95761E))) & M
    st = v0
    out = []
    for _ in range(n):
        st = (1664525 * st + 1013904223) & M
        out.append(np.float32(st & 0xFFFFFF) / np.float32(1 << 24))
    return np.array(out, np.float32)


@pytest.mark.parametrize("seed", [(0, 0), (1, 0), (123456, 7), (1920 * 1080 - 1, 255)])
def test_lcg_stream(seed):
    assert np.array_equal(ora.lcg_stream(seed[0], seed[1], 64), tea_lcg_numpy(seed[0], seed[1], 64))


def test_lcg_uniform():
    xs = np.concatenate([ora.lcg_stream(p, 3, 200) for p in range(500)])
    hist = np.histogram(xs, bins=20, range=(0, 1))[0]
    exp = len(xs) / 20
    chi2 = ((hist - exp) ** 2 / exp).sum()
    assert chi2 < 45  # 19 dof, p ~ 1e-3
    assert xs.min() >= 0.0 and xs.max() < 1.0


# --------------------------------------------------------------------------
# transfer function: texture linear filter, border 0
# --------------------------------------------------------------------------
def test_tf_texel_centres_border_and_weights():
    rng = np.random.default_rng(2)
    N = 16
    tab = rng.uniform(0, 1, size=(N, 4)).astype(np.float32)
    dom = (2.0, 6.0)
    for i in range(N):
        v = dom[0] + (i + 0.5) / N * (dom[1] - dom[0])
        assert np.allclose(ora.tf_lookup(tab, dom, v), tab[i], rtol=0, atol=1e-6)
    # ends: half of the end texel (border texel is 0)
    assert np.allclose(ora.tf_lookup(tab, dom, dom[0]), 0.5 * tab[0], atol=1e-6)
    assert np.allclose(ora.tf_lookup(tab, dom, dom[1]), 0.5 * tab[N - 1], atol=1e-6)
    # outside: 0
    assert not ora.tf_lookup(tab, dom, dom[0] - 1.0).any()
    assert not ora.tf_lookup(tab, dom, dom[1] + 1.0).any()
    # weights in steps of 1/256 between texel centres
    for k in (0, 1, 64, 128, 255):
        v = dom[0] + (3 + 0.5 + k / 256.0) / N * (dom[1] - dom[0])
        a = k / 256.0
        assert np.allclose(ora.tf_lookup(tab, dom, v), (1 - a) * tab[3] + a * tab[4], atol=2e-6)
    # empty domain
    assert not ora.tf_lookup(tab, (1.0, 1.0), 1.0).any()


# --------------------------------------------------------------------------
# element intersection
# --------------------------------------------------------------------------
REF = {
    4: np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64),
    5: np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 0.5, 1]], np.float64),
    6: np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1]], np.float64),
    8: np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]],
                np.float64),
}


def random_affine(rng):
    while True:
        A = rng.uniform(-1, 1, size=(3, 3)) + np.eye(3) * 2.0
        if abs(np.linalg.det(A)) > 0.5:
            return A * rng.uniform(0.5, 3.0), rng.uniform(-50, 50, 3)


def element_points(nv, rng, inside):
    """points of the reference element (inside) or clearly outside"""
    if nv == 4:
        b = rng.dirichlet(np.ones(4))
        if not inside:
            b = b.copy()
            b[rng.integers(4)] = -0.3
            b /= b.sum()
        return b @ REF[4]
    if nv == 8 or nv == 5:
        r, s, t = rng.uniform(0.05, 0.95, 3)
        if not inside:
            r = 1.3
        if nv == 8:
            return np.array([r, s, t])
        # pyramid: base bilinear in (r,s) shrinking to the apex with t
        base = (1 - r) * (1 - s) * REF[5][0] + r * (1 - s) * REF[5][1] + r * s * REF[5][2] + (1 - r) * s * REF[5][3]
        return (1 - t) * base + t * REF[5][4]
    r, s = rng.dirichlet(np.ones(3))[:2]
    t = rng.uniform(0.05, 0.95)
    if not inside:
        t = 1.4
    return np.array([r, s, t])


def call_elem(nv, verts4, P):
    v = np.ascontiguousarray(verts4, np.float32).reshape(-1)
    p = f32(*P)
    val = C.c_float()
    fn = {4: L().ora_intersect_tet_f32, 5: L().ora_intersect_pyr_f32, 6: L().ora_intersect_wedge_f32,
          8: L().ora_intersect_hex_f32}[nv]
    hit = fn(C.byref(val), ora.ptr(p), ora.ptr(v))
    return hit, val.value


@pytest.mark.parametrize("nv", [4, 5, 6, 8])
def test_element_reproduces_affine_field(nv):
    """Isoparametric interpolation reproduces affine fields exactly (up to
    rounding) on affinely mapped elements; Newton converges; outside misses."""
    rng = np.random.default_rng(nv)
    for _ in range(300):
        A, b = random_affine(rng)
        g, c0 = rng.uniform(-1, 1, 3), rng.uniform(-1, 1)
        X = REF[nv] @ A.T + b
        w = X @ g + c0
        verts = np.concatenate([X, w[:, None]], axis=1)
        Pref = element_points(nv, rng, True)
        P = Pref @ A.T + b
        hit, val = call_elem(nv, verts, P)
        assert hit, (nv, Pref)
        assert abs(val - (P @ g + c0)) <= 1e-4 * (1 + abs(P @ g + c0))
        Pout = element_points(nv, rng, False) @ A.T + b
        assert not call_elem(nv, verts, Pout)[0]


def test_element_dispatch_and_f64_agree():
    """The int[8] row dispatch picks the shape by its number of valid indices
    (programs/ExaStitchSampler.cu:133-161); the f64 build agrees."""
    rng = np.random.default_rng(9)
    for nv in (4, 5, 6, 8):
        A, b = random_affine(rng)
        X = REF[nv] @ A.T + b
        w = rng.uniform(0, 1, nv)
        verts = np.ascontiguousarray(np.concatenate([X, w[:, None]], axis=1), np.float32)
        idx = np.full(8, -1, np.int32)
        idx[:nv] = np.arange(nv)
        for _ in range(50):
            P = element_points(nv, rng, True) @ A.T + b
            v32 = C.c_float()
            h32 = L().ora_intersect_element_f32(C.byref(v32), ora.ptr(f32(*P)), ora.ptr(idx), ora.ptr(verts))
            v64 = C.c_double()
            h64 = L().ora_intersect_element_f64(C.byref(v64), ora.ptr(np.array(P, np.float64)), ora.ptr(idx),
                                                ora.ptr(verts))
            assert h32 == h64 == 1
            assert abs(v32.value - v64.value) < 1e-5
            # value inside the range of the vertex values (convex combination)
            assert w.min() - 1e-5 <= v32.value <= w.max() + 1e-5
    idx = np.array([0, 1, 2, 3, 4, 5, 6, -1], np.int32)  # 7 vertices: no shape
    v = C.c_float()
    assert L().ora_intersect_element_f32(C.byref(v), ora.ptr(f32(0.2, 0.2, 0.2)), ora.ptr(idx),
                                         ora.ptr(np.zeros((8, 4), np.float32))) == 0


def test_gridlet_trilinear_and_nan_holes():
    rng = np.random.default_rng(4)
    for level in (0, 1, 2):
        dims = rng.integers(1, 9, 3)
        lower = rng.integers(0, 20, 3)
        g = np.array([*lower, level, *dims, 0], np.int32)
        gcoef, c0 = rng.uniform(-1, 1, 3), rng.uniform(-1, 1)
        w = 1 << level
        ii = np.stack(np.meshgrid(*(np.arange(d + 1) for d in dims), indexing="ij"), -1)  # x,y,z
        pts = (lower + ii + 0.5) * w
        sc = (pts @ gcoef + c0).transpose(2, 1, 0).astype(np.float32).reshape(-1)  # z-major
        lo = (lower + 0.5) * w
        hi = (lower + dims + 0.5) * w
        for _ in range(200):
            P = rng.uniform(lo, hi).astype(np.float32)
            val = C.c_float()
            cid = C.c_int32()
            assert L().ora_intersect_gridlet_f32(C.byref(val), C.byref(cid), ora.ptr(P), ora.ptr(g), ora.ptr(sc))
            assert abs(val.value - (P.astype(np.float64) @ gcoef + c0)) < 1e-4
            cell = np.minimum(((P - lo) / w).astype(int), dims)
            assert cid.value == cell[2] * dims[1] * dims[0] + cell[1] * dims[0] + cell[0]
        # a NaN corner turns the cells touching it into misses
        sc2 = sc.copy()
        sc2[0] = np.nan
        val = C.c_float()
        cid = C.c_int32()
        P = (lo + 0.25 * w).astype(np.float32)
        assert not L().ora_intersect_gridlet_f32(C.byref(val), C.byref(cid), ora.ptr(P), ora.ptr(g), ora.ptr(sc2))
        Pout = (hi + 0.1).astype(np.float32)
        assert not L().ora_intersect_gridlet_f32(C.byref(val), C.byref(cid), ora.ptr(Pout), ora.ptr(g), ora.ptr(sc))


# --------------------------------------------------------------------------
# DDA and majorant grid
# --------------------------------------------------------------------------
def test_dda_intervals_tile_the_segment():
    rng = np.random.default_rng(5)
    dims = np.array([7, 5, 9], np.int32)
    bounds = np.array([1.0, -2.0, 0.5, 15.0, 8.0, 20.0], np.float32)
    lo, hi = bounds[:3].astype(np.float64), bounds[3:].astype(np.float64)
    for _ in range(300):
        org = rng.uniform(lo, hi).astype(np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        cells, t0, t1 = ora.dda_trace(org, d, 0.0, 1e10, dims, bounds)
        assert len(cells) >= 1
        assert t0[0] == 0.0
        assert np.all(t0[1:] == t1[:-1])
        # exit parameter of the box
        with np.errstate(divide="ignore"):
            ts = np.maximum((lo - org) / d, (hi - org) / d)
        assert abs(t1[-1] - ts.min()) < 1e-3 * (1 + ts.min())
        size = (hi - lo) / dims
        for c, a, b in zip(cells, t0, t1):
            if b - a < 1e-4:
                continue
            mid = org + d * (0.5 * (a + b))
            ci = np.floor((mid - lo) / size).astype(int)
            assert c == (ci[2] * dims[1] + ci[1]) * dims[0] + ci[0]


def test_majorant_ranges_contain_every_sample():
    """Grid.cu:43-246 invariant: the range of the macrocell containing a
    point contains the value sampled there (majorant conservativeness)."""
    sc = scenes.small()
    om = ora.Model(sc.verts, sc.idx, sc.gridlets, sc.gscalars)
    dims = np.array([8, 8, 8], np.int32)
    bounds = sc.cell_bounds
    ranges = om.grid_build(dims, bounds)
    rng = np.random.default_rng(6)
    pos = rng.uniform(bounds[:3], bounds[3:], size=(20000, 3)).astype(np.float32)
    out = om.locate_binned(pos)
    vals = out[:, 2].copy().view(np.float32)
    hit = out[:, 0] >= 0
    mc = np.zeros((len(pos), 3), np.int32)
    for i in np.nonzero(hit)[0][:5000]:
        L().ora_project_on_grid(ora.ptr(pos[i]), ora.ptr(dims), ora.ptr(bounds), ora.ptr(mc[i]))
        r = ranges[(mc[i, 2] * dims[1] + mc[i, 1]) * dims[0] + mc[i, 0]]
        assert r[0] - 1e-6 <= vals[i] <= r[1] + 1e-6


def test_max_opacity_index_mapping():
    """Grid.cu:706-718 on hand-computed cases: lo = int(l*(N-1)),
    hi = int(u*(N-1)) + 1, clamped; empty ranges give 0."""
    N = 5
    rgba = np.zeros((N, 4), np.float32)
    rgba[:, 3] = [0.1, 0.2, 0.3, 0.4, 0.5]
    xr = np.array([0.0, 1.0], np.float32)
    cases = {
        (0.0, 0.0): 0.2,      # lo 0, hi 1
        (0.3, 0.3): 0.3,      # int(1.2)=1, int(1.2)+1=2
        (0.3, 0.6): 0.4,      # 1..3
        (0.9, 1.0): 0.5,      # 3..4 (clamped)
        (-5.0, -4.0): 0.1,    # both clamp to 0
        (1e30, -1e30): 0.0,   # empty
    }
    ranges = np.array(list(cases.keys()), np.float32)
    got = ora.map_max_opacity(ranges, rgba, xr)
    assert np.allclose(got, list(cases.values()))


def test_locate_binned_equals_brute_force():
    sc = scenes.small()
    om = ora.Model(sc.verts, sc.idx, sc.gridlets, sc.gscalars)
    rng = np.random.default_rng(8)
    pos = rng.uniform(-1, 33, size=(4000, 3)).astype(np.float32)
    pos[:500] = (np.round(pos[:500] * 2) / 2).astype(np.float32)  # lattice ties
    b, cnt = om.locate_brute(pos)
    assert np.array_equal(b, om.locate_binned(pos))
    assert (cnt >= 1).sum() > 2500  # hull [1,31]^3 of [-1,33]^3: 69% of the points


def test_logf_coefficients_come_from_the_fit():
    """The polynomial in ora_logf (and exa_logf) is the one oracle/fit_logf.py
    derives; the device copy holds the same constants."""
    import os
    import re
    from oracle import fit_logf
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

    def constants(path, fn):
        src = open(os.path.join(root, path)).read()
        body = src[src.index(fn):]
        body = body[:body.index("\n}\n")]
        first = re.search(r"float q = (-?[0-9.e+-]+)f;", body).group(1)
        rest = re.findall(r"q = fmaf\(q, f, (-?[0-9.e+-]+)f\);", body)
        return [float(x) for x in [first] + rest][::-1]  # lowest degree first

    fitted = fit_logf.fit().astype(np.float64)
    for path, fn in (("oracle/ora_f32.c", "float ora_logf("),
                     ("owlexastitcher_amd/csrc/exa_device.h", "EXA_HD float exa_logf(")):
        c = np.array(constants(path, fn))
        assert len(c) == len(fitted) == 9
        assert np.allclose(c, fitted, rtol=1e-6, atol=0), (path, c, fitted)
