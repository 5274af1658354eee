"""ctypes binding of the CPU oracle (oracle/_build/libexaoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the product never loads it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libexaoracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class OraModel(C.Structure):
    _fields_ = [
        ("vertices", C.c_void_p), ("numVertices", C.c_int64),
        ("indices", C.c_void_p), ("numElements", C.c_int64),
        ("gridlets", C.c_void_p), ("numGridlets", C.c_int64),
        ("gridletScalars", C.c_void_p), ("numGridletScalars", C.c_int64),
        ("gridletMaxOpacity", C.c_void_p), ("elementMaxOpacity", C.c_void_p),
    ]


class OraLight(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("intensity", C.c_float), ("on", C.c_int32), ("type", C.c_int32)]


class OraRenderParams(C.Structure):
    _fields_ = [
        ("W", C.c_int32), ("H", C.c_int32), ("spp", C.c_int32), ("accumID", C.c_int32),
        ("integrator", C.c_int32), ("maxBounces", C.c_int32),
        ("camOrg", C.c_float * 3), ("camDir00", C.c_float * 3), ("camDu", C.c_float * 3), ("camDv", C.c_float * 3),
        ("worldBounds", C.c_float * 6),
        ("voxelXfm", C.c_float * 12), ("voxelXfmInv", C.c_float * 12), ("lightXfm", C.c_float * 12),
        ("gridDims", C.c_int32 * 3), ("gridBounds", C.c_float * 6),
        ("majorants", C.c_void_p), ("tfRGBA", C.c_void_p), ("tfN", C.c_int32),
        ("tfDomain", C.c_float * 2), ("opacityScale", C.c_float),
        ("clipEnabled", C.c_int32), ("clipN", C.c_float * 3), ("clipD", C.c_float),
        ("lights", OraLight * 4),
        ("debugCrosshair", C.c_int32), ("shadeMode", C.c_int32), ("dt", C.c_float),
        ("bricks", C.c_void_p),
        ("subImageActive", C.c_int32), ("subImage", C.c_float * 4),
        ("subImageSelecting", C.c_int32), ("subImageSel", C.c_float * 4),
    ]


class OraBrickScene(C.Structure):
    _fields_ = [
        ("bricks", C.c_void_p), ("numBricks", C.c_int64), ("scalars", C.c_void_p),
        ("abrs", C.c_void_p), ("numAbrs", C.c_int64), ("leafList", C.c_void_p),
        ("abrMaxOpacity", C.c_void_p),
    ]


P = C.c_void_p
_SIGS = {
    "ora_intersect_gridlet_f32": (C.c_int, [P, P, P, P, P]),
    "ora_intersect_tet_f32": (C.c_int, [P, P, P]),
    "ora_intersect_pyr_f32": (C.c_int, [P, P, P]),
    "ora_intersect_wedge_f32": (C.c_int, [P, P, P]),
    "ora_intersect_hex_f32": (C.c_int, [P, P, P]),
    "ora_intersect_element_f32": (C.c_int, [P, P, P, P]),
    "ora_intersect_gridlet_f64": (C.c_int, [P, P, P, P, P]),
    "ora_intersect_element_f64": (C.c_int, [P, P, P, P]),
    "ora_locate_brute_f32": (None, [P, P, C.c_int64, P, P]),
    "ora_bins_build": (P, [P, C.c_int]),
    "ora_bins_free": (None, [P]),
    "ora_locate_binned_f32": (None, [P, P, P, C.c_int64, P]),
    "ora_project_on_grid": (None, [P, P, P, P]),
    "ora_grid_build": (None, [P, P, P, P]),
    "ora_gridlet_value_ranges": (None, [P, P]),
    "ora_map_max_opacity": (None, [P, C.c_int64, P, C.c_int32, P, P]),
    "ora_element_max_opacity": (None, [P, P, C.c_int32, P, P]),
    "ora_dda_trace_f32": (C.c_int, [P, P, C.c_float, C.c_float, P, P, P, P, P, C.c_int]),
    "ora_tf_lookup_f32": (None, [P, C.c_int32, P, C.c_float, P]),
    "ora_orthonormal_basis_f32": (None, [P, P, P]),
    "ora_hg_sample_f32": (None, [P, C.c_uint32, C.c_uint32, C.c_int64, P]),
    "ora_lcg_init": (C.c_uint32, [C.c_uint32, C.c_uint32]),
    "ora_lcg_next": (C.c_float, [P]),
    "ora_logf": (C.c_float, [C.c_float]),
    "ora_sincosf": (None, [C.c_float, P, P]),
    "ora_render_f32": (None, [P, P, P, P, P, C.c_int32, C.c_int32]),
    "ora_render_f64": (None, [P, P, P, P, C.c_int32, C.c_int32]),
    "ora_render_queries_f32": (None, [P, P, P, P, C.c_int32, C.c_int32, P, C.c_int64, P]),
    "ora_dual_build": (P, [P, P, C.c_int64]),
    "ora_dual_count": (C.c_int64, [P, C.c_int]),
    "ora_dual_get": (None, [P, P, P, P, P]),
    "ora_dual_free": (None, [P]),
    "ora_model_bounds": (None, [P, P, P]),
    "ora_expf": (C.c_float, [C.c_float]),
    "ora_powf": (C.c_float, [C.c_float, C.c_float]),
    "ora_abrs_build": (P, [P, C.c_int64, P]),
    "ora_abrs_count": (C.c_int64, [P, C.c_int]),
    "ora_abrs_get": (None, [P, P, P]),
    "ora_abrs_free": (None, [P]),
    "ora_brick_blend_brute_f32": (None, [P, C.c_int64, P, P, C.c_int64, P, P]),
    "ora_brick_locate_f32": (None, [P, P, C.c_int64, P]),
}

# numpy layouts of the ExaBricks records (model/ABRs.h:28-75)
BRICK_DTYPE = np.dtype([("lower", "<i4", 3), ("size", "<i4", 3), ("level", "<i4"), ("begin", "<u4")])
ABR_DTYPE = np.dtype([("domain", "<f4", 6), ("valueRange", "<f4", 2), ("leafListBegin", "<i4"),
                      ("leafListSize", "<i4"), ("finestLevelCellWidth", "<f4")])

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    # data_as keeps a reference to the array for the lifetime of the pointer
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Model:
    """Keeps numpy arrays alive behind an OraModel struct."""

    def __init__(self, vertices, indices, gridlets, gridlet_scalars, gridlet_max_op=None, element_max_op=None):
        self.vertices = np.ascontiguousarray(vertices, np.float32).reshape(-1, 4)
        self.indices = np.ascontiguousarray(indices, np.int32).reshape(-1, 8)
        self.gridlets = np.ascontiguousarray(gridlets, np.int32).reshape(-1, 8)
        self.scalars = np.ascontiguousarray(gridlet_scalars, np.float32).reshape(-1)
        self.gmo = None if gridlet_max_op is None else np.ascontiguousarray(gridlet_max_op, np.float32)
        self.emo = None if element_max_op is None else np.ascontiguousarray(element_max_op, np.float32)
        self.s = OraModel(ptr(self.vertices).value, len(self.vertices), ptr(self.indices).value, len(self.indices),
                          ptr(self.gridlets).value, len(self.gridlets), ptr(self.scalars).value, len(self.scalars),
                          ptr(self.gmo).value if self.gmo is not None else None,
                          ptr(self.emo).value if self.emo is not None else None)
        self._bins = None

    def set_culling(self, gridlet_max_op, element_max_op):
        self.gmo = None if gridlet_max_op is None else np.ascontiguousarray(gridlet_max_op, np.float32)
        self.emo = None if element_max_op is None else np.ascontiguousarray(element_max_op, np.float32)
        self.s.gridletMaxOpacity = ptr(self.gmo).value if self.gmo is not None else None
        self.s.elementMaxOpacity = ptr(self.emo).value if self.emo is not None else None

    @property
    def ref(self):
        return C.byref(self.s)

    def bins(self, res=64):
        if self._bins is None:
            self._bins = lib().ora_bins_build(self.ref, res)
        return self._bins

    def locate_brute(self, pos):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        out = np.zeros((len(pos), 3), np.int32)
        cnt = np.zeros(len(pos), np.int32)
        lib().ora_locate_brute_f32(self.ref, ptr(pos), len(pos), ptr(out), ptr(cnt))
        return out, cnt

    def locate_binned(self, pos):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        out = np.zeros((len(pos), 3), np.int32)
        lib().ora_locate_binned_f32(self.bins(), self.ref, ptr(pos), len(pos), ptr(out))
        return out

    def grid_build(self, dims, bounds):
        dims = np.ascontiguousarray(dims, np.int32)
        bounds = np.ascontiguousarray(bounds, np.float32)
        out = np.zeros((int(np.prod(dims)), 2), np.float32)
        lib().ora_grid_build(self.ref, ptr(dims), ptr(bounds), ptr(out))
        return out

    def gridlet_value_ranges(self):
        out = np.zeros((len(self.gridlets), 2), np.float32)
        lib().ora_gridlet_value_ranges(self.ref, ptr(out))
        return out

    def element_max_opacity(self, rgba, xf_range):
        rgba = np.ascontiguousarray(rgba, np.float32).reshape(-1, 4)
        xr = np.ascontiguousarray(xf_range, np.float32)
        out = np.zeros(len(self.indices), np.float32)
        lib().ora_element_max_opacity(self.ref, ptr(rgba), len(rgba), ptr(xr), ptr(out))
        return out

    def bounds(self):
        cb = np.zeros(6, np.float32)
        vr = np.zeros(2, np.float32)
        lib().ora_model_bounds(self.ref, ptr(cb), ptr(vr))
        return cb, vr

    def __del__(self):
        if self._bins is not None and _lib is not None:
            _lib.ora_bins_free(self._bins)
            self._bins = None


def map_max_opacity(ranges, rgba, xf_range):
    ranges = np.ascontiguousarray(ranges, np.float32).reshape(-1, 2)
    rgba = np.ascontiguousarray(rgba, np.float32).reshape(-1, 4)
    xr = np.ascontiguousarray(xf_range, np.float32)
    out = np.zeros(len(ranges), np.float32)
    lib().ora_map_max_opacity(ptr(ranges), len(ranges), ptr(rgba), len(rgba), ptr(xr), ptr(out))
    return out


def dual_mesh(cells, scalars):
    cells = np.ascontiguousarray(cells, np.int32).reshape(-1, 4)
    scalars = np.ascontiguousarray(scalars, np.float32)
    d = lib().ora_dual_build(ptr(cells), ptr(scalars), len(cells))
    names = ("tets", "pyramids", "wedges", "hexes", "elements", "vertices", "gridlets", "gridlet_scalars",
             "cubes", "conflicts", "degenerate")
    counts = {n: int(lib().ora_dual_count(d, i)) for i, n in enumerate(names)}
    idx = np.empty((counts["elements"], 8), np.int32)
    verts = np.empty((counts["vertices"], 4), np.float32)
    gl = np.empty((counts["gridlets"], 8), np.int32)
    gs = np.empty(counts["gridlet_scalars"], np.float32)
    lib().ora_dual_get(d, ptr(idx), ptr(verts), ptr(gl), ptr(gs))
    lib().ora_dual_free(d)
    return counts, idx, verts, gl, gs


def dda_trace(org, direction, tmin, tmax, dims, bounds, max_cells=4096):
    org = np.ascontiguousarray(org, np.float32)
    direction = np.ascontiguousarray(direction, np.float32)
    dims = np.ascontiguousarray(dims, np.int32)
    bounds = np.ascontiguousarray(bounds, np.float32)
    cells = np.zeros(max_cells, np.int64)
    t0 = np.zeros(max_cells, np.float32)
    t1 = np.zeros(max_cells, np.float32)
    n = lib().ora_dda_trace_f32(ptr(org), ptr(direction), tmin, tmax, ptr(dims), ptr(bounds), ptr(cells), ptr(t0),
                                ptr(t1), max_cells)
    n = min(n, max_cells)
    return cells[:n], t0[:n], t1[:n]


def tf_lookup(rgba, domain, v):
    rgba = np.ascontiguousarray(rgba, np.float32).reshape(-1, 4)
    dom = np.ascontiguousarray(domain, np.float32)
    out = np.zeros(4, np.float32)
    lib().ora_tf_lookup_f32(ptr(rgba), len(rgba), ptr(dom), float(v), ptr(out))
    return out


def orthonormal_basis(w):
    w = np.ascontiguousarray(w, np.float32)
    u = np.zeros(3, np.float32)
    v = np.zeros(3, np.float32)
    lib().ora_orthonormal_basis_f32(ptr(w), ptr(u), ptr(v))
    return u, v


def hg_sample(wo, v0, v1, n):
    wo = np.ascontiguousarray(wo, np.float32)
    out = np.zeros((n, 3), np.float32)
    lib().ora_hg_sample_f32(ptr(wo), v0, v1, n, ptr(out))
    return out


def lcg_stream(v0, v1, n):
    st = np.array([lib().ora_lcg_init(v0, v1)], np.uint32)
    return np.array([lib().ora_lcg_next(ptr(st)) for _ in range(n)], np.float32)


def expf(x):
    return lib().ora_expf(float(x))


def powf(x, y):
    return lib().ora_powf(float(x), float(y))


def abrs_build(bricks, scalars):
    """ABRs::buildFrom (ABRs.cpp:272-352): (abrs [ABR_DTYPE], leafList)."""
    bricks = np.ascontiguousarray(bricks, BRICK_DTYPE)
    scalars = np.ascontiguousarray(scalars, np.float32)
    a = lib().ora_abrs_build(ptr(bricks), len(bricks), ptr(scalars))
    abrs = np.zeros(lib().ora_abrs_count(a, 0), ABR_DTYPE)
    leaf = np.zeros(lib().ora_abrs_count(a, 1), np.int32)
    lib().ora_abrs_get(a, ptr(abrs), ptr(leaf))
    lib().ora_abrs_free(a)
    return abrs, leaf


def brick_blend_brute(bricks, scalars, pos):
    """Blend over every brick whose closed domain contains each point:
    (value, sum of weights)."""
    bricks = np.ascontiguousarray(bricks, BRICK_DTYPE)
    scalars = np.ascontiguousarray(scalars, np.float32)
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    v = np.zeros(len(pos), np.float32)
    w = np.zeros(len(pos), np.float32)
    lib().ora_brick_blend_brute_f32(ptr(bricks), len(bricks), ptr(scalars), ptr(pos), len(pos), ptr(v), ptr(w))
    return v, w


class BrickScene:
    """Bricks, scalars, the oracle's own regions and per-region max
    opacities behind an OraBrickScene struct."""

    def __init__(self, bricks, scalars, abr_max_op=None):
        self.bricks = np.ascontiguousarray(bricks, BRICK_DTYPE)
        self.scalars = np.ascontiguousarray(scalars, np.float32)
        self.abrs, self.leaf = abrs_build(self.bricks, self.scalars)
        self.s = OraBrickScene(ptr(self.bricks).value, len(self.bricks), ptr(self.scalars).value,
                               ptr(self.abrs).value, len(self.abrs), ptr(self.leaf).value, None)
        self.set_max_opacity(abr_max_op)

    def set_max_opacity(self, m):
        self.mo = None if m is None else np.ascontiguousarray(m, np.float32)
        self.s.abrMaxOpacity = ptr(self.mo).value if self.mo is not None else None

    def value_ranges(self):
        return np.ascontiguousarray(self.abrs["valueRange"], np.float32)

    @property
    def ref(self):
        return C.byref(self.s)

    def locate(self, pos):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        out = np.zeros((len(pos), 3), np.int32)
        lib().ora_brick_locate_f32(self.ref, ptr(pos), len(pos), ptr(out))
        return out

    def bounds(self):
        """cellBounds = union of brick bounds, valueRange = union of region
        ranges (ExaBrickModel.cpp:177-187)"""
        b = self.bricks
        w = (1 << b["level"].astype(np.int64))[:, None]
        lo = b["lower"].astype(np.int64)
        hi = lo + b["size"].astype(np.int64) * w
        cb = np.concatenate([lo.min(0), hi.max(0)]).astype(np.float32)
        vr = np.array([self.abrs["valueRange"][:, 0].min(), self.abrs["valueRange"][:, 1].max()], np.float32)
        return cb, vr


def render_bricks_f32(params, scene, accum=None, threads=8):
    """One frame of the ExaBricks renderer (params.bricks is set here)."""
    from concurrent.futures import ThreadPoolExecutor
    W, H = params.W, params.H
    if accum is None:
        accum = np.zeros((H, W, 4), np.float32)
    fb = np.zeros((H, W), np.uint32)
    params.bricks = C.cast(C.pointer(scene.s), C.c_void_p).value
    rows = max(1, H // (threads * 4))

    def work(y0):
        lib().ora_render_f32(C.byref(params), None, None, ptr(accum), ptr(fb), y0, min(H, y0 + rows))

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(work, range(0, H, rows)))
    return accum, fb


def render_params_from_state(st, majorants, tf_rgba):
    """Fill oracle render params from an exa_render_state snapshot."""
    p = OraRenderParams()
    for f in ("W", "H", "spp", "accumID", "integrator", "maxBounces", "opacityScale", "clipEnabled", "clipD",
              "tfN", "debugCrosshair", "shadeMode", "dt", "subImageActive", "subImageSelecting"):
        setattr(p, f, getattr(st, f))
    for f in ("camOrg", "camDir00", "camDu", "camDv", "worldBounds", "voxelXfm", "voxelXfmInv", "lightXfm",
              "gridDims", "gridBounds", "tfDomain", "clipN", "subImage", "subImageSel"):
        C.memmove(getattr(p, f), getattr(st, f), C.sizeof(getattr(p, f)))
    for i in range(4):
        for k in range(3):
            p.lights[i].pos[k] = st.lightPos[i][k]
        p.lights[i].intensity = st.lightIntensity[i]
        p.lights[i].on = st.lightOn[i]
        p.lights[i].type = st.lightType[i]
    p._keep = (majorants, tf_rgba)
    p.majorants = ptr(majorants).value
    p.tfRGBA = ptr(tf_rgba).value
    return p


def render_f32(params, model, accum=None, threads=8):
    """Render one frame on host threads (ctypes releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    W, H = params.W, params.H
    if accum is None:
        accum = np.zeros((H, W, 4), np.float32)
    fb = np.zeros((H, W), np.uint32)
    bins = model.bins()
    rows = max(1, H // (threads * 4))

    def work(y0):
        lib().ora_render_f32(C.byref(params), model.ref, bins, ptr(accum), ptr(fb), y0, min(H, y0 + rows))

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(work, range(0, H, rows)))
    return accum, fb


def render_queries_f32(params, model, cap, threads=8):
    """One frame as render_f32, returning the positions of its point-location
    queries (n x 3 float32; order across threads unspecified)."""
    from concurrent.futures import ThreadPoolExecutor
    W, H = params.W, params.H
    accum = np.zeros((H, W, 4), np.float32)
    out = np.zeros((cap, 3), np.float32)
    count = np.zeros(1, np.int64)
    bins = model.bins()

    def work(y):
        lib().ora_render_queries_f32(C.byref(params), model.ref, bins, ptr(accum), y, y + 1, ptr(out), cap,
                                     ptr(count))

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(work, range(H)))
    n = int(count[0])
    if n > cap:
        raise ValueError(f"frame made {n} queries, more than cap={cap}")
    return out[:n]


def render_f64(params, model, accum=None, threads=8):
    from concurrent.futures import ThreadPoolExecutor
    W, H = params.W, params.H
    if accum is None:
        accum = np.zeros((H, W, 4), np.float64)
    bins = model.bins()
    rows = max(1, H // (threads * 4))

    def work(y0):
        lib().ora_render_f64(C.byref(params), model.ref, bins, ptr(accum), y0, min(H, y0 + rows))

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(work, range(0, H, rows)))
    return accum
