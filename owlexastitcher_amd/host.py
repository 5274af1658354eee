"""Python host mirror of the reference's exaStitch interfaces over the C ABI.

Class and method names follow the reference so scripts read like its own
host code:
  ExaStitchModel   model/ExaStitchModel.h:24-57 + model/Model.h:29-65
  ExaStitchSampler sampler/Sampler.h:28-69, sampler/ExaStitchSampler.h:37-106
  ExaBrickModel    model/ExaBrickModel.h:27-56 + model/ABRs.h:28-100
  ExaBrickSampler  sampler/ExaBrickSampler.h:30-241 (ABR mode)
  OWLRenderer      OWLRenderer.h:29-175
Device buffers are torch tensors on the current HIP device (plumbing only);
every computation runs in libexastitch.so.
"""
import ctypes as C
import os

import numpy as np

from ._capi import (AmrCell, Gridlet, RenderState, StitchModelDesc, check, lib)

PATH_TRACER = 0
DIRECT_LIGHTING = 1
RAY_MARCHER = 2
ACCEL_TABLES = 0
ACCEL_BVH = 1
RENDER_MEGAKERNEL = 0
RENDER_PERSISTENT = 1
ELEMENT_IN_WAVE = 0
ELEMENT_WORKGROUP_QUEUE = 1
ELEMENT_QUEUE_ROUNDS = 2


def _ptr(a):
    # data_as keeps a reference to the array for the lifetime of the pointer
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _path(p):
    if p is None:
        return None
    return os.fsencode(p)


def _f32(x, n=None):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} floats, got {a.size}")
    return a


class AmrScene:
    """Leaf cells (position, level) + per-cell scalars of an AMR hierarchy."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def gaussians(cls, base_res=64, coarse_level=1, fine_lo=32, fine_hi=96, num_gaussians=8, seed=1):
        h = C.c_void_p()
        check(lib().exa_amr_create_gaussians(base_res, coarse_level, fine_lo, fine_hi, num_gaussians, seed,
                                             C.byref(h)))
        return cls(h)

    @classmethod
    def noise(cls, base_res=128, base_level=2, thresh1=0.55, thresh2=0.7, octaves=6, contrast=2.2, seed=1):
        h = C.c_void_p()
        check(lib().exa_amr_create_noise(base_res, base_level, thresh1, thresh2, octaves, contrast, seed,
                                         C.byref(h)))
        return cls(h)

    @classmethod
    def shock(cls, base_res=256, base_level=4, radius=1200.0, thickness=40.0, theta_core=0.74, theta_wake=0.43,
              seed=1):
        """Expanding spherical shock shell with turbulence in its wake (5 levels
        at the defaults, over [0, base_res * 2^base_level]^3)."""
        h = C.c_void_p()
        check(lib().exa_amr_create_shock(int(base_res), int(base_level), float(radius), float(thickness),
                                         float(theta_core), float(theta_wake), int(seed), C.byref(h)))
        return cls(h)

    @staticmethod
    def shock_counts(base_res=256, base_level=4, radius=1200.0, thickness=40.0, theta_core=0.74, theta_wake=0.43,
                     seed=1):
        """Leaf cells per level of AmrScene.shock(...) without building it."""
        c = np.zeros(16, np.int64)
        check(lib().exa_amr_count_shock(int(base_res), int(base_level), float(radius), float(thickness),
                                        float(theta_core), float(theta_wake), int(seed), _ptr(c)))
        return c[:base_level + 1]

    @classmethod
    def from_cells(cls, cells, scalars):
        cells = np.ascontiguousarray(cells, dtype=np.int32).reshape(-1, 4)
        scalars = _f32(scalars)
        h = C.c_void_p()
        check(lib().exa_amr_create_from_cells(_ptr(cells), _ptr(scalars), len(cells), C.byref(h)))
        return cls(h)

    @classmethod
    def load(cls, cells_path, scalars_path):
        """Leaf cells from a .cells file (16-byte {x, y, z, level} records) and
        one float32 scalar per cell from a .scalars file."""
        h = C.c_void_p()
        check(lib().exa_amr_load(_path(cells_path), _path(scalars_path), C.byref(h)))
        return cls(h)

    def save(self, cells_path=None, scalars_path=None):
        check(lib().exa_amr_save(self.h, _path(cells_path), _path(scalars_path)))

    def write_bricks(self, bricks_path, scalars_path=None, block_cells=16):
        """.bricks file of the brick maker's bricks (+ the .scalars file)"""
        check(lib().exa_amr_write_bricks(self.h, int(block_cells), _path(bricks_path), _path(scalars_path)))

    def write_dual_mesh(self, element_mesh_path, cubes_prefix=None, grids_path=None):
        """amrMakeDualMesh + amrMakeGrids to files: the element mesh, the
        per-level <cubes_prefix>_<L>.cubes files and the .grids file (None
        skips a file).  Returns the dual mesh counts."""
        c = np.zeros(11, np.int64)
        check(lib().exa_amr_write_dual_mesh(self.h, _path(element_mesh_path), _path(cubes_prefix),
                                            _path(grids_path), _ptr(c)))
        return dict(zip(DualMesh.COUNT_NAMES, (int(v) for v in c)))

    def __len__(self):
        return int(lib().exa_amr_num_cells(self.h))

    def arrays(self):
        n = len(self)
        cells = np.empty((n, 4), np.int32)
        scalars = np.empty(n, np.float32)
        check(lib().exa_amr_get(self.h, _ptr(cells), _ptr(scalars)))
        return cells, scalars

    def params(self):
        n = lib().exa_amr_get_params(self.h, None, 0)
        out = np.empty(max(n, 1), np.float32)
        lib().exa_amr_get_params(self.h, _ptr(out), n)
        return out[:n]

    def __del__(self):
        if getattr(self, "h", None):
            lib().exa_amr_destroy(self.h)
            self.h = None


class DualMesh:
    """Output of the dual-mesh generator (tools/amrMakeDualMesh)."""

    COUNT_NAMES = ("tets", "pyramids", "wedges", "hexes", "elements", "vertices", "gridlets",
                   "gridlet_scalars", "cubes", "conflicts", "degenerate")

    def __init__(self, amr):
        self.h = C.c_void_p()
        check(lib().exa_dual_mesh_build(amr.h, C.byref(self.h)))

    def counts(self):
        c = np.zeros(11, np.int64)
        check(lib().exa_dual_mesh_counts(self.h, _ptr(c)))
        return dict(zip(self.COUNT_NAMES, (int(v) for v in c)))

    def arrays(self):
        c = self.counts()
        idx = np.empty((c["elements"], 8), np.int32)
        verts = np.empty((c["vertices"], 4), np.float32)
        gl = (Gridlet * max(c["gridlets"], 1))()
        gs = np.empty(c["gridlet_scalars"], np.float32)
        check(lib().exa_dual_mesh_get(self.h, _ptr(idx), _ptr(verts), gl, _ptr(gs)))
        gla = np.ctypeslib.as_array(gl).view(np.int32).reshape(-1, 8)[: c["gridlets"]].copy()
        return idx, verts, gla, gs

    def __del__(self):
        if getattr(self, "h", None):
            lib().exa_dual_mesh_destroy(self.h)
            self.h = None


class ExaStitchModel:
    """ExaStitchModel: element vertices/indices + gridlets (copied in)."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def from_dual_mesh(cls, dm):
        h = C.c_void_p()
        check(lib().exa_model_create_from_dual_mesh(dm.h, C.byref(h)))
        return cls(h)

    @classmethod
    def load(cls, umesh_file_name, grids_file_name, scalar_file_name):
        """ExaStitchModel::load(umesh, grids, scalars) (ExaStitchModel.cpp:24-257);
        the element mesh is this library's EXAELEM1 file; any name may be
        None or empty."""
        h = C.c_void_p()
        check(lib().exa_model_load(_path(umesh_file_name), _path(grids_file_name), _path(scalar_file_name),
                                   C.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, vertices, indices, gridlets, gridlet_scalars):
        v = _f32(vertices).reshape(-1)
        i = np.ascontiguousarray(indices, dtype=np.int32).reshape(-1)
        g = np.ascontiguousarray(gridlets, dtype=np.int32).reshape(-1)
        s = _f32(gridlet_scalars).reshape(-1)
        d = StitchModelDesc(_ptr(v).value, v.size // 4, _ptr(i).value, i.size // 8, _ptr(g).value, g.size // 8,
                            _ptr(s).value, s.size)
        h = C.c_void_p()
        check(lib().exa_model_create(C.byref(d), C.byref(h)))
        return cls(h)

    def bounds(self):
        cb = np.zeros(6, np.float32)
        vr = np.zeros(2, np.float32)
        check(lib().exa_model_get_bounds(self.h, _ptr(cb), _ptr(vr)))
        return cb, vr

    def counts(self):
        c = np.zeros(4, np.int64)
        check(lib().exa_model_get_counts(self.h, _ptr(c)))
        return dict(zip(("vertices", "elements", "gridlets", "gridlet_scalars"), (int(x) for x in c)))

    def arrays(self):
        """(vertices float4 [V,4], indices int[8] [E,8], gridlets int32 [G,8],
        gridlet scalars [S]) as held by the model."""
        c = self.counts()
        v = np.empty((c["vertices"], 4), np.float32)
        i = np.empty((c["elements"], 8), np.int32)
        g = np.empty((c["gridlets"], 8), np.int32)
        s = np.empty(c["gridlet_scalars"], np.float32)
        check(lib().exa_model_get(self.h, _ptr(v), _ptr(i), _ptr(g), _ptr(s)))
        return v, i, g, s

    def setNumGridCells(self, num_mcs):
        d = np.ascontiguousarray(num_mcs, dtype=np.int32)
        check(lib().exa_model_set_num_grid_cells(self.h, _ptr(d)))

    def setVoxelSpaceTransform(self, remap_from, remap_to):
        check(lib().exa_model_set_voxel_space_transform(self.h, _ptr(_f32(remap_from, 6)), _ptr(_f32(remap_to, 6))))

    def __del__(self):
        if getattr(self, "h", None):
            lib().exa_model_destroy(self.h)
            self.h = None


def _stream(stream):
    if stream is None:
        return None
    return C.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))


class ExaStitchSampler:
    """ExaStitchSampler: build / computeMaxOpacities / device sample()."""

    def __init__(self, handle=None, owner=None):
        self._owned = handle is None
        if handle is None:
            handle = C.c_void_p()
            check(lib().exa_sampler_create(C.byref(handle)))
        self.h = handle
        self._owner = owner

    def build(self, model, stream=None):
        check(lib().exa_sampler_build(self.h, model.h, _stream(stream)))
        return True

    def setAccel(self, accel):
        check(lib().exa_sampler_set_accel(self.h, int(accel)))

    @property
    def accel(self):
        return int(lib().exa_sampler_get_accel(self.h))

    def computeMaxOpacities(self, rgba, xf_range, stream=None):
        cm = _f32(rgba).reshape(-1, 4)
        check(lib().exa_sampler_compute_max_opacities(self.h, _ptr(cm), len(cm), float(xf_range[0]),
                                                      float(xf_range[1]), _stream(stream)))

    def locate(self, d_pos, d_out, stream=None):
        """d_pos: float32 tensor (n,3) on the GPU; d_out: int32 tensor (n,3)
        viewed as Sample records."""
        n = d_pos.shape[0]
        check(lib().exa_sampler_locate(self.h, C.c_void_p(d_pos.data_ptr()), n, C.c_void_p(d_out.data_ptr()),
                                       _stream(stream)))

    def locate_counted(self, d_pos, d_out, stream=None):
        n = d_pos.shape[0]
        c = np.zeros(6, np.uint64)
        check(lib().exa_sampler_locate_counted(self.h, C.c_void_p(d_pos.data_ptr()), n,
                                               C.c_void_p(d_out.data_ptr()), _ptr(c), _stream(stream)))
        return dict(zip(("gridlet", "tet", "pyr", "wedge", "hex", "miss"), (int(x) for x in c)))

    def testSample(self, d_pos, d_prim, d_out, stream=None):
        n = d_pos.shape[0]
        check(lib().exa_sampler_test_sample(self.h, C.c_void_p(d_pos.data_ptr()), C.c_void_p(d_prim.data_ptr()),
                                            n, C.c_void_p(d_out.data_ptr()), _stream(stream)))

    def grid_info(self):
        dims = np.zeros(3, np.int32)
        b = np.zeros(6, np.float32)
        check(lib().exa_sampler_get_grid_info(self.h, _ptr(dims), _ptr(b)))
        return dims, b

    def value_ranges(self):
        dims, _ = self.grid_info()
        out = np.empty((int(np.prod(dims)), 2), np.float32)
        check(lib().exa_sampler_get_value_ranges(self.h, _ptr(out)))
        return out

    def max_opacities(self):
        dims, _ = self.grid_info()
        out = np.empty(int(np.prod(dims)), np.float32)
        check(lib().exa_sampler_get_max_opacities(self.h, _ptr(out)))
        return out

    def gridlet_value_ranges(self, num_gridlets):
        out = np.empty((num_gridlets, 2), np.float32)
        check(lib().exa_sampler_get_gridlet_value_ranges(self.h, _ptr(out)))
        return out

    def prim_max_opacities(self, num_gridlets, num_elements):
        g = np.empty(max(num_gridlets, 1), np.float32)
        e = np.empty(max(num_elements, 1), np.float32)
        check(lib().exa_sampler_get_prim_max_opacities(self.h, _ptr(g), _ptr(e)))
        return g[:num_gridlets], e[:num_elements]

    def info(self):
        i = np.zeros(8, np.int64)
        check(lib().exa_sampler_get_info(self.h, _ptr(i)))
        return dict(zip(("gridlet_nodes", "gridlet_depth", "gridlet_refs", "element_nodes", "element_depth",
                         "element_refs", "model_bytes", "bvh_bytes"), (int(x) for x in i)))

    def __del__(self):
        if self._owned and getattr(self, "h", None):
            lib().exa_sampler_destroy(self.h)
            self.h = None


BRICK_DTYPE = np.dtype([("lower", "<i4", 3), ("size", "<i4", 3), ("level", "<i4"), ("begin", "<u4")])
ABR_DTYPE = np.dtype([("domain", "<f4", 6), ("valueRange", "<f4", 2), ("leafListBegin", "<i4"),
                      ("leafListSize", "<i4"), ("finestLevelCellWidth", "<f4")])


class ExaBrickModel:
    """ExaBrickModel: bricks of AMR cells + their active brick regions."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def load(cls, brick_file_name, scalar_file_name):
        """ExaBrickModel::load(brickFile, scalarFile, "") (ExaBrickModel.cpp:27-117)"""
        h = C.c_void_p()
        check(lib().exa_brick_model_load(_path(brick_file_name), _path(scalar_file_name), C.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, bricks, scalars):
        """ExaBrickModel::load(bricks, scalars, numBricks) (ExaBrickModel.cpp:119-145)"""
        b = np.ascontiguousarray(bricks, BRICK_DTYPE)
        s = _f32(scalars).reshape(-1)
        h = C.c_void_p()
        check(lib().exa_brick_model_create(_ptr(b), _ptr(s), len(b), C.byref(h)))
        return cls(h)

    @classmethod
    def from_amr(cls, amr, block_cells=16):
        """bricks of an AMR scene's leaf cells (the brick maker)"""
        h = C.c_void_p()
        check(lib().exa_brick_model_from_amr(amr.h, int(block_cells), C.byref(h)))
        return cls(h)

    def counts(self):
        c = np.zeros(6, np.int64)
        check(lib().exa_brick_model_get_counts(self.h, _ptr(c)))
        return dict(zip(("bricks", "cells", "abrs", "leaf_list", "kd_nodes", "kd_depth"), (int(x) for x in c)))

    def arrays(self):
        """(bricks [BRICK_DTYPE], scalars, abrs [ABR_DTYPE], leafList)"""
        c = self.counts()
        b = np.empty(c["bricks"], BRICK_DTYPE)
        s = np.empty(c["cells"], np.float32)
        a = np.empty(c["abrs"], ABR_DTYPE)
        ll = np.empty(c["leaf_list"], np.int32)
        check(lib().exa_brick_model_get(self.h, _ptr(b), _ptr(s), _ptr(a), _ptr(ll)))
        return b, s, a, ll

    def kd_nodes(self):
        """(split float32, info int32) per kd node"""
        n = np.empty((self.counts()["kd_nodes"], 2), np.int32)
        check(lib().exa_brick_model_get_kd_nodes(self.h, _ptr(n)))
        return n[:, 0].view(np.float32).copy(), n[:, 1].copy()

    def bounds(self):
        cb = np.zeros(6, np.float32)
        vr = np.zeros(2, np.float32)
        check(lib().exa_brick_model_get_bounds(self.h, _ptr(cb), _ptr(vr)))
        return cb, vr

    def __del__(self):
        if getattr(self, "h", None):
            lib().exa_brick_model_destroy(self.h)
            self.h = None


class ExaBrickSampler:
    """ExaBrickSampler in ABR mode: computeMaxOpacities / device sample()."""

    def __init__(self, model=None, stream=None, handle=None, owner=None):
        self._owned = handle is None
        if handle is None:
            handle = C.c_void_p()
            check(lib().exa_brick_sampler_create(model.h, _stream(stream), C.byref(handle)))
        self.h = handle
        self._owner = owner
        self._n = None if model is None else model.counts()["abrs"]

    def computeMaxOpacities(self, rgba, xf_range, stream=None):
        cm = _f32(rgba).reshape(-1, 4)
        check(lib().exa_brick_sampler_compute_max_opacities(self.h, _ptr(cm), len(cm), float(xf_range[0]),
                                                            float(xf_range[1]), _stream(stream)))

    def max_opacities(self, num_abrs=None):
        n = num_abrs if num_abrs is not None else self._n
        out = np.empty(n, np.float32)
        check(lib().exa_brick_sampler_get_max_opacities(self.h, _ptr(out)))
        return out

    def locate(self, d_pos, d_out, stream=None):
        """d_pos: float32 tensor (n,3) on the GPU; d_out: int32 tensor (n,3)
        viewed as Sample records: {0, -1, value} or {0, 0, 0}."""
        check(lib().exa_brick_sampler_locate(self.h, C.c_void_p(d_pos.data_ptr()), d_pos.shape[0],
                                             C.c_void_p(d_out.data_ptr()), _stream(stream)))

    def __del__(self):
        if self._owned and getattr(self, "h", None):
            lib().exa_brick_sampler_destroy(self.h)
            self.h = None


class OWLRenderer:
    """OWLRenderer: model -> sampler -> Woodcock path-traced frames.  An
    ExaBrickModel gives the ExaBricks renderer (OWLRenderer.cpp:269-291)."""

    def __init__(self, model, num_mcs=(128, 128, 128), remap_from=None, remap_to=None):
        self.h = C.c_void_p()
        mcs = np.ascontiguousarray(num_mcs, dtype=np.int32)
        rf = _f32(remap_from, 6) if remap_from is not None else None
        rt = _f32(remap_to, 6) if remap_to is not None else None
        self.bricks = isinstance(model, ExaBrickModel)
        if self.bricks:
            check(lib().exa_renderer_create_bricks(model.h, _ptr(mcs), _ptr(rf), _ptr(rt), C.byref(self.h)))
            self.sampler = ExaBrickSampler(handle=lib().exa_renderer_get_brick_sampler(self.h), owner=self)
            self.sampler._n = model.counts()["abrs"]
        else:
            check(lib().exa_renderer_create(model.h, _ptr(mcs), _ptr(rf), _ptr(rt), C.byref(self.h)))
            self.sampler = ExaStitchSampler(lib().exa_renderer_get_sampler(self.h), owner=self)

    # --- OWLRenderer.h:88-149 ---
    def setCamera(self, org, dir_00, dir_du, dir_dv):
        check(lib().exa_renderer_set_camera(self.h, _ptr(_f32(org, 3)), _ptr(_f32(dir_00, 3)),
                                            _ptr(_f32(dir_du, 3)), _ptr(_f32(dir_dv, 3))))

    def resize(self, w, h):
        check(lib().exa_renderer_resize(self.h, int(w), int(h)))
        self.size = (int(w), int(h))

    def render(self, d_fb=None, stream=None):
        ptr = C.c_void_p(d_fb.data_ptr()) if d_fb is not None else None
        check(lib().exa_renderer_render(self.h, ptr, _stream(stream)))

    def render_counted(self, d_fb=None, stream=None):
        ptr = C.c_void_p(d_fb.data_ptr()) if d_fb is not None else None
        c = np.zeros(11, np.uint64)
        check(lib().exa_renderer_render_counted(self.h, ptr, _ptr(c), _stream(stream)))
        if self.bricks:
            return {"samples": int(c[0]), "bricks": int(c[1]), "kd_leaves": int(c[2]), "regions": int(c[7]),
                    "pixels": int(c[8]), "tf": int(c[9])}
        names = ("gridlet", "tet", "pyr", "wedge", "hex", "miss", "cached", "dda_cells", "pixels", "tf", "neural")
        return dict(zip(names, (int(x) for x in c)))

    def accum(self):
        w, h = self.size
        out = np.empty((h, w, 4), np.float32)
        check(lib().exa_renderer_get_accum(self.h, _ptr(out)))
        return out

    def setType(self, t):
        check(lib().exa_renderer_set_type(self.h, int(t)))

    def setColorMap(self, rgba):
        cm = _f32(rgba).reshape(-1, 4)
        check(lib().exa_renderer_set_color_map(self.h, _ptr(cm), len(cm)))

    def setRange(self, lo, hi):
        check(lib().exa_renderer_set_range(self.h, float(lo), float(hi)))

    def setRelDomain(self, lo, hi):
        check(lib().exa_renderer_set_rel_domain(self.h, float(lo), float(hi)))

    def setOpacityScale(self, s):
        check(lib().exa_renderer_set_opacity_scale(self.h, float(s)))

    def setClipPlane(self, i, enabled, N, d):
        check(lib().exa_renderer_set_clip_plane(self.h, int(i), int(bool(enabled)), _ptr(_f32(N, 3)), float(d)))

    def setShadeMode(self, sm):
        check(lib().exa_renderer_set_shade_mode(self.h, int(sm)))

    @staticmethod
    def shadeModes():
        return {0: "default", 1: "gridlets", 2: "teaser"}

    def setDt(self, dt):
        """Ray-marcher step (the reference fixes it at 1)."""
        check(lib().exa_renderer_set_dt(self.h, float(dt)))

    def setHeatMap(self, enabled, scale=1.0):
        check(lib().exa_renderer_set_heat_map(self.h, int(bool(enabled)), float(scale)))

    def setLightSource(self, i, pos, intensity, on):
        check(lib().exa_renderer_set_light_source(self.h, int(i), _ptr(_f32(pos, 3)), float(intensity),
                                                  int(bool(on))))

    def setDirectionalLight(self, i, direction, intensity, on):
        check(lib().exa_renderer_set_directional_light(self.h, int(i), _ptr(_f32(direction, 3)), float(intensity),
                                                       int(bool(on))))

    def setLightSpaceTransform(self, xfm12):
        check(lib().exa_renderer_set_light_space_transform(self.h, _ptr(_f32(xfm12, 12))))

    def setSubImage(self, box, active):
        check(lib().exa_renderer_set_sub_image(self.h, _ptr(_f32(box, 4)), int(bool(active))))

    def setSubImageSelection(self, box, active):
        check(lib().exa_renderer_set_sub_image_selection(self.h, _ptr(_f32(box, 4)), int(bool(active))))

    def resetAccum(self):
        check(lib().exa_renderer_reset_accum(self.h))

    def setSpp(self, spp):
        check(lib().exa_renderer_set_spp(self.h, int(spp)))

    def setMaxBounces(self, n):
        check(lib().exa_renderer_set_max_bounces(self.h, int(n)))

    def setAccel(self, accel):
        check(lib().exa_renderer_set_accel(self.h, int(accel)))

    def setNeural(self, field):
        """Use a validated NeuralField for the element stage of its macrocells
        (None: exact sampler everywhere)."""
        check(lib().exa_renderer_set_neural(self.h, field.h if field is not None else None))
        self.neural = field  # keep it alive while in use

    def setScheme(self, scheme):
        """RENDER_PERSISTENT (default) or RENDER_MEGAKERNEL; same images."""
        check(lib().exa_renderer_set_scheme(self.h, int(scheme)))

    @property
    def scheme(self):
        rc = lib().exa_renderer_get_scheme(self.h)
        if rc < 0:
            check(rc)
        return rc

    def setPersistentTuning(self, ctrlSteps=4, elemBatch=16, rareBatch=24):
        """Scheduling knobs of RENDER_PERSISTENT; images do not depend on them."""
        check(lib().exa_renderer_set_persistent_tuning(self.h, int(ctrlSteps), int(elemBatch), int(rareBatch)))

    def setLaunchTuning(self, element_stage=ELEMENT_IN_WAVE, queue_rounds=1, queue_depth=0, block_size=256):
        """Launch shape of RENDER_PERSISTENT: where the element stage runs
        (ELEMENT_IN_WAVE, ELEMENT_WORKGROUP_QUEUE, ELEMENT_QUEUE_ROUNDS), the
        queue rounds per frame and entries per queue of the last (0: one per
        pixel), threads per workgroup; images do not depend on it."""
        check(lib().exa_renderer_set_launch_tuning(self.h, int(element_stage), int(queue_rounds), int(queue_depth),
                                                   int(block_size)))

    @property
    def tuning(self):
        """(ctrlSteps, elemBatch, rareBatch, elementStage, queueRounds, queueDepth, blockSize)"""
        t = np.zeros(7, np.int32)
        check(lib().exa_renderer_get_tuning(self.h, _ptr(t)))
        return tuple(int(x) for x in t)

    def setDebugCrosshair(self, on):
        check(lib().exa_renderer_set_debug_crosshair(self.h, int(bool(on))))

    @property
    def modelBounds(self):
        b = np.zeros(6, np.float32)
        check(lib().exa_renderer_get_bounds(self.h, _ptr(b)))
        return b

    @property
    def valueRange(self):
        v = np.zeros(2, np.float32)
        check(lib().exa_renderer_get_value_range(self.h, _ptr(v)))
        return v

    @property
    def accumID(self):
        return int(lib().exa_renderer_get_accum_id(self.h))

    @accumID.setter
    def accumID(self, v):
        check(lib().exa_renderer_set_accum_id(self.h, int(v)))

    def state(self):
        s = RenderState()
        check(lib().exa_renderer_get_state(self.h, C.byref(s)))
        return s

    def __del__(self):
        if getattr(self, "h", None):
            lib().exa_renderer_destroy(self.h)
            self.h = None


def camera_look_at(frm, at, up, fovy, W, H):
    org, d00, du, dv = (np.zeros(3, np.float32) for _ in range(4))
    check(lib().exa_camera_look_at(_ptr(_f32(frm, 3)), _ptr(_f32(at, 3)), _ptr(_f32(up, 3)), float(fovy), int(W),
                                   int(H), _ptr(org), _ptr(d00), _ptr(du), _ptr(dv)))
    return org, d00, du, dv


def tf_rasterize(points, N=128):
    pts = _f32(points).reshape(-1, 5)
    out = np.empty((N, 4), np.float32)
    check(lib().exa_tf_rasterize(_ptr(pts), len(pts), int(N), _ptr(out)))
    return out


def make_grids(cubes_paths, grids_path):
    """amrMakeGrids (makeGrids.cpp:262-343): bricks of each _<L>.cubes file,
    appended to one .grids file in argument order.  Returns the number of
    inconsistent brick-vertex writes (0 for a valid dual mesh)."""
    arr = (C.c_char_p * max(len(cubes_paths), 1))(*[_path(p) for p in cubes_paths])
    bad = C.c_int64()
    check(lib().exa_make_grids(arr, len(cubes_paths), _path(grids_path), C.byref(bad)))
    return int(bad.value)


def xf_save(path, xf):
    """Writes a .xf file (the layout xf_load reads)."""
    cm = _f32(xf["colorMap"]).reshape(-1, 4)
    params = _f32([xf["opacityScale"], *xf["absDomain"], *xf["relDomain"]], 5)
    check(lib().exa_xf_write(_path(path), _ptr(params), _ptr(cm), len(cm)))


def xf_load(path):
    os_ = C.c_float()
    ad = np.zeros(2, np.float32)
    rd = np.zeros(2, np.float32)
    n = C.c_int()
    check(lib().exa_xf_load(_path(path), C.byref(os_), _ptr(ad), _ptr(rd), C.byref(n), None, 0))
    cm = np.empty((n.value, 4), np.float32)
    check(lib().exa_xf_load(_path(path), None, None, None, None, _ptr(cm), n.value))
    return {"opacityScale": os_.value, "absDomain": ad, "relDomain": rd, "colorMap": cm}


class NeuralField:
    """Neural field sampler: multiresolution hash encoding (levels x 2^log2T
    entries x 2 bf16 features) + MLP 32-64-64-1, trained on samples of the
    renderer's exact sampler; per-macrocell validation decides where it stands
    in for element point location."""

    def __init__(self, renderer, levels=16, log2T=19, n_min=16.0, n_max=512.0, seed=1):
        self.renderer = renderer
        self.h = C.c_void_p()
        check(lib().exa_neural_create(renderer.h, int(levels), int(log2T), float(n_min), float(n_max),
                                      int(seed), C.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None):
            lib().exa_neural_destroy(self.h)
            self.h = None

    def config(self):
        info = np.zeros(4, np.int32)
        res = np.zeros(16, np.int32)
        check(lib().exa_neural_get_config(self.h, _ptr(info), _ptr(res)))
        return {"levels": int(info[0]), "log2T": int(info[1]), "step": int(info[2]), "validated": bool(info[3]),
                "res": [int(x) for x in res[:info[0]]]}

    def params(self):
        c = self.config()
        table = np.empty((c["levels"], 1 << c["log2T"], 2), np.float32)
        W1 = np.empty((32, 64), np.float32)
        W2 = np.empty((64, 64), np.float32)
        w3 = np.empty(64, np.float32)
        check(lib().exa_neural_get_params(self.h, _ptr(table), _ptr(W1), _ptr(W2), _ptr(w3)))
        return table, W1, W2, w3

    def setParams(self, table, W1, W2, w3):
        arrs = [np.ascontiguousarray(a, np.float32) for a in (table, W1, W2, w3)]
        check(lib().exa_neural_set_params(self.h, *[_ptr(a) for a in arrs]))

    def makePool(self, count, elem_fraction=0.7, seed=7):
        nv = C.c_int64()
        check(lib().exa_neural_make_pool(self.h, self.renderer.h, int(count), float(elem_fraction), int(seed),
                                         C.byref(nv)))
        return nv.value

    def train(self, steps, batch=1 << 18, lr=1e-2, lr_end=None):
        """Adam; the learning rate decays exponentially from lr to lr_end
        (default: constant).  Returns the loss of every step."""
        loss = np.zeros(max(int(steps), 1), np.float32)
        check(lib().exa_neural_train(self.h, int(steps), int(batch), float(lr),
                                     float(lr if lr_end is None else lr_end), _ptr(loss)))
        return loss[:int(steps)]

    def gradients(self, x01, target):
        """Loss and gradients of one batch (rows in order; len % 64 == 0)."""
        c = self.config()
        x = np.ascontiguousarray(x01, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        loss = C.c_float()
        dt = np.empty((c["levels"], 1 << c["log2T"], 2), np.float32)
        d1 = np.empty((32, 64), np.float32)
        d2 = np.empty((64, 64), np.float32)
        d3 = np.empty(64, np.float32)
        check(lib().exa_neural_gradients(self.h, _ptr(x), _ptr(t), len(t), C.byref(loss), _ptr(dt), _ptr(d1),
                                         _ptr(d2), _ptr(d3)))
        return loss.value, (dt, d1, d2, d3)

    def applyAdam(self, lr):
        check(lib().exa_neural_apply_adam(self.h, float(lr)))

    def evaluate(self, d_pos, d_out, stream=None):
        """Values at device positions (n x 3 float32, voxel space)."""
        check(lib().exa_neural_evaluate(self.h, C.c_void_p(d_pos.data_ptr()), d_pos.shape[0],
                                        C.c_void_p(d_out.data_ptr()), _stream(stream)))

    def validate(self, points_per_mc=64, bound=None, seed=11):
        """bound: max abs error in scalar units (default 1e-3 of the value range)."""
        if bound is None:
            vr = self.renderer.valueRange
            bound = 1e-3 * float(vr[1] - vr[0])
        out = np.zeros(3, np.float32)
        check(lib().exa_neural_validate(self.h, self.renderer.h, int(points_per_mc), float(bound), int(seed),
                                        _ptr(out)))
        return {"neural_mcs": int(out[0]), "element_mcs": int(out[1]), "max_error": float(out[2]), "bound": bound}

    def flags(self, num_mcs):
        f = np.zeros(num_mcs, np.uint8)
        check(lib().exa_neural_get_flags(self.h, _ptr(f)))
        return f


class RenderBatch:
    """Many render requests (views, image sizes, transfer functions) in one
    persistent launch per frame, sharing a renderer's model and tables.  Each
    request gives the image the renderer would give for the same settings."""

    COUNTER_NAMES = ("gridlet", "tet", "pyr", "wedge", "hex", "miss", "cached", "dda_cells", "pixels", "tf",
                     "neural")

    def __init__(self, renderer):
        self.renderer = renderer  # keeps the model / sampler alive
        self.h = C.c_void_p()
        check(lib().exa_batch_create(renderer.h, C.byref(self.h)))
        self.sizes = []

    def __del__(self):
        if getattr(self, "h", None):
            lib().exa_batch_destroy(self.h)
            self.h = None

    def addTransferFunction(self, rgba, lo, hi, opacity_scale):
        """rgba: (N,4) texels; [lo, hi]: absolute domain; opacity_scale multiplies
        extinction directly (the renderer's slider s stores 1.1**(s-100))."""
        a = np.ascontiguousarray(rgba, dtype=np.float32)
        idx = C.c_int()
        check(lib().exa_batch_add_transfer_function(self.h, _ptr(a), len(a), float(lo), float(hi),
                                                    float(opacity_scale), C.byref(idx)))
        return idx.value

    def addRequest(self, width, height, spp, tf, camera, integrator=0):
        """camera = (org, dir_00, dir_du, dir_dv) as from camera_look_at."""
        from ._capi import BatchRequestDesc
        d = BatchRequestDesc()
        d.width, d.height, d.spp, d.integrator, d.tf = int(width), int(height), int(spp), int(integrator), int(tf)
        for name, v in zip(("camOrg", "camDir00", "camDu", "camDv"), camera):
            getattr(d, name)[:] = [float(x) for x in _f32(v, 3)]
        idx = C.c_int()
        check(lib().exa_batch_add_request(self.h, C.byref(d), C.byref(idx)))
        self.sizes.append((int(width), int(height)))
        return idx.value

    def __len__(self):
        return len(self.sizes)

    def setTuning(self, ctrl_steps, elem_batch, rare_batch):
        check(lib().exa_batch_set_tuning(self.h, int(ctrl_steps), int(elem_batch), int(rare_batch)))

    @property
    def tuning(self):
        t = np.zeros(3, np.int32)
        check(lib().exa_batch_get_tuning(self.h, _ptr(t)))
        return tuple(int(x) for x in t)

    def setLaunchTuning(self, element_stage=0, queue_rounds=1, queue_depth=0, block_size=256):
        check(lib().exa_batch_set_launch_tuning(self.h, int(element_stage), int(queue_rounds), int(queue_depth),
                                                int(block_size)))

    @property
    def launchTuning(self):
        """(elementStage, queueRounds, queueDepth, blockSize)"""
        t = np.zeros(4, np.int32)
        check(lib().exa_batch_get_launch_tuning(self.h, _ptr(t)))
        return tuple(int(x) for x in t)

    def render(self, stream=None):
        check(lib().exa_batch_render(self.h, _stream(stream)))

    def render_counted(self, stream=None):
        c = np.zeros(11, np.uint64)
        check(lib().exa_batch_render_counted(self.h, _ptr(c), _stream(stream)))
        return dict(zip(self.COUNTER_NAMES, (int(x) for x in c)))

    def captureGraph(self, stream):
        check(lib().exa_batch_capture_graph(self.h, _stream(stream)))

    def renderGraph(self, stream=None):
        check(lib().exa_batch_render_graph(self.h, _stream(stream)))

    def resetAccum(self):
        check(lib().exa_batch_reset_accum(self.h))

    def accumID(self, i):
        return check_nonneg(lib().exa_batch_get_accum_id(self.h, int(i)))

    def accum(self, i):
        w, h = self.sizes[i]
        out = np.empty((h, w, 4), np.float32)
        check(lib().exa_batch_get_accum(self.h, int(i), _ptr(out)))
        return out

    def frame(self, i):
        w, h = self.sizes[i]
        out = np.empty((h, w), np.uint32)
        check(lib().exa_batch_get_frame(self.h, int(i), _ptr(out)))
        return out

    def autotune(self, ctrl_steps=(4, 8, 16, 32), elem_batch=(0, 8, 16, 32), rare_batch=(0, 8, 16), stream=None):
        """Times one frame per knob combination, keeps the fastest; resets the
        accumulation.  Returns (best_knobs, best_ms)."""
        c = np.ascontiguousarray(ctrl_steps, np.int32)
        e = np.ascontiguousarray(elem_batch, np.int32)
        r = np.ascontiguousarray(rare_batch, np.int32)
        best = np.zeros(3, np.int32)
        ms = C.c_float()
        check(lib().exa_batch_autotune(self.h, _ptr(c), len(c), _ptr(e), len(e), _ptr(r), len(r),
                                       _stream(stream), _ptr(best), C.byref(ms)))
        return tuple(int(x) for x in best), ms.value

    def autotuneLaunch(self, ctrl_steps=(6,), elem_batch=(16,), rare_batch=(16,), element_stage=(0, 1),
                       queue_rounds=(1,), queue_depth=(0,), block_size=(64, 128, 256), stream=None):
        """The autotuner over seven axes: the scheduling knobs, where the
        element stage runs, the queue rounds and depth, and the workgroup
        size.  Returns ((ctrlSteps, elemBatch, rareBatch, elementStage,
        queueRounds, queueDepth, blockSize), best_ms)."""
        ax = [np.ascontiguousarray(a, np.int32) for a in (ctrl_steps, elem_batch, rare_batch, element_stage,
                                                          queue_rounds, queue_depth, block_size)]
        best = np.zeros(7, np.int32)
        ms = C.c_float()
        args = []
        for a in ax:
            args += [_ptr(a), len(a)]
        check(lib().exa_batch_autotune_launch(self.h, *args, _stream(stream), _ptr(best), C.byref(ms)))
        return tuple(int(x) for x in best), ms.value


def check_nonneg(rc):
    if rc < 0:
        check(rc)
    return rc
