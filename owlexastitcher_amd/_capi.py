"""ctypes binding of libexastitch.so (include/exa/exa.h).

The library is built in-tree (``owlexastitcher_amd/libexastitch.so``) by
``__graft_entry__.build()``; importing this module fails loudly when it is
missing -- there is no fallback implementation.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libexastitch.so")

EXA_OK = 0
EXA_ERR_INVALID = -1
EXA_ERR_HIP = -2
EXA_ERR_STATE = -3
EXA_ERR_ALLOC = -4


class ExaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"exa error {code}: {msg}")
        self.code = code


class Gridlet(C.Structure):
    """Gridlet.h:22-37"""
    _fields_ = [("lower", C.c_int32 * 3), ("level", C.c_int32), ("dims", C.c_int32 * 3), ("begin", C.c_uint32)]


class Sample(C.Structure):
    """sampler/Sampler.h:72-76"""
    _fields_ = [("primID", C.c_int32), ("cellID", C.c_int32), ("value", C.c_float)]


class AmrCell(C.Structure):
    _fields_ = [("pos", C.c_int32 * 3), ("level", C.c_int32)]


class StitchModelDesc(C.Structure):
    _fields_ = [
        ("vertices", C.c_void_p), ("numVertices", C.c_int64),
        ("indices", C.c_void_p), ("numElements", C.c_int64),
        ("gridlets", C.c_void_p), ("numGridlets", C.c_int64),
        ("gridletScalars", C.c_void_p), ("numGridletScalars", C.c_int64),
    ]


class BatchRequestDesc(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32), ("integrator", C.c_int32),
        ("tf", C.c_int32),
        ("camOrg", C.c_float * 3), ("camDir00", C.c_float * 3), ("camDu", C.c_float * 3),
        ("camDv", C.c_float * 3),
    ]


class RenderState(C.Structure):
    _fields_ = [
        ("W", C.c_int32), ("H", C.c_int32), ("spp", C.c_int32), ("accumID", C.c_int32),
        ("integrator", C.c_int32), ("shadeMode", C.c_int32), ("maxBounces", C.c_int32),
        ("debugCrosshair", C.c_int32),
        ("camOrg", C.c_float * 3), ("camDir00", C.c_float * 3), ("camDu", C.c_float * 3),
        ("camDv", C.c_float * 3),
        ("worldBounds", C.c_float * 6),
        ("voxelXfm", C.c_float * 12), ("voxelXfmInv", C.c_float * 12), ("lightXfm", C.c_float * 12),
        ("gridDims", C.c_int32 * 3), ("gridBounds", C.c_float * 6),
        ("tfDomain", C.c_float * 2), ("opacityScale", C.c_float), ("tfN", C.c_int32),
        ("clipEnabled", C.c_int32), ("clipN", C.c_float * 3), ("clipD", C.c_float),
        ("lightPos", (C.c_float * 3) * 4), ("lightIntensity", C.c_float * 4),
        ("lightOn", C.c_int32 * 4), ("lightType", C.c_int32 * 4),
        ("dt", C.c_float), ("heatMapEnabled", C.c_int32), ("heatMapScale", C.c_float),
        ("subImageActive", C.c_int32), ("subImage", C.c_float * 4),
        ("subImageSelecting", C.c_int32), ("subImageSel", C.c_float * 4),
    ]


P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
F = C.c_float

# name -> (restype, argtypes)
_SIGS = {
    "exa_get_last_error": (C.c_char_p, []),
    "exa_version": (C.c_int, []),
    "exa_amr_create_gaussians": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, P]),
    "exa_amr_create_noise": (C.c_int, [C.c_int, C.c_int, F, F, C.c_int, F, C.c_uint64, P]),
    "exa_amr_create_shock": (C.c_int, [C.c_int, C.c_int, F, F, F, F, C.c_uint64, P]),
    "exa_amr_count_shock": (C.c_int, [C.c_int, C.c_int, F, F, F, F, C.c_uint64, P]),
    "exa_amr_create_from_cells": (C.c_int, [P, P, I64, P]),
    "exa_amr_load": (C.c_int, [C.c_char_p, C.c_char_p, P]),
    "exa_amr_save": (C.c_int, [P, C.c_char_p, C.c_char_p]),
    "exa_amr_write_dual_mesh": (C.c_int, [P, C.c_char_p, C.c_char_p, C.c_char_p, P]),
    "exa_make_grids": (C.c_int, [P, C.c_int, C.c_char_p, P]),
    "exa_model_load": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, P]),
    "exa_model_get": (C.c_int, [P, P, P, P, P]),
    "exa_xf_write": (C.c_int, [C.c_char_p, P, P, C.c_int]),
    "exa_amr_num_cells": (I64, [P]),
    "exa_amr_get": (C.c_int, [P, P, P]),
    "exa_amr_get_params": (C.c_int, [P, P, C.c_int]),
    "exa_amr_destroy": (None, [P]),
    "exa_dual_mesh_build": (C.c_int, [P, P]),
    "exa_dual_mesh_counts": (C.c_int, [P, P]),
    "exa_dual_mesh_get": (C.c_int, [P, P, P, P, P]),
    "exa_dual_mesh_destroy": (None, [P]),
    "exa_model_create": (C.c_int, [P, P]),
    "exa_model_create_from_dual_mesh": (C.c_int, [P, P]),
    "exa_model_get_bounds": (C.c_int, [P, P, P]),
    "exa_model_get_counts": (C.c_int, [P, P]),
    "exa_model_set_num_grid_cells": (C.c_int, [P, P]),
    "exa_model_set_voxel_space_transform": (C.c_int, [P, P, P]),
    "exa_model_destroy": (None, [P]),
    "exa_sampler_create": (C.c_int, [P]),
    "exa_sampler_build": (C.c_int, [P, P, P]),
    "exa_sampler_compute_max_opacities": (C.c_int, [P, P, C.c_int, F, F, P]),
    "exa_sampler_locate": (C.c_int, [P, P, I64, P, P]),
    "exa_sampler_locate_counted": (C.c_int, [P, P, I64, P, P, P]),
    "exa_sampler_locate_noalloc": (C.c_int, [P, P, I64, P, P]),
    "exa_sampler_test_sample": (C.c_int, [P, P, P, I64, P, P]),
    "exa_sampler_get_grid_info": (C.c_int, [P, P, P]),
    "exa_sampler_get_value_ranges": (C.c_int, [P, P]),
    "exa_sampler_get_max_opacities": (C.c_int, [P, P]),
    "exa_sampler_get_gridlet_value_ranges": (C.c_int, [P, P]),
    "exa_sampler_get_prim_max_opacities": (C.c_int, [P, P, P]),
    "exa_sampler_get_info": (C.c_int, [P, P]),
    "exa_sampler_set_accel": (C.c_int, [P, C.c_int]),
    "exa_sampler_get_accel": (C.c_int, [P]),
    "exa_sampler_destroy": (None, [P]),
    "exa_renderer_create": (C.c_int, [P, P, P, P, P]),
    "exa_renderer_set_camera": (C.c_int, [P, P, P, P, P]),
    "exa_renderer_resize": (C.c_int, [P, C.c_int, C.c_int]),
    "exa_renderer_render": (C.c_int, [P, P, P]),
    "exa_renderer_render_counted": (C.c_int, [P, P, P, P]),
    "exa_renderer_get_accum": (C.c_int, [P, P]),
    "exa_renderer_set_type": (C.c_int, [P, C.c_int]),
    "exa_renderer_set_color_map": (C.c_int, [P, P, C.c_int]),
    "exa_renderer_set_range": (C.c_int, [P, F, F]),
    "exa_renderer_set_rel_domain": (C.c_int, [P, F, F]),
    "exa_renderer_set_opacity_scale": (C.c_int, [P, F]),
    "exa_renderer_set_clip_plane": (C.c_int, [P, C.c_int, C.c_int, P, F]),
    "exa_renderer_set_shade_mode": (C.c_int, [P, C.c_int]),
    "exa_renderer_set_dt": (C.c_int, [P, F]),
    "exa_renderer_set_heat_map": (C.c_int, [P, C.c_int, F]),
    "exa_renderer_set_light_source": (C.c_int, [P, C.c_int, P, F, C.c_int]),
    "exa_renderer_set_directional_light": (C.c_int, [P, C.c_int, P, F, C.c_int]),
    "exa_renderer_set_light_space_transform": (C.c_int, [P, P]),
    "exa_renderer_set_sub_image": (C.c_int, [P, P, C.c_int]),
    "exa_renderer_set_sub_image_selection": (C.c_int, [P, P, C.c_int]),
    "exa_renderer_reset_accum": (C.c_int, [P]),
    "exa_renderer_set_spp": (C.c_int, [P, C.c_int]),
    "exa_renderer_set_max_bounces": (C.c_int, [P, C.c_int]),
    "exa_renderer_set_debug_crosshair": (C.c_int, [P, C.c_int]),
    "exa_renderer_get_bounds": (C.c_int, [P, P]),
    "exa_renderer_get_value_range": (C.c_int, [P, P]),
    "exa_renderer_get_accum_id": (C.c_int, [P]),
    "exa_renderer_set_accum_id": (C.c_int, [P, C.c_int]),
    "exa_renderer_get_sampler": (P, [P]),
    "exa_renderer_set_accel": (C.c_int, [P, C.c_int]),
    "exa_renderer_set_scheme": (C.c_int, [P, C.c_int]),
    "exa_renderer_get_scheme": (C.c_int, [P]),
    "exa_renderer_set_persistent_tuning": (C.c_int, [P, C.c_int, C.c_int, C.c_int]),
    "exa_renderer_set_launch_tuning": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int]),
    "exa_renderer_get_tuning": (C.c_int, [P, P]),
    "exa_renderer_get_state": (C.c_int, [P, P]),
    "exa_renderer_destroy": (None, [P]),
    "exa_batch_create": (C.c_int, [P, P]),
    "exa_batch_destroy": (None, [P]),
    "exa_batch_add_transfer_function": (C.c_int, [P, P, C.c_int, F, F, F, P]),
    "exa_batch_add_request": (C.c_int, [P, P, P]),
    "exa_batch_num_requests": (C.c_int, [P]),
    "exa_batch_set_tuning": (C.c_int, [P, C.c_int, C.c_int, C.c_int]),
    "exa_batch_set_launch_tuning": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int]),
    "exa_batch_get_launch_tuning": (C.c_int, [P, P]),
    "exa_batch_get_tuning": (C.c_int, [P, P]),
    "exa_batch_render": (C.c_int, [P, P]),
    "exa_batch_render_counted": (C.c_int, [P, P, P]),
    "exa_batch_capture_graph": (C.c_int, [P, P]),
    "exa_batch_render_graph": (C.c_int, [P, P]),
    "exa_batch_reset_accum": (C.c_int, [P]),
    "exa_batch_get_accum_id": (C.c_int, [P, C.c_int]),
    "exa_batch_get_accum": (C.c_int, [P, C.c_int, P]),
    "exa_batch_get_frame": (C.c_int, [P, C.c_int, P]),
    "exa_batch_autotune": (C.c_int, [P, P, C.c_int, P, C.c_int, P, C.c_int, P, P, P]),
    "exa_batch_autotune_launch": (C.c_int, [P, P, C.c_int, P, C.c_int, P, C.c_int, P, C.c_int, P, C.c_int, P,
                                            C.c_int, P, C.c_int, P, P, P]),
    "exa_neural_create": (C.c_int, [P, C.c_int, C.c_int, F, F, C.c_uint32, P]),
    "exa_neural_destroy": (None, [P]),
    "exa_neural_get_config": (C.c_int, [P, P, P]),
    "exa_neural_get_params": (C.c_int, [P, P, P, P, P]),
    "exa_neural_set_params": (C.c_int, [P, P, P, P, P]),
    "exa_neural_make_pool": (C.c_int, [P, P, C.c_int64, F, C.c_uint32, P]),
    "exa_neural_train": (C.c_int, [P, C.c_int, C.c_int, F, F, P]),
    "exa_neural_gradients": (C.c_int, [P, P, P, C.c_int, P, P, P, P, P]),
    "exa_neural_apply_adam": (C.c_int, [P, F]),
    "exa_neural_evaluate": (C.c_int, [P, P, C.c_int64, P, P]),
    "exa_neural_validate": (C.c_int, [P, P, C.c_int, F, C.c_uint32, P]),
    "exa_neural_get_flags": (C.c_int, [P, P]),
    "exa_renderer_set_neural": (C.c_int, [P, P]),
    "exa_camera_look_at": (C.c_int, [P, P, P, F, C.c_int, C.c_int, P, P, P, P]),
    "exa_xf_load": (C.c_int, [C.c_char_p, P, P, P, P, P, C.c_int]),
    "exa_tf_rasterize": (C.c_int, [P, C.c_int, C.c_int, P]),
    "exa_brick_model_create": (C.c_int, [P, P, C.c_size_t, P]),
    "exa_brick_model_load": (C.c_int, [C.c_char_p, C.c_char_p, P]),
    "exa_brick_model_from_amr": (C.c_int, [P, C.c_int, P]),
    "exa_amr_write_bricks": (C.c_int, [P, C.c_int, C.c_char_p, C.c_char_p]),
    "exa_brick_model_get_counts": (C.c_int, [P, P]),
    "exa_brick_model_get": (C.c_int, [P, P, P, P, P]),
    "exa_brick_model_get_kd_nodes": (C.c_int, [P, P]),
    "exa_brick_model_get_bounds": (C.c_int, [P, P, P]),
    "exa_brick_model_destroy": (None, [P]),
    "exa_brick_sampler_create": (C.c_int, [P, P, P]),
    "exa_brick_sampler_compute_max_opacities": (C.c_int, [P, P, C.c_int, F, F, P]),
    "exa_brick_sampler_get_max_opacities": (C.c_int, [P, P]),
    "exa_brick_sampler_locate": (C.c_int, [P, P, I64, P, P]),
    "exa_brick_sampler_destroy": (None, [P]),
    "exa_renderer_create_bricks": (C.c_int, [P, P, P, P, P]),
    "exa_renderer_get_brick_sampler": (P, [P]),
}

EXPORTED = sorted(_SIGS)
_lib = None


def lib():
    """Load libexastitch.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built: run __graft_entry__.build() "
                              "(make -C owlexastitcher_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != EXA_OK:
        msg = lib().exa_get_last_error()
        raise ExaError(rc, msg.decode() if msg else "")
    return rc
