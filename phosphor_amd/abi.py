"""numpy/ctypes views of the C ABI in include/phosphor_hip.h.

The numpy structured dtypes restate the reference's POD layouts byte for byte
(include/common/*.h); their sizes are asserted at import.  Struct arguments
cross the boundary as pointers to numpy buffers (numpy data is 16-byte
aligned), never by value.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

F4 = ("<f4", (4,))

CAMERA_DTYPE = np.dtype([  # camera.h:7-22
    ("position", *F4), ("lower_left_corner", *F4), ("horizontal", *F4), ("vertical", *F4),
    ("defocus_disk_u", *F4), ("defocus_disk_v", *F4),
    ("defocus_angle", "<f4"), ("focus_distance", "<f4"), ("_padding", "u1", (8,)),
])
MATERIAL_DTYPE = np.dtype([  # material.h:36-59
    ("base_color", *F4), ("emissive", *F4), ("att_color", *F4),
    ("metallic", "<f4"), ("roughness", "<f4"), ("transmission", "<f4"), ("ior", "<f4"),
    ("att_dist", "<f4"), ("thickness", "<f4"),
    ("diff_index", "<u4"), ("emis_index", "<u4"), ("norm_index", "<u4"), ("occlusion_index", "<u4"),
    ("metal_rough_index", "<u4"), ("transmission_index", "<u4"),
])
UV_TRANSFORM_DTYPE = np.dtype([  # material.h:7-17
    ("uv_offset", "<f4", (2,)), ("uv_scale", "<f4", (2,)), ("uv_rotation", "<f4"), ("_padding", "u1", (12,)),
])
UV_BLOCK_DTYPE = np.dtype([(k, UV_TRANSFORM_DTYPE) for k in
                           ("diff", "emis", "norm", "occlusion", "metal_rough", "transmission")])
LIGHT_DTYPE = np.dtype([  # light.h:14-31
    ("position", *F4), ("power", *F4), ("direction", *F4), ("tangent", *F4), ("bitangent", *F4),
    ("origin", *F4), ("aux", *F4), ("kind", "u1"), ("_padding", "u1", (15,)),
])
HASH_DTYPE = np.dtype([  # photon_hash.h:11-19
    ("origin", *F4), ("cell_sizes", *F4), ("grid_res", "<u4"), ("_padding", "u1", (12,)),
])
TEX_META_DTYPE = np.dtype([("atlas_offset", "<u4"), ("width", "<u4"), ("height", "<u4"), ("channels", "<u4")])

for _dt, _size in ((CAMERA_DTYPE, 112), (MATERIAL_DTYPE, 96), (UV_BLOCK_DTYPE, 192), (LIGHT_DTYPE, 128),
                   (HASH_DTYPE, 48), (TEX_META_DTYPE, 16)):
    assert _dt.itemsize == _size, (_dt, _dt.itemsize, _size)

LIGHT_POINT, LIGHT_SPOT, LIGHT_DIRECTIONAL, LIGHT_TEXTURED = 0, 1, 2, 3
NO_TEXTURE = 0xFFFFFFFF

PHX_OK = 0
PHX_ERR_INVALID_ARG = 10001
PHX_ERR_NO_PHOTONS = 10002
PHX_ERR_WORKSPACE = 10003
PHX_ERR_NO_DEVICE = 10004
PHX_ERR_TOO_LARGE = 10005


class RenderArgs(C.Structure):
    """PhxRenderArgs — the ArgsList fields the pass reads (cmd_args.hpp:36-65)."""

    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32), ("photons_per_round", C.c_uint32),
        ("max_photons_per_batch", C.c_uint32), ("max_photon_bounces", C.c_uint32), ("grid_res", C.c_uint32),
        ("direct_samples", C.c_uint32), ("seed", C.c_uint32), ("sppm_alpha", C.c_float),
        ("initial_radius", C.c_float), ("trace_mode", C.c_uint32), ("flags", C.c_uint32),
        ("emit_wave_bounces", C.c_uint32), ("bvh_leaf_size", C.c_uint32),
        ("camera_block", C.c_uint32), ("shadow_block", C.c_uint32), ("emit_block", C.c_uint32),
        ("gather_tile_width", C.c_uint32),
    ]


PHX_PASS_GRAPH = 1
PHX_REFERENCE_BVH = 2
PHX_PASS_SERIAL = 4
PHX_FUSED_TRACE = 8
PHX_PAIR_MAX_LEAF = 4


class BatchOptions(C.Structure):
    """PhxBatchOptions — group sizes of a multi-scene pass."""

    _fields_ = [("trace_group_pixels", C.c_uint32), ("emit_group_photons", C.c_uint32), ("flags", C.c_uint32),
                ("reserved", C.c_uint32)]


class KernelInfo(C.Structure):
    """PhxKernelInfo — registers and resident waves of one kernel."""

    _fields_ = [("name", C.c_char * 32), ("vgprs", C.c_uint32), ("lds_bytes", C.c_uint32),
                ("scratch_bytes", C.c_uint32), ("block", C.c_uint32), ("waves_per_simd", C.c_uint32)]


class SceneDesc(C.Structure):
    """PhxSceneDesc — host pointers of one scene (SceneData after the SoA split)."""

    _fields_ = [
        ("tri_verts", C.c_void_p), ("tri_normals", C.c_void_p), ("tri_tangents", C.c_void_p),
        ("tri_uvs", C.c_void_p), ("tri_mat_index", C.c_void_p), ("n_triangles", C.c_uint32),
        ("etri_verts", C.c_void_p), ("etri_normals", C.c_void_p), ("etri_uvs", C.c_void_p),
        ("etri_mat_index", C.c_void_p), ("n_e_triangles", C.c_uint32),
        ("materials", C.c_void_p), ("uv_transforms", C.c_void_p), ("n_materials", C.c_uint32),
        ("lights", C.c_void_p), ("n_lights", C.c_uint32),
        ("tex_meta", C.c_void_p), ("n_textures", C.c_uint32), ("tex_atlas", C.c_void_p),
        ("tex_atlas_bytes", C.c_uint64),
        ("envmap", C.c_void_p), ("envmap_width", C.c_uint32), ("envmap_height", C.c_uint32),
        ("camera", C.c_void_p),
        ("bvh_nodes", C.c_void_p), ("n_bvh_nodes", C.c_uint32),
    ]


def _sig(lib, name, restype, *argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = list(argtypes)
    return fn


VP, U32, F32, SZ, INT = C.c_void_p, C.c_uint32, C.c_float, C.c_size_t, C.c_int

# name -> (restype, argtypes): every symbol include/phosphor_hip.h declares
SIGNATURES = {
    "phx_last_error": (C.c_char_p,),
    "phx_version": (C.c_char_p,),
    "phx_trace_rays": (INT, VP, U32, U32, U32, U32, VP, VP, VP, VP, VP, VP, U32, VP, VP, VP, VP, VP, U32, VP, F32,
                       VP, F32, VP, VP, VP, VP, VP, U32, U32, U32, VP, VP, VP, VP, VP, VP, VP),
    "phx_trace_rays_wavefront_workspace_size": (SZ, U32, U32, U32),
    "phx_trace_rays_wavefront": (INT, VP, U32, U32, U32, U32, VP, VP, VP, VP, VP, VP, U32, VP, VP, VP, VP, VP, U32,
                                 VP, F32, VP, F32, VP, VP, VP, VP, VP, U32, U32, U32, VP, VP, VP, VP, VP, VP, U32,
                                 VP, SZ, VP),
    "phx_emit_photons_workspace_size": (SZ, U32, U32),
    "phx_emit_photons": (INT, VP, U32, U32, U32, U32, U32, U32, VP, VP, VP, VP, VP, VP, U32, VP, VP, VP, VP, VP, VP,
                         VP, VP, VP, F32, VP, F32, VP, VP, VP, VP, VP, SZ, VP),
    "phx_hash_photons": (INT, VP, U32, VP, VP, VP, VP),
    "phx_scan_cells_workspace_size": (SZ, U32),
    "phx_scan_cells": (INT, VP, U32, VP, VP, VP, SZ, VP),
    "phx_scan_cells_noalloc": (INT, VP, U32, VP, VP, VP),
    "phx_scatter_photons": (INT, VP, VP, VP, VP, VP, U32, VP, VP, VP, VP),
    "phx_sort_photons_workspace_size": (SZ, U32, U32),
    "phx_sort_photons": (INT, VP, VP, VP, VP, VP, VP, U32, U32, VP, VP, VP, VP, SZ, VP),
    "phx_gather_flux": (INT, VP, VP, VP, VP, VP, VP, U32, VP, VP, VP, U32, VP, VP, VP, F32, VP, VP, VP, VP),
    "phx_accel_eligible": (INT, VP, U32, VP, U32),
    "phx_accel_attach": (INT, VP, VP, VP, U32, VP, U32, U32, VP),
    "phx_accel_detach": (INT, VP),
    "phx_sort_photons_by_octant_workspace_size": (SZ, U32, U32),
    "phx_sort_photons_by_octant": (INT, VP, VP, VP, U32, VP, U32, VP, VP, VP, VP, VP, VP, VP, VP, SZ, VP),
    "phx_gather_flux_octants": (INT, VP, VP, VP, VP, VP, VP, U32, U32, VP, VP, U32, VP, VP, VP, VP, F32, VP, VP, VP,
                                VP),
    "phx_device_count": (INT,),
    "phx_kernel_info": (INT, C.POINTER(KernelInfo), U32),
    "phx_device_malloc": (INT, C.POINTER(VP), SZ),
    "phx_device_free": (INT, VP),
    "phx_memcpy_htod": (INT, VP, VP, SZ),
    "phx_memcpy_dtoh": (INT, VP, VP, SZ),
    "phx_memset_device": (INT, VP, INT, SZ),
    "phx_device_copy": (INT, VP, VP, C.c_size_t, VP),
    "phx_device_synchronize": (INT,),
    "phx_build_bvh_median": (U32, VP, U32, VP),
    "phx_build_bvh_sah": (U32, VP, U32, U32, U32, VP, VP),
    "phx_build_bvh_pairs": (INT, VP, U32, VP, VP),
    "phx_build_photon_hash_info": (INT, VP, U32, VP),
    "phx_renderer_create": (INT, C.POINTER(RenderArgs), C.POINTER(SceneDesc), C.POINTER(VP)),
    "phx_renderer_destroy": (None, VP),
    "phx_renderer_run_pass": (INT, VP),
    "phx_renderer_synchronize": (INT, VP),
    "phx_renderer_last_stage_ms": (INT, VP, VP),
    "phx_renderer_gather_ms": (INT, VP, VP, U32, VP),
    "phx_renderer_reset_gather_ms": (INT, VP),
    "phx_renderer_last_pass_ms": (INT, VP, VP),
    "phx_renderer_checkpoint_bytes": (C.c_size_t, VP),
    "phx_renderer_save_checkpoint": (INT, VP, VP, C.c_size_t),
    "phx_renderer_load_checkpoint": (INT, VP, VP, C.c_size_t),
    "phx_renderer_last_photon_count": (U32, VP),
    "phx_renderer_rounds_done": (U32, VP),
    "phx_renderer_read_state": (INT, VP, VP, VP, VP),
    "phx_renderer_read_hitpoints": (INT, VP, VP),
    "phx_renderer_read_photon_map": (INT, VP, VP, VP, VP, VP, VP),
    "phx_renderer_device_buffers": (INT, VP, VP, U32),
    "phx_renderer_estimate": (INT, VP, VP),
    "phx_adjoint_create": (INT, C.POINTER(RenderArgs), C.POINTER(SceneDesc), U32, C.POINTER(VP)),
    "phx_adjoint_destroy": (None, VP),
    "phx_adjoint_num_params": (U32, VP),
    "phx_adjoint_forward": (INT, VP, VP, U32, VP),
    "phx_adjoint_backward": (INT, VP, VP, VP),
    "phx_adjoint_set_replay_check": (INT, VP, U32),
    "phx_adjoint_read_replay": (INT, VP, VP, VP, VP, VP),
    "phx_adjoint_renderer": (VP, VP),
    "phx_adjoint_ggx_rejection": (INT, F32, F32, VP),
    "phx_batch_create": (INT, VP, VP, U32, VP, C.POINTER(VP)),
    "phx_batch_destroy": (None, VP),
    "phx_batch_run_pass": (INT, VP),
    "phx_batch_synchronize": (INT, VP),
    "phx_batch_last_pass_ms": (INT, VP, VP),
    "phx_batch_rounds_done": (U32, VP),
    "phx_batch_photon_counts": (INT, VP, VP),
    "phx_batch_read_state": (INT, VP, U32, VP, VP, VP),
    "phx_batch_read_hitpoints": (INT, VP, U32, VP),
    "phx_batch_read_photon_map": (INT, VP, U32, VP, VP, VP, VP, VP),
    "phx_batch_estimate": (INT, VP, U32, VP),
}


def bind(lib):
    for name, (res, *args) in SIGNATURES.items():
        _sig(lib, name, res, *args)
    return lib


def ptr(a) -> int | None:
    """Address of a numpy array's data (None for None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays crossing the C ABI must be contiguous"
    return a.ctypes.data
