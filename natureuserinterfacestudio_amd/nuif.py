"""ctypes binding of the C ABI (include/nuif.h) plus small device-array helpers.

This is the Python counterpart of the binding a host application would add;
the product computation all happens in libnuif.so.
"""
import ctypes as C
import os

import numpy as np

from . import libkinfu, libnuif

c_void_p = C.c_void_p
c_int = C.c_int
c_float = C.c_float
c_size_t = C.c_size_t


class NuiCLCameraParams(C.Structure):
    _fields_ = [("fx", c_float), ("fy", c_float), ("fx_inv", c_float), ("fy_inv", c_float), ("cx", c_float),
                ("cy", c_float), ("depthImageWidth", C.c_uint32), ("depthImageHeight", C.c_uint32),
                ("sensorDepthWorldMin", c_float), ("sensorDepthWorldMax", c_float)]


class NuiCLRigidTransform(C.Structure):
    _fields_ = [("R", c_float * 9), ("R_inv", c_float * 9), ("t", c_float * 3)]


class TsdfParams(C.Structure):
    _fields_ = [("resolution", c_float * 3), ("dimension", c_float * 3), ("cell_size", c_float * 3),
                ("tranc_dist", c_float)]


class HashParams(C.Structure):
    _fields_ = [("voxel_size", c_float), ("trunc_scale", c_float), ("truncation", c_float), ("weight_max", c_int),
                ("buckets_log2", c_int), ("bricks", c_int), ("storage", c_int), ("color", c_int)]


def hash_params(voxel=0.005, trunc_scale=0.02, truncation=0.06, weight_max=255, buckets_log2=20, bricks=1 << 18,
                storage=0, color=False):
    """nuif_hash_params with the reference's defaults (VO/NuiHashingSDFConfig.cpp:3-15)."""
    p = HashParams()
    p.voxel_size, p.trunc_scale, p.truncation, p.weight_max = voxel, trunc_scale, truncation, weight_max
    p.buckets_log2, p.bricks, p.storage, p.color = buckets_log2, bricks, storage, 1 if color else 0
    return p


ICP_TERMS = 29
ICP_MAX_LEVELS = 4
POSE_ROW = 16
TSDF_S16, TSDF_F16, TSDF_F32 = 0, 1, 2
ON_FAILURE_KEEP, ON_FAILURE_RESET, ON_FAILURE_EVICT = 0, 1, 2


class IcpResult(C.Structure):
    _fields_ = [("R", c_float * 9), ("t", c_float * 3), ("error", c_float), ("count", c_float),
                ("status", C.c_int32), ("iterations", C.c_int32), ("integrate", C.c_int32), ("frame", C.c_int32),
                ("sums", C.c_double * ICP_TERMS), ("Rd", C.c_double * 9), ("td", C.c_double * 3),
                ("level_iterations", C.c_int32 * ICP_MAX_LEVELS)]


class IcpConfig(C.Structure):
    _fields_ = [("levels", c_int), ("iterations", c_int * ICP_MAX_LEVELS), ("dist_threshold", c_float),
                ("gate_threshold", c_float), ("reduction", c_int)]


ICP_REDUCE_REFERENCE, ICP_REDUCE_COMPENSATED, ICP_REDUCE_JOINT = 0, 1, 2


class IcpPhotometric(C.Structure):
    _fields_ = [("intensity_pyr", c_void_p * ICP_MAX_LEVELS), ("ref_model", c_void_p), ("weight", c_float),
                ("max_difference", c_float), ("min_gradient", c_float)]


class ColorIcpResult(C.Structure):
    _fields_ = [("ok", C.c_int32), ("iterations", C.c_int32), ("error", c_float), ("count", c_float),
                ("R", c_float * 9), ("t", c_float * 3), ("sums", c_float * 29)]


class KinfuDesc(C.Structure):
    _fields_ = [("width", c_int), ("height", c_int), ("fx", c_float), ("fy", c_float), ("cx", c_float),
                ("cy", c_float), ("depth_min", c_float), ("depth_max", c_float), ("filter_radius", c_int),
                ("sigma_space2_inv_half", c_float), ("sigma_depth2_inv_half", c_float),
                ("depth_threshold", c_float), ("dist_threshold", c_float), ("levels", c_int),
                ("iterations", c_int * 4), ("integration_threshold", c_float), ("resolution", c_int * 3),
                ("dimensions", c_float * 3), ("tranc_dist", c_float), ("color", c_int), ("storage", c_int),
                ("use_graph", c_int), ("icp_reduction", c_int), ("scene", c_int), ("hash_voxel_size", c_float),
                ("hash_trunc_scale", c_float), ("hash_truncation", c_float), ("hash_weight_max", c_int),
                ("hash_bricks", c_int), ("hash_forward_render", c_int), ("voxel_shift", c_int),
                ("volume_basis", c_float * 3), ("hash_swapping", c_int), ("hash_buckets", c_int)]


SCENE_DENSE, SCENE_HASH, SCENE_SHIFT = 0, 1, 2
RENDER_CASTRAY, RENDER_FORWARD, RENDER_FORWARD_REFERENCE = 0, 1, 2


_SIGS = {
    "nuif_init": (c_int, [c_int]),
    "nuif_device_count": (c_int, [C.POINTER(c_int)]),
    "nuif_device_info": (c_int, [C.c_char_p, c_int, C.POINTER(c_int), C.POINTER(c_size_t), C.POINTER(c_int)]),
    "nuif_last_error": (C.c_char_p, []),
    "nuif_version": (C.c_char_p, []),
    "nuif_malloc": (c_int, [C.POINTER(c_void_p), c_size_t]),
    "nuif_free": (c_int, [c_void_p]),
    "nuif_mem_in_use": (c_int, [C.POINTER(c_size_t)]),
    "nuif_host_alloc": (c_int, [C.POINTER(c_void_p), c_size_t]),
    "nuif_host_free": (c_int, [c_void_p]),
    "nuif_memcpy_h2d": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "nuif_memcpy_d2h": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "nuif_memcpy_d2d": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "nuif_fill_u8": (c_int, [c_void_p, C.c_uint8, c_size_t, c_void_p]),
    "nuif_fill_u32": (c_int, [c_void_p, C.c_uint32, c_size_t, c_void_p]),
    "nuif_stream_create": (c_int, [C.POINTER(c_void_p)]),
    "nuif_stream_destroy": (c_int, [c_void_p]),
    "nuif_stream_synchronize": (c_int, [c_void_p]),
    "nuif_device_synchronize": (c_int, []),
    "nuif_event_create": (c_int, [C.POINTER(c_void_p)]),
    "nuif_event_destroy": (c_int, [c_void_p]),
    "nuif_event_record": (c_int, [c_void_p, c_void_p]),
    "nuif_event_synchronize": (c_int, [c_void_p]),
    "nuif_event_elapsed_ms": (c_int, [c_void_p, c_void_p, C.POINTER(c_float)]),
    "nuif_stream_wait_event": (c_int, [c_void_p, c_void_p]),
    "nuif_graph_begin": (c_int, [c_void_p]),
    "nuif_graph_end": (c_int, [c_void_p, C.POINTER(c_void_p)]),
    "nuif_graph_launch": (c_int, [c_void_p, c_void_p]),
    "nuif_graph_destroy": (c_int, [c_void_p]),
    "nuif_depth_passing_filter": (c_int, [c_void_p, c_void_p, c_size_t, c_float, c_float, c_void_p]),
    "nuif_bilateral_filter_depth": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_int, c_float, c_float,
                                            c_void_p]),
    "nuif_half_sample": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_float, c_void_p]),
    "nuif_depth2vertex": (c_int, [c_void_p, c_void_p, c_int, c_int, C.POINTER(NuiCLCameraParams), c_int,
                                  c_void_p]),
    "nuif_estimate_normals_simple": (c_int, [c_void_p, c_void_p, c_int, c_int, c_float, c_void_p]),
    "nuif_transform_maps": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "nuif_scale_depths": (c_int, [c_void_p, c_int, c_int, C.POINTER(NuiCLCameraParams), c_void_p]),
    "nuif_frame_preprocess": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                                      C.POINTER(NuiCLCameraParams), c_void_p, c_int, c_float, c_float, c_int,
                                      c_void_p]),
    "nuif_icp_state_bytes": (c_size_t, []),
    "nuif_icp_partials_bytes": (c_size_t, [c_int, c_int]),
    "nuif_icp_state_init": (c_int, [c_void_p, c_void_p, c_void_p]),
    "nuif_icp_state_read": (c_int, [c_void_p, C.POINTER(IcpResult), c_void_p]),
    "nuif_icp_begin": (c_int, [c_void_p, c_void_p, c_void_p]),
    "nuif_icp_iteration": (c_int, [c_void_p, c_void_p, C.POINTER(NuiCLCameraParams), c_void_p, c_int, c_void_p,
                                   c_void_p, c_void_p, c_float, C.c_uint32, c_int, c_int, c_void_p]),
    "nuif_icp_set_translation": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "nuif_icp_end": (c_int, [c_void_p, c_void_p, c_float, c_void_p, c_int, c_void_p]),
    "nuif_icp_record": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "nuif_icp_batch_table_bytes": (c_size_t, [c_int]),
    "nuif_icp_static_scratch_bytes": (c_size_t, []),
    "nuif_fusion_round_table_bytes": (c_size_t, [c_int]),
    "nuif_fusion_round": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "nuif_fusion_round_prepare": (c_int, [c_void_p, c_int, c_void_p, c_void_p, C.POINTER(C.c_uint64)]),
    "nuif_icp_estimate_static": (c_int, [c_void_p, C.POINTER(NuiCLCameraParams), c_void_p, c_void_p, c_void_p,
                                         c_void_p, C.POINTER(IcpConfig), c_void_p, c_int, c_void_p]),
    "nuif_icp_estimate_batch": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "nuif_fusion_round_timed": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nuif_tsdf_footprint_bytes": (c_size_t, [c_void_p]),
    "nuif_schedule": (c_int, [c_void_p]),
    "nuif_map_depth_to_color": (c_int, [c_void_p, c_void_p, C.c_uint32, c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "nuif_schedule_segment": (c_int, [c_int]),
    "nuif_tsdf_raycast_footprint": (C.c_longlong, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                                    c_void_p, c_void_p, c_void_p]),
    "nuif_icp_estimate": (c_int, [c_void_p, c_void_p, C.POINTER(NuiCLCameraParams), c_void_p, c_void_p, c_void_p,
                                  c_void_p, C.POINTER(IcpConfig), c_void_p, c_int, c_void_p]),
    "nuif_icp_gate_flag": (c_void_p, [c_void_p]),
    "nuif_icp_ok_flag": (c_void_p, [c_void_p]),
    "nuif_icp_count_thresholds": (c_int, [c_int, c_int, C.POINTER(IcpConfig), C.POINTER(C.c_uint32)]),
    "nuif_tsdf_voxel_bytes": (c_size_t, [c_int]),
    "nuif_tsdf_params": (c_int, [c_int, c_int, c_int, c_float, c_float, c_float, c_float, C.POINTER(TsdfParams)]),
    "nuif_tsdf_reset": (c_int, [c_void_p, c_void_p, C.POINTER(TsdfParams), c_int, c_void_p]),
    "nuif_tsdf_integrate": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, C.POINTER(TsdfParams),
                                    C.POINTER(NuiCLCameraParams), c_void_p, c_int, c_int, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_void_p]),
    "nuif_tsdf_depth_bounds_bytes": (c_size_t, [c_int, c_int]),
    "nuif_tsdf_depth_bounds": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "nuif_tsdf_raycast": (c_int, [c_void_p, c_void_p, C.POINTER(TsdfParams), C.POINTER(NuiCLCameraParams),
                                  c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "nuif_tsdf_raycast_gated": (c_int, [c_void_p, c_void_p, C.POINTER(TsdfParams), C.POINTER(NuiCLCameraParams),
                                        c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                        c_void_p]),
    "nuif_tsdf_surface_blocks_bytes": (c_size_t, [C.POINTER(TsdfParams)]),
    "nuif_tsdf_surface_blocks_clear": (c_int, [c_void_p, C.POINTER(TsdfParams), c_void_p]),
    "nuif_scan_scratch_bytes": (c_size_t, [c_size_t]),
    "nuif_exclusive_scan_u32": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p, c_void_p, c_void_p]),
    "nuif_exclusive_scan_flags": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p, c_void_p, c_void_p]),
    "nuif_compact_flags": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p, c_void_p]),
    "nuif_mc_scratch_bytes": (c_size_t, [C.POINTER(TsdfParams)]),
    "nuif_mc_dense": (c_int, [c_void_p, c_void_p, C.POINTER(TsdfParams), c_int, c_void_p, c_void_p, C.c_uint32,
                              c_void_p, c_void_p, c_void_p]),
    "nuif_tsdf_walk_states_bytes": (c_size_t, [C.POINTER(TsdfParams)]),
    "nuif_tsdf_integrate_wrapped": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, C.POINTER(TsdfParams),
                                            C.POINTER(NuiCLCameraParams), c_void_p, c_int, c_int, c_void_p, c_void_p,
                                            c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nuif_tsdf_raycast_wrapped": (c_int, [c_void_p, c_void_p, C.POINTER(TsdfParams), C.POINTER(NuiCLCameraParams),
                                          c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                          c_void_p]),
    "nuif_tsdf_reset_slice": (c_int, [c_void_p, c_void_p, C.POINTER(TsdfParams), c_int, c_void_p, c_void_p,
                                      c_void_p]),
    "nuif_points_dense_slice": (c_int, [c_void_p, c_void_p, C.POINTER(TsdfParams), c_int, c_void_p, c_void_p,
                                        c_void_p, c_void_p, c_void_p, C.c_uint32, c_void_p, c_void_p, c_void_p]),
    "nuif_points_dense_scratch_bytes": (c_size_t, [C.POINTER(TsdfParams)]),
    "nuif_points_dense": (c_int, [c_void_p, c_void_p, C.POINTER(TsdfParams), c_int, c_void_p, c_void_p, c_void_p,
                                  C.c_uint32, c_void_p, c_void_p, c_void_p]),
    "nuif_mesh_dense_scratch_bytes": (c_size_t, [C.POINTER(TsdfParams), C.c_uint32]),
    "nuif_mesh_dense": (c_int, [c_void_p, C.POINTER(TsdfParams), c_int, C.c_uint32, c_void_p, c_void_p,
                                C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), c_void_p, c_void_p]),
    "nuif_hash_scene_bytes": (c_size_t, [C.POINTER(HashParams)]),
    "nuif_hash_layout": (c_int, [C.POINTER(HashParams), C.POINTER(c_size_t)]),
    "nuif_hash_reset": (c_int, [c_void_p, C.POINTER(HashParams), c_void_p]),
    "nuif_hash_allocate": (c_int, [c_void_p, C.POINTER(HashParams), c_void_p, C.POINTER(NuiCLCameraParams), c_void_p, c_void_p,
                                   c_void_p]),
    "nuif_hash_integrate": (c_int, [c_void_p, C.POINTER(HashParams), c_void_p, c_void_p, C.POINTER(NuiCLCameraParams), c_void_p,
                                    c_void_p, c_void_p, c_void_p]),
    "nuif_hash_raycast": (c_int, [c_void_p, C.POINTER(HashParams), C.POINTER(NuiCLCameraParams), c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p, c_void_p]),
    "nuif_color_half_sample": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "nuif_color_gradient": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "nuif_color_pyramid": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_void_p]),
    "nuif_color_difference": (c_int, [c_void_p, c_void_p, c_void_p, C.POINTER(NuiCLCameraParams), c_int, c_void_p,
                                      c_void_p, c_void_p]),
    "nuif_color_icp_state_bytes": (c_size_t, []),
    "nuif_color_icp_partials_bytes": (c_size_t, [c_int, c_int]),
    "nuif_color_icp_begin": (c_int, [c_void_p, c_void_p, c_int, c_void_p]),
    "nuif_color_icp_iteration": (c_int, [c_void_p, c_void_p, C.POINTER(NuiCLCameraParams), c_int, c_int, c_float,
                                         c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nuif_color_icp_end": (c_int, [c_void_p, c_void_p, c_void_p]),
    "nuif_color_icp_estimate": (c_int, [c_void_p, c_void_p, C.POINTER(NuiCLCameraParams), c_int, c_void_p, c_void_p,
                                        c_float, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                        c_void_p]),
    "nuif_color_icp_read": (c_int, [c_void_p, C.POINTER(ColorIcpResult), c_void_p]),
    "nuif_bgra_to_intensity": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "nuif_intensity_derivatives": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "nuif_intensity_half_sample": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "nuif_icp_photometric_inputs": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, C.POINTER(IcpPhotometric),
                                            c_void_p]),
    "nuif_icp_estimate_joint": (c_int, [c_void_p, c_void_p, C.POINTER(NuiCLCameraParams), c_void_p, c_void_p,
                                        c_void_p, c_void_p, C.POINTER(IcpConfig), C.POINTER(IcpPhotometric),
                                        c_void_p, c_int, c_void_p]),
    "nuif_intensity_icp_workspace_bytes": (c_size_t, [c_int, c_int]),
    "nuif_intensity_icp_sums": (c_int, [C.POINTER(NuiCLCameraParams), c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                        c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, C.POINTER(c_float),
                                        C.POINTER(c_float), c_void_p, c_void_p, c_void_p]),
    "nuif_hash_forward_render_bytes": (c_size_t, [c_int, c_int]),
    "nuif_hash_forward_render": (c_int, [c_void_p, C.POINTER(HashParams), C.POINTER(NuiCLCameraParams), c_void_p,
                                         c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p,
                                         c_void_p]),
    "nuif_hash_mesh_scratch_bytes": (c_size_t, [C.POINTER(HashParams)]),
    "nuif_hash_mesh": (c_int, [c_void_p, C.POINTER(HashParams), c_void_p, c_void_p, C.c_uint32, c_void_p, c_void_p, c_void_p]),
    "nuif_hash_swap_layout": (c_int, [C.POINTER(HashParams), C.POINTER(c_size_t)]),
    "nuif_hash_swap_prepare": (c_int, [c_void_p, C.POINTER(HashParams), c_void_p, c_void_p]),
    "nuif_hash_swap_scratch_bytes": (c_size_t, [C.POINTER(HashParams)]),
    "nuif_hash_swap_list": (c_int, [c_void_p, C.POINTER(HashParams), c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nuif_hash_swap_in": (c_int, [c_void_p, C.POINTER(HashParams), c_void_p, C.c_uint32, c_void_p, c_void_p,
                                  c_void_p]),
    "nuif_hash_swap_out": (c_int, [c_void_p, C.POINTER(HashParams), c_void_p, C.c_uint32, c_void_p, c_void_p,
                                   c_void_p]),
    "nuif_hash_points_scratch_bytes": (c_size_t, [C.POINTER(HashParams)]),
    "nuif_hash_points": (c_int, [c_void_p, C.POINTER(HashParams), c_void_p, c_void_p, C.c_uint32, c_void_p, c_void_p,
                                 c_void_p]),
    "nuif_hash_stats": (c_int, [c_void_p, C.POINTER(C.c_uint32), c_void_p]),
}

class LoadedFrame(C.Structure):
    """nuif_loaded_frame (include/nuif_io.h)."""
    _fields_ = [("frame_id", C.c_uint32), ("parts", C.c_uint32), ("slot", C.c_int32), ("pad", C.c_uint32),
                ("timestamp", C.c_int64 * 4), ("width", C.c_uint32 * 4), ("height", C.c_uint32 * 4),
                ("depth", c_void_p), ("color", c_void_p), ("cmap", c_void_p), ("vertex", c_void_p),
                ("device_depth", c_void_p), ("device_color", c_void_p),
                ("fx", c_float), ("fy", c_float), ("cx", c_float), ("cy", c_float)]


FRAME_DEPTH, FRAME_COLOR, FRAME_CMAP, FRAME_VERTEX, FRAME_CAMERA = 1, 2, 4, 8, 16
ERR_IO = -6

_KSIGS = {
    "nuifk_last_error": (C.c_char_p, []),
    "nuifk_create": (c_int, [C.POINTER(KinfuDesc), C.POINTER(c_void_p)]),
    "nuifk_destroy": (c_int, [c_void_p]),
    "nuifk_run": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int]),
    "nuifk_synchronize": (c_int, [c_void_p]),
    "nuifk_reset": (c_int, [c_void_p]),
    "nuifk_poses": (c_int, [c_void_p, c_void_p, c_int]),
    "nuifk_set_history_capacity": (c_int, [c_void_p, c_int]),
    "nuifk_state": (c_int, [c_void_p, C.POINTER(IcpResult)]),
    "nuifk_read_volume": (c_int, [c_void_p, c_void_p, c_size_t]),
    "nuifk_read_maps": (c_int, [c_void_p, c_void_p, c_void_p]),
    "nuifk_read_feedback_colors": (c_int, [c_void_p, c_void_p]),
    "nuifk_set_photometric": (c_int, [c_void_p, c_float, c_float, c_float]),
    "nuifk_stream": (c_void_p, [c_void_p]),
    "nuifk_volume_ptr": (c_void_p, [c_void_p]),
    "nuifk_volume_bytes": (c_size_t, [c_void_p]),
    "nuifk_trunc": (c_float, [c_void_p]),
    "nuifk_set_stage_timing": (c_int, [c_void_p, c_int]),
    "nuifk_take_stage_times": (c_int, [c_void_p, C.POINTER(C.c_double)]),
    "nuifk_updated_voxels": (c_int, [c_void_p, C.POINTER(C.c_ulonglong)]),
    "nuifk_load_tracker_config": (c_int, [C.c_char_p, C.POINTER(KinfuDesc)]),
    "nuifk_save_tracker_config": (c_int, [C.c_char_p, C.POINTER(KinfuDesc)]),
    "nuifk_shift_state": (c_int, [c_void_p, c_void_p, c_void_p]),
    "nuifk_shift_cache": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "nuifk_hash_stored_blocks": (c_int, [c_void_p, C.POINTER(c_size_t)]),
    "nuifs_create": (c_int, [c_size_t, C.POINTER(c_void_p)]),
    "nuifs_destroy": (c_int, [c_void_p]),
    "nuifs_session_bytes": (c_size_t, [C.POINTER(KinfuDesc)]),
    "nuifs_open": (c_int, [c_void_p, C.POINTER(KinfuDesc)]),
    "nuifs_close": (c_int, [c_void_p, c_int]),
    "nuifs_bytes_in_use": (c_size_t, [c_void_p]),
    "nuifs_session_count": (c_int, [c_void_p]),
    "nuifs_engine": (c_void_p, [c_void_p, c_int]),
    "nuifs_submit": (c_int, [c_void_p, c_int, C.POINTER(c_int), C.POINTER(c_void_p)]),
    "nuifs_submit_color": (c_int, [c_void_p, c_int, C.POINTER(c_int), C.POINTER(c_void_p), C.POINTER(c_void_p)]),
    "nuifs_set_round_graph": (c_int, [c_void_p, c_int]),
    "nuifs_wait": (c_int, [c_void_p, C.POINTER(c_float)]),
    "nuifs_set_failure_policy": (c_int, [c_void_p, c_int]),
    "nuifs_set_stage_timing": (c_int, [c_void_p, c_int]),
    "nuifs_take_stage_times": (c_int, [c_void_p, c_void_p]),
    "nuifk_raycast_footprint": (c_int, [c_void_p, c_void_p]),
    "nuifk_mesh_shape": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p, c_size_t, c_void_p, c_void_p]),
    "nuifk_map_color": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int]),
    "nuifk_frame_colors": (c_int, [c_void_p, c_void_p]),
    "nuifs_last_outcomes": (c_int, [c_void_p, c_void_p, c_int]),
    "nuifk_scene_failures": (c_int, [c_void_p, c_void_p]),
    "nuifk_hash_stats": (c_int, [c_void_p, c_void_p]),
    "nuifs_set_batching": (c_int, [c_void_p, c_int]),
    "nuifs_last_batch": (c_int, [c_void_p]),
    "nuif_io_write_frame_image": (c_int, [C.c_char_p, c_int, C.c_int64, C.c_uint32, C.c_uint32, c_void_p,
                                          C.c_uint32]),
    "nuif_io_read_frame_header": (c_int, [C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_uint32),
                                          C.POINTER(C.c_uint32)]),
    "nuif_io_read_frame_buffer": (c_int, [C.c_char_p, c_void_p, C.c_uint32]),
    "nuif_io_write_camera": (c_int, [C.c_char_p, c_float, c_float, c_float, c_float]),
    "nuif_io_read_camera": (c_int, [C.c_char_p, C.POINTER(c_float), C.POINTER(c_float), C.POINTER(c_float),
                                    C.POINTER(c_float)]),
    "nuif_frame_loader_create": (c_int, [C.c_char_p, c_int, c_int, c_int, C.POINTER(c_void_p)]),
    "nuif_frame_loader_next": (c_int, [c_void_p, C.POINTER(LoadedFrame)]),
    "nuif_frame_loader_release": (c_int, [c_void_p, c_int]),
    "nuif_frame_loader_reset": (c_int, [c_void_p]),
    "nuif_frame_loader_destroy": (c_int, [c_void_p]),
    "nuif_frame_save": (c_int, [C.c_char_p, C.POINTER(LoadedFrame), c_int]),
}

_bound = {}


def _bind(libfn, sigs):
    lib = libfn()
    key = id(lib)
    if key in _bound:
        return lib
    for name, (res, args) in sigs.items():
        if not hasattr(lib, name):
            continue
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _bound[key] = True
    return lib


def L():
    return _bind(libnuif, _SIGS)


def K():
    L()
    return _bind(libkinfu, _KSIGS)


def declared_symbols():
    """Names of every function declared in include/nuif.h that this binding knows."""
    return sorted(_SIGS)


def kcheck(rc, what=""):
    """check() for the libnuif_kinfu entry points (message from nuifk_last_error)."""
    if rc < 0:
        msg = K().nuifk_last_error()
        raise NuifError(f"{what}: status {rc}: {msg.decode() if msg else ''}")
    return rc


def declared_kinfu_symbols():
    """Names of every function declared in include/nuif_kinfu.h that this binding knows."""
    return sorted(_KSIGS)


class NuifError(RuntimeError):
    pass


def check(rc, what=""):
    if rc != 0:
        msg = L().nuif_last_error()
        raise NuifError(f"{what}: status {rc}: {msg.decode() if msg else ''}")
    return rc


_initialised = False


def init(device=0):
    global _initialised
    if not _initialised:
        check(L().nuif_init(device), "nuif_init")
        _initialised = True


def device_info():
    init()
    name = C.create_string_buffer(256)
    cus, hbm, clk = c_int(), c_size_t(), c_int()
    check(L().nuif_device_info(name, 256, C.byref(cus), C.byref(hbm), C.byref(clk)), "nuif_device_info")
    return {"name": name.value.decode(), "cu": cus.value, "hbm_bytes": hbm.value, "clock_khz": clk.value}


class DeviceArray:
    """A device allocation with a numpy shape/dtype for transfers."""

    def __init__(self, shape, dtype):
        init()
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (shape,)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = c_void_p()
        check(L().nuif_malloc(C.byref(p), max(self.nbytes, 1)), "nuif_malloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, arr, stream=None):
        arr = np.ascontiguousarray(arr)
        d = cls(arr.shape, arr.dtype)
        d.upload(arr, stream)
        return d

    def upload(self, arr, stream=None):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes == self.nbytes, (arr.nbytes, self.nbytes)
        check(L().nuif_memcpy_h2d(self.ptr, arr.ctypes.data, self.nbytes, stream), "h2d")
        check(L().nuif_stream_synchronize(stream), "sync")

    def download(self, stream=None):
        out = np.empty(self.shape, self.dtype)
        check(L().nuif_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes, stream), "d2h")
        check(L().nuif_stream_synchronize(stream), "sync")
        return out

    def free(self):
        if self.ptr:
            L().nuif_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def camera_params(W, H, fx, fy, cx, cy, dmin=0.5, dmax=4.5):
    """NuiCLCameraParams as OCL/NuiKinfuOpenCLCameraState.cpp:33-62 fills it."""
    c = NuiCLCameraParams()
    c.fx = fx
    c.fy = fy
    c.fx_inv = np.float32(1.0) / np.float32(fx)
    c.fy_inv = np.float32(1.0) / np.float32(fy)
    c.cx = cx
    c.cy = cy
    c.depthImageWidth = W
    c.depthImageHeight = H
    c.sensorDepthWorldMin = dmin
    c.sensorDepthWorldMax = dmax
    return c


def inverse3_f32(R):
    """Eigen-style float cofactor inverse of a row-major 3x3 (matches the device helper)."""
    m = np.asarray(R, np.float32).reshape(9)
    f = np.float32
    c00 = f(m[4] * m[8]) - f(m[5] * m[7])
    c01 = f(m[5] * m[6]) - f(m[3] * m[8])
    c02 = f(m[3] * m[7]) - f(m[4] * m[6])
    det = f(f(f(m[0] * c00) + f(m[1] * c01)) + f(m[2] * c02))
    inv = f(1.0) / det
    out = np.empty(9, np.float32)
    out[0] = c00 * inv
    out[3] = c01 * inv
    out[6] = c02 * inv
    out[1] = f(f(m[2] * m[7]) - f(m[1] * m[8])) * inv
    out[4] = f(f(m[0] * m[8]) - f(m[2] * m[6])) * inv
    out[7] = f(f(m[1] * m[6]) - f(m[0] * m[7])) * inv
    out[2] = f(f(m[1] * m[5]) - f(m[2] * m[4])) * inv
    out[5] = f(f(m[2] * m[3]) - f(m[0] * m[5])) * inv
    out[8] = f(f(m[0] * m[4]) - f(m[1] * m[3])) * inv
    return out


def rigid(R, t):
    g = NuiCLRigidTransform()
    R = np.asarray(R, np.float32).reshape(9)
    Ri = inverse3_f32(R)
    for k in range(9):
        g.R[k] = R[k]
        g.R_inv[k] = Ri[k]
    for k in range(3):
        g.t[k] = float(np.float32(t[k]))
    return g


def upload_struct(st):
    d = DeviceArray((C.sizeof(st),), np.uint8)
    check(L().nuif_memcpy_h2d(d.ptr, C.addressof(st), C.sizeof(st), None), "h2d")
    check(L().nuif_stream_synchronize(None), "sync")
    return d


def tsdf_params(N, dim, trunc):
    p = TsdfParams()
    check(L().nuif_tsdf_params(N, N, N, dim, dim, dim, trunc, C.byref(p)), "nuif_tsdf_params")
    return p


def points_dense(volume, params, storage, color_volume=None, surface_blocks=None, capacity=None, stream=None):
    """nuif_points_dense over device arrays: (points float32[n,3], colours float32[n,4], total)
    in (y, x, z) column order; capacity None sizes the output from a count pass."""
    scratch = DeviceArray((L().nuif_points_dense_scratch_bytes(C.byref(params)),), np.uint8)
    tot = DeviceArray((1,), np.uint32)
    cv = color_volume.ptr if color_volume is not None else None
    sb = surface_blocks.ptr if surface_blocks is not None else None
    if capacity is None:
        check(L().nuif_points_dense(volume.ptr, cv, C.byref(params), storage, sb, None, None, 0, tot.ptr,
                                    scratch.ptr, stream), "nuif_points_dense")
        capacity = int(tot.download()[0])
    v = DeviceArray((max(capacity, 1), 3), np.float32)
    c = DeviceArray((max(capacity, 1), 4), np.float32)
    check(L().nuif_points_dense(volume.ptr, cv, C.byref(params), storage, sb, v.ptr, c.ptr, capacity, tot.ptr,
                                scratch.ptr, stream), "nuif_points_dense")
    total = int(tot.download()[0])
    n = min(capacity, total)
    return v.download()[:n], c.download()[:n], total


class Event:
    def __init__(self):
        init()
        e = c_void_p()
        check(L().nuif_event_create(C.byref(e)), "event")
        self.ptr = e.value

    def record(self, stream=None):
        check(L().nuif_event_record(self.ptr, stream), "event record")

    def synchronize(self):
        check(L().nuif_event_synchronize(self.ptr), "event sync")

    def elapsed_ms(self, later):
        ms = c_float()
        check(L().nuif_event_elapsed_ms(self.ptr, later.ptr, C.byref(ms)), "elapsed")
        return ms.value


class Engine:
    """One tracking session: NuiKinfuTrackingEngine + dense NuiKinfuHIPScene."""

    def __init__(self, desc, handle=None):
        init()
        K()
        self.desc = desc
        self._owned = handle is None
        if handle is not None:  # a session owned by a Server
            self._h = c_void_p(handle)
            return
        self._h = c_void_p()
        rc = K().nuifk_create(C.byref(desc), C.byref(self._h))
        if rc != 0:
            raise NuifError("nuifk_create: " + (K().nuifk_last_error() or b"").decode())

    def _check(self, rc, what):
        if rc < 0:
            raise NuifError(f"{what}: " + (K().nuifk_last_error() or b"").decode())
        return rc

    def run(self, depth, bgra=None, sync=True):
        """depth: numpy uint16 (host) or DeviceArray (device)."""
        if isinstance(depth, DeviceArray):
            bp = bgra.ptr if isinstance(bgra, DeviceArray) else None
            return self._check(K().nuifk_run(self._h, depth.ptr, 1, bp, 1 if sync else 0), "run")
        depth = np.ascontiguousarray(depth, np.uint16)
        bp = np.ascontiguousarray(bgra, np.uint8).ctypes.data if bgra is not None else None
        return self._check(K().nuifk_run(self._h, depth.ctypes.data, 0, bp, 1), "run")

    def synchronize(self):
        return self._check(K().nuifk_synchronize(self._h), "synchronize")

    def shift_state(self):
        """Shifting volume: (voxel offset int32[3], its size in metres float32[3])."""
        o = np.zeros(3, np.int32)
        m = np.zeros(3, np.float32)
        self._check(K().nuifk_shift_state(self._h, o.ctypes.data, m.ctypes.data), "shift state")
        return o, m

    def shift_cache(self):
        """Shifting volume: points (float32[n,3]) and colours (float32[n,4]) of the evicted slabs."""
        n = self._check(K().nuifk_shift_cache(self._h, None, None, 0), "shift cache")
        p = np.empty((max(n, 1), 3), np.float32)
        c = np.empty((max(n, 1), 4), np.float32)
        self._check(K().nuifk_shift_cache(self._h, p.ctypes.data, c.ctypes.data, n), "shift cache")
        return p[:n], c[:n]

    def mesh_shape(self):
        """Volume2Mesh(NuiMeshShape*): (points float32[n,3], triangle indices int32[m])."""
        npts, nidx = c_size_t(), c_size_t()
        self._check(K().nuifk_mesh_shape(self._h, None, 0, None, 0, C.byref(npts), C.byref(nidx)), "mesh shape")
        pts = np.empty((max(npts.value, 1), 3), np.float32)
        idx = np.empty(max(nidx.value, 1), np.int32)
        self._check(K().nuifk_mesh_shape(self._h, pts.ctypes.data, npts.value, idx.ctypes.data, nidx.value,
                                         C.byref(npts), C.byref(nidx)), "mesh shape")
        return pts[:npts.value], idx[:nidx.value]

    def map_color(self, map_xy, image, lut=None):
        """The frame's colours from `image` (HxWx4 BGRA) through the depth-to-colour map."""
        m = np.ascontiguousarray(map_xy, np.float32)
        img = np.ascontiguousarray(image, np.uint8)
        lp = np.ascontiguousarray(lut, np.uint32).ctypes.data if lut is not None else None
        self._check(K().nuifk_map_color(self._h, m.ctypes.data, lp, img.ctypes.data, img.shape[1], img.shape[0]),
                    "map colour")

    def frame_colors(self):
        out = np.empty((self.desc.height, self.desc.width, 4), np.uint8)
        self._check(K().nuifk_frame_colors(self._h, out.ctypes.data), "frame colours")
        return out

    def raycast_footprint(self):
        """Distinct voxels the dense raycast reads at the current pose (U of B*U + 28*W*H)."""
        out = C.c_ulonglong()
        self._check(K().nuifk_raycast_footprint(self._h, C.byref(out)), "raycast footprint")
        return out.value

    def scene_failures(self):
        """Hash scene: insertions dropped for want of a brick or a bucket slot (0 if dense)."""
        out = C.c_ulonglong()
        self._check(K().nuifk_scene_failures(self._h, C.byref(out)), "scene failures")
        return out.value

    def hash_stats(self):
        """Hash scene: (bricks handed out, visible blocks of the last frame, failed
        insertions, occupied entries); zeros for a dense scene."""
        out = (C.c_uint32 * 4)()
        self._check(K().nuifk_hash_stats(self._h, out), "hash stats")
        return tuple(int(x) for x in out)

    def set_history_capacity(self, capacity):
        """Rows of the device pose ring; rows are copied to the host before it wraps."""
        self._check(K().nuifk_set_history_capacity(self._h, capacity), "history capacity")

    def poses(self, n):
        out = np.zeros((n, POSE_ROW), np.float32)
        m = self._check(K().nuifk_poses(self._h, out.ctypes.data, n), "poses")
        return out[:m]

    def state(self):
        r = IcpResult()
        self._check(K().nuifk_state(self._h, C.byref(r)), "state")
        return r

    def volume_raw(self):
        nb = K().nuifk_volume_bytes(self._h)
        out = np.empty(nb, np.uint8)
        self._check(K().nuifk_read_volume(self._h, out.ctypes.data, nb), "read volume")
        return out

    def maps(self):
        n = self.desc.width * self.desc.height
        v = np.empty((n, 3), np.float32)
        nm = np.empty((n, 3), np.float32)
        self._check(K().nuifk_read_maps(self._h, v.ctypes.data, nm.ctypes.data), "maps")
        return v, nm

    def set_photometric(self, weight, max_difference=0.2, min_gradient=0.002):
        """The joint ICP order's photometric weight, intensity difference and minimum gradient."""
        self._check(K().nuifk_set_photometric(self._h, weight, max_difference, min_gradient), "photometric")

    def feedback_colors(self):
        """The feedback frame's BGRA colours (H*W x 4 uint8)."""
        out = np.empty((self.desc.width * self.desc.height, 4), np.uint8)
        self._check(K().nuifk_read_feedback_colors(self._h, out.ctypes.data), "feedback colours")
        return out

    def set_stage_timing(self, on):
        self._check(K().nuifk_set_stage_timing(self._h, 1 if on else 0), "stage timing")

    def take_stage_times(self):
        """(frames, [preprocess, icp, integrate, raycast] summed ms)."""
        ms = (C.c_double * 4)()
        n = self._check(K().nuifk_take_stage_times(self._h, ms), "stage times")
        return n, list(ms)

    def updated_voxels(self):
        v = C.c_ulonglong()
        self._check(K().nuifk_updated_voxels(self._h, C.byref(v)), "updated voxels")
        return v.value

    @property
    def stream(self):
        return K().nuifk_stream(self._h)

    @property
    def trunc(self):
        return K().nuifk_trunc(self._h)

    def close(self):
        if self._h and self._owned:
            K().nuifk_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def decode_volume(raw, storage):
    """(tsdf float32 normalised, weight int32) arrays from a raw volume dump."""
    if storage == TSDF_S16:
        v = raw.view(np.int16).reshape(-1, 2)
        return v[:, 0].astype(np.float32) / np.float32(32766.0), v[:, 1].astype(np.int32)
    if storage == TSDF_F16:
        v = raw.view(np.uint16).reshape(-1, 2)
        return v[:, 0].view(np.float16).astype(np.float32), v[:, 1].astype(np.int32)
    v = raw.view(np.float32).reshape(-1, 2)
    return v[:, 0].copy(), v[:, 1].astype(np.int32)


class Server:
    """Multi-session fusion server (nuifs_*): sessions resident on the device,
    one frame per session per round, each session on its own stream."""

    def __init__(self, budget_bytes):
        init()
        self._h = c_void_p()
        self._check(K().nuifs_create(c_size_t(int(budget_bytes)), C.byref(self._h)), "nuifs_create")
        self.sessions = {}

    @staticmethod
    def _check(rc, what):
        if rc < 0:
            raise NuifError(f"{what}: status {rc}: " + (K().nuifk_last_error() or b"").decode())
        return rc

    @staticmethod
    def session_bytes(desc):
        return K().nuifs_session_bytes(C.byref(desc))

    def open(self, desc):
        """Admit a session; returns its id, or None when it does not fit the budget."""
        sid = K().nuifs_open(self._h, C.byref(desc))
        if sid == -3:
            return None
        self._check(sid, "nuifs_open")
        self.sessions[sid] = Engine(desc, handle=K().nuifs_engine(self._h, sid))
        return sid

    def close_session(self, sid):
        self.sessions.pop(sid)._h = None
        self._check(K().nuifs_close(self._h, sid), "nuifs_close")

    @property
    def bytes_in_use(self):
        return K().nuifs_bytes_in_use(self._h)

    def submit(self, ids, frames, colors=None):
        """ids: session ids; frames: DeviceArray uint16 depth per id; colors: None or a
        DeviceArray of BGRA bytes (or None) per id."""
        n = len(ids)
        self._ids = (c_int * n)(*ids)
        self._ptrs = (c_void_p * n)(*[f.ptr for f in frames])
        self._n = n
        if colors is None:
            self._check(K().nuifs_submit(self._h, n, self._ids, self._ptrs), "nuifs_submit")
            return
        self._cptrs = (c_void_p * n)(*[c.ptr if c is not None else None for c in colors])
        self._check(K().nuifs_submit_color(self._h, n, self._ids, self._ptrs, self._cptrs), "nuifs_submit_color")

    def set_round_graph(self, on):
        """Batched rounds replayed from a captured graph (default) or enqueued launch by launch."""
        self._check(K().nuifs_set_round_graph(self._h, 1 if on else 0), "nuifs_set_round_graph")

    def set_batching(self, on):
        """Batched rounds (nuif_fusion_round) on or off."""
        self._check(K().nuifs_set_batching(self._h, 1 if on else 0), "nuifs_set_batching")

    @property
    def last_batch(self):
        return K().nuifs_last_batch(self._h)

    def wait(self):
        """Waits for the round; returns per-frame latency (ms, round start -> frame end).
        The failure policy has then been applied (evicted sessions leave self.sessions)."""
        lat = (c_float * max(self._n, 1))()
        self._check(K().nuifs_wait(self._h, lat), "nuifs_wait")
        if getattr(self, "_policy", ON_FAILURE_KEEP) == ON_FAILURE_EVICT:
            for sid, ok in zip(list(self._ids), self.last_outcomes()):
                if not ok and sid in self.sessions:
                    self.sessions.pop(sid)._h = None
        return np.array(lat[:self._n], np.float64)

    def set_stage_timing(self, on):
        self._check(K().nuifs_set_stage_timing(self._h, 1 if on else 0), "nuifs_set_stage_timing")

    def take_stage_times(self):
        """(rounds, [preprocess, bounds, icp, integrate, raycast] summed ms) of the batched rounds."""
        ms = (C.c_double * 5)()
        n = self._check(K().nuifs_take_stage_times(self._h, ms), "nuifs_take_stage_times")
        return n, list(ms)

    def set_failure_policy(self, policy):
        """ON_FAILURE_KEEP / ON_FAILURE_RESET / ON_FAILURE_EVICT."""
        self._check(K().nuifs_set_failure_policy(self._h, policy), "nuifs_set_failure_policy")
        self._policy = policy

    def last_outcomes(self):
        """Tracking outcome of each frame of the last waited round (1 tracked, 0 failed)."""
        n = self._check(K().nuifs_last_outcomes(self._h, None, 0), "nuifs_last_outcomes")
        out = (C.c_int32 * max(n, 1))()
        self._check(K().nuifs_last_outcomes(self._h, out, n), "nuifs_last_outcomes")
        return np.array(out[:n], np.int32)

    def close(self):
        if self._h:
            for e in self.sessions.values():
                e._h = None
            self.sessions = {}
            K().nuifs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass



class HashScene:
    """A voxel-hashing scene in device memory (nuif_hash_*), driven with explicit poses."""

    ENTRY = np.dtype([("key", "<u8"), ("ptr", "<i4"), ("pad", "<u4")])

    def __init__(self, params, stream=None):
        init()
        self.p = params
        self.stream = stream
        nb = L().nuif_hash_scene_bytes(C.byref(params))
        if nb == 0:
            raise NuifError("nuif_hash_scene_bytes: bad parameters: " + L().nuif_last_error().decode())
        self.mem = DeviceArray((nb,), np.uint8)
        off = (c_size_t * 8)()
        check(L().nuif_hash_layout(C.byref(params), off), "nuif_hash_layout")
        self.offsets = list(off)
        check(L().nuif_hash_reset(self.mem.ptr, C.byref(params), stream), "nuif_hash_reset")

    def allocate(self, depth, cam, pose, gate=None):
        check(L().nuif_hash_allocate(self.mem.ptr, C.byref(self.p), depth.ptr, C.byref(cam), pose.ptr,
                                     gate, self.stream), "nuif_hash_allocate")

    def integrate(self, depth, cam, pose, bgra=None, gate=None, updated=None):
        check(L().nuif_hash_integrate(self.mem.ptr, C.byref(self.p), depth.ptr, bgra.ptr if bgra else None,
                                      C.byref(cam), pose.ptr, gate, updated, self.stream), "nuif_hash_integrate")

    def raycast(self, cam, pose, color=False):
        n = cam.depthImageWidth * cam.depthImageHeight
        rg = DeviceArray((n, 2), np.float32)
        v = DeviceArray((n, 3), np.float32)
        nm = DeviceArray((n, 3), np.float32)
        cm = DeviceArray((n, 4), np.uint8) if color else None
        check(L().nuif_hash_raycast(self.mem.ptr, C.byref(self.p), C.byref(cam), pose.ptr, rg.ptr, v.ptr, nm.ptr,
                                    cm.ptr if cm else None, None, self.stream), "nuif_hash_raycast")
        return v.download(), nm.download(), rg.download(), cm.download() if cm else None

    def forward_render(self, cam, pose, vmap, cmap=None, reference_transform=False):
        """Forward render from the previous world vertex map `vmap` (and colours `cmap`)
        at `pose`: returns (vmap, nmap, range, cmap) as host arrays."""
        W, H = cam.depthImageWidth, cam.depthImageHeight
        n = W * H
        rg = DeviceArray((n, 2), np.float32)
        v = DeviceArray.from_host(np.ascontiguousarray(vmap, np.float32).reshape(n, 3))
        nm = DeviceArray((n, 3), np.float32)
        cm = DeviceArray.from_host(np.ascontiguousarray(cmap, np.uint8).reshape(n, 4)) if cmap is not None else None
        work = DeviceArray((L().nuif_hash_forward_render_bytes(W, H),), np.uint8)
        check(L().nuif_hash_forward_render(self.mem.ptr, C.byref(self.p), C.byref(cam), pose.ptr, rg.ptr, v.ptr,
                                           nm.ptr, cm.ptr if cm else None, work.ptr, int(bool(reference_transform)),
                                           None, self.stream), "nuif_hash_forward_render")
        return v.download(), nm.download(), rg.download(), cm.download() if cm else None

    def mesh(self, capacity=None, colors=False):
        scratch = DeviceArray((L().nuif_hash_mesh_scratch_bytes(C.byref(self.p)),), np.uint8)
        tot = DeviceArray((1,), np.uint32)
        check(L().nuif_hash_mesh(self.mem.ptr, C.byref(self.p), None, None, 0, tot.ptr, scratch.ptr, self.stream),
              "nuif_hash_mesh")
        n = int(tot.download()[0]) if capacity is None else capacity
        v = DeviceArray((max(n, 1), 3), np.float32)
        c = DeviceArray((max(n, 1), 4), np.float32) if colors else None
        check(L().nuif_hash_mesh(self.mem.ptr, C.byref(self.p), v.ptr, c.ptr if c else None, n, tot.ptr,
                                 scratch.ptr, self.stream), "nuif_hash_mesh")
        total = int(tot.download()[0])
        return v.download()[:min(n, total)], (c.download()[:min(n, total)] if c else None), total

    def points(self, capacity=None):
        """nuif_hash_points: (points float32[n,3], colours float32[n,4], total) in visible-list order."""
        scratch = DeviceArray((L().nuif_hash_points_scratch_bytes(C.byref(self.p)),), np.uint8)
        tot = DeviceArray((1,), np.uint32)
        if capacity is None:
            check(L().nuif_hash_points(self.mem.ptr, C.byref(self.p), None, None, 0, tot.ptr, scratch.ptr,
                                       self.stream), "nuif_hash_points")
            capacity = int(tot.download()[0])
        v = DeviceArray((max(capacity, 1), 3), np.float32)
        c = DeviceArray((max(capacity, 1), 4), np.float32)
        check(L().nuif_hash_points(self.mem.ptr, C.byref(self.p), v.ptr, c.ptr, capacity, tot.ptr, scratch.ptr,
                                   self.stream), "nuif_hash_points")
        total = int(tot.download()[0])
        n = min(capacity, total)
        return v.download()[:n], c.download()[:n], total

    def visible_blocks(self):
        """(pos int32[n,3], brick ptr int32[n]) of the visible list in list order, the raw
        brick words uint32[bricks*512] and the voxel colours uint8[bricks*512*4] or None."""
        raw = self.mem.download()
        nvis = int(raw[4:8].view(np.uint32)[0])
        ids = raw[self.offsets[3]:self.offsets[3] + nvis * 4].view(np.int32)
        ne = self.offsets[6]
        ent = raw[self.offsets[1]:self.offsets[1] + ne * 16].view(self.ENTRY)[ids]
        k = ent["key"]
        pos = np.stack([((k >> 32) & 0xFFFF).astype(np.uint16).view(np.int16),
                        ((k >> 16) & 0xFFFF).astype(np.uint16).view(np.int16),
                        (k & 0xFFFF).astype(np.uint16).view(np.int16)], 1).astype(np.int32)
        nbr = self.p.bricks
        words = raw[self.offsets[4]:self.offsets[4] + nbr * 512 * 4].view(np.uint32).copy()
        cols = raw[self.offsets[5]:self.offsets[5] + nbr * 512 * 4].copy() if self.p.color else None
        return pos, ent["ptr"].astype(np.int32), words, cols

    def swap_prepare(self, gate=None):
        check(L().nuif_hash_swap_prepare(self.mem.ptr, C.byref(self.p), gate, self.stream), "nuif_hash_swap_prepare")

    def swap_cycle(self, store):
        """The host side of one frame's swapping after integration (the role of
        NuiKinfuOpenCLHashGlobalCache): stored blocks in, unseen blocks out.  store: dict entry
        id -> (words uint32[512], colours uint8[512,4] or None), updated in place.  Returns the
        (swapped in, swapped out) entry ids."""
        ne = self.offsets[6]
        if not hasattr(self, "_swap_scratch"):
            self._swap_scratch = DeviceArray((L().nuif_hash_swap_scratch_bytes(C.byref(self.p)),), np.uint8)
            self._swap_ids = DeviceArray((ne,), np.int32)
            self._swap_cnt = DeviceArray((1,), np.uint32)
        moved = []
        for which in (0, 1):
            check(L().nuif_hash_swap_list(self.mem.ptr, C.byref(self.p), which, self._swap_ids.ptr,
                                          self._swap_cnt.ptr, self._swap_scratch.ptr, self.stream), "swap list")
            n = int(self._swap_cnt.download()[0])
            ids = self._swap_ids.download()[:n].copy()
            moved.append(ids)
            if n == 0:
                continue
            words = np.zeros((n, 512), np.uint32)
            cols = np.zeros((n, 512, 4), np.uint8) if self.p.color else None
            if which == 0:
                for i, e in enumerate(ids):
                    if int(e) in store:
                        words[i] = store[int(e)][0]
                        if cols is not None:
                            cols[i] = store[int(e)][1]
                dw, dc = DeviceArray.from_host(words), (DeviceArray.from_host(cols) if cols is not None else None)
                check(L().nuif_hash_swap_in(self.mem.ptr, C.byref(self.p), self._swap_ids.ptr, n, dw.ptr,
                                            dc.ptr if dc else None, self.stream), "nuif_hash_swap_in")
            else:
                dw = DeviceArray((n, 512), np.uint32)
                dc = DeviceArray((n, 512, 4), np.uint8) if cols is not None else None
                check(L().nuif_hash_swap_out(self.mem.ptr, C.byref(self.p), self._swap_ids.ptr, n, dw.ptr,
                                             dc.ptr if dc else None, self.stream), "nuif_hash_swap_out")
                w, c = dw.download(), (dc.download() if dc else None)
                for i, e in enumerate(ids):
                    store[int(e)] = (w[i].copy(), None if c is None else c[i].copy())
        return moved[0], moved[1]

    def swap_states(self):
        """(keys int16[n,3] sorted by (x,y,z), state uint8[n]) of every occupied entry."""
        raw = self.mem.download()
        off = (c_size_t * 2)()
        check(L().nuif_hash_swap_layout(C.byref(self.p), off), "nuif_hash_swap_layout")
        ne = self.offsets[6]
        ent = raw[self.offsets[1]:self.offsets[1] + ne * 16].view(self.ENTRY)
        st = raw[off[0]:off[0] + ne]
        sel = np.nonzero(ent["key"] != 0)[0]
        k = ent["key"][sel]
        keys = np.stack([((k >> 32) & 0xFFFF).astype(np.uint16).view(np.int16),
                         ((k >> 16) & 0xFFFF).astype(np.uint16).view(np.int16),
                         (k & 0xFFFF).astype(np.uint16).view(np.int16)], 1)
        order = np.lexsort((keys[:, 2], keys[:, 1], keys[:, 0]))
        return keys[order], st[sel][order]

    def entry_keys(self, ids):
        """int16[n,3] keys of entry ids."""
        raw = self.mem.download()
        ne = self.offsets[6]
        k = raw[self.offsets[1]:self.offsets[1] + ne * 16].view(self.ENTRY)["key"][np.asarray(ids, np.int64)]
        return np.stack([((k >> 32) & 0xFFFF).astype(np.uint16).view(np.int16),
                         ((k >> 16) & 0xFFFF).astype(np.uint16).view(np.int16),
                         (k & 0xFFFF).astype(np.uint16).view(np.int16)], 1)

    def stats(self):
        out = (C.c_uint32 * 4)()
        check(L().nuif_hash_stats(self.mem.ptr, out, self.stream), "nuif_hash_stats")
        return {"bricks": out[0], "visible": out[1], "failed": out[2], "entries": out[3]}

    def _region(self, start, count, dtype):
        raw = self.mem.download()
        return raw[start:start + count * np.dtype(dtype).itemsize].view(dtype)

    def blocks(self):
        """(keys int16[n,3] sorted by (x,y,z), sdf float32[n,512] (decoded), weight[n,512], raw words[n,512],
        rgb uint8[n,512,3] or None) of every entry that holds a brick."""
        return decode_hash_blocks(self.mem.download(), self.p)

    def visible_keys(self):
        """Keys of the visible list (int16[n,3], sorted by (x,y,z))."""
        raw = self.mem.download()
        nvis = int(raw[4:8].view(np.uint32)[0])
        ids = raw[self.offsets[3]:self.offsets[3] + nvis * 4].view(np.int32)
        ne = self.offsets[6]
        ent = raw[self.offsets[1]:self.offsets[1] + ne * 16].view(self.ENTRY)
        k = ent["key"][ids]
        keys = np.stack([((k >> 32) & 0xFFFF).astype(np.uint16).view(np.int16),
                         ((k >> 16) & 0xFFFF).astype(np.uint16).view(np.int16),
                         (k & 0xFFFF).astype(np.uint16).view(np.int16)], 1)
        return keys[np.lexsort((keys[:, 2], keys[:, 1], keys[:, 0]))]


def hash_params_of_desc(d):
    """The nuif_hash_params a hash session builds from its descriptor (buckets: the
    smallest 2^b with 2^b >= hash_buckets (0: 512 x 512) and 4 * 2^b >= 2 * bricks)."""
    want = d.hash_buckets if d.hash_buckets > 0 else 512 * 512
    b = 4
    while b < 24 and ((4 << b) < 2 * d.hash_bricks or (1 << b) < want):
        b += 1
    return hash_params(voxel=d.hash_voxel_size, trunc_scale=d.hash_trunc_scale, truncation=d.hash_truncation,
                       weight_max=d.hash_weight_max, buckets_log2=b, bricks=d.hash_bricks, storage=d.storage,
                       color=bool(d.color))


def decode_hash_blocks(raw, params):
    """Blocks of a downloaded hash scene: (keys int16[n,3] sorted by (x,y,z), sdf float32[n,512],
    weight int32[n,512], raw words uint32[n,512], rgb uint8[n,512,3] or None)."""
    off = (c_size_t * 8)()
    check(L().nuif_hash_layout(C.byref(params), off), "nuif_hash_layout")
    ne = off[6]
    ent = raw[off[1]:off[1] + ne * 16].view(HashScene.ENTRY)
    sel = np.nonzero((ent["key"] != 0) & (ent["ptr"] >= 0))[0]
    k = ent["key"][sel]
    keys = np.stack([((k >> 32) & 0xFFFF).astype(np.uint16).view(np.int16),
                     ((k >> 16) & 0xFFFF).astype(np.uint16).view(np.int16),
                     (k & 0xFFFF).astype(np.uint16).view(np.int16)], 1)
    order = np.lexsort((keys[:, 2], keys[:, 1], keys[:, 0]))
    keys = keys[order]
    ptr = ent["ptr"][sel][order]
    nbr = params.bricks
    words = raw[off[4]:off[4] + nbr * 512 * 4].view(np.uint32).reshape(nbr, 512)[ptr]
    lo = (words & 0xFFFF).astype(np.uint16)
    if params.storage == TSDF_S16:
        sdf = lo.view(np.int16).astype(np.float32) / np.float32(32767.0)
    else:
        sdf = lo.view(np.float16).astype(np.float32)
    weight = (words >> 16).astype(np.int32)
    rgb = None
    if params.color:
        rgb = raw[off[5]:off[5] + nbr * 512 * 4].reshape(nbr, 512, 4)[ptr][:, :, :3].copy()
    return keys, sdf, weight, words, rgb



class ColorTracker:
    """The photometric tracker (OCL/NuiKinfuOpenCLColorTracker.cpp) on nuif_color_*:
    the frame's colour pyramid and gradients, then ColorIterativeClosestPoint from a
    camera pose against a model (world vertices + colours of the previous render)."""

    def __init__(self, W, H, levels=3, corrected=False, stream=None):
        self.W, self.H, self.levels, self.corrected, self.stream = W, H, levels, bool(corrected), stream
        self.colors = [DeviceArray(((H >> l) * (W >> l) * 4,), np.uint8) for l in range(levels)]
        self.gx = [DeviceArray(((H >> l) * (W >> l) * 4,), np.uint8) for l in range(levels)]
        self.gy = [DeviceArray(((H >> l) * (W >> l) * 4,), np.uint8) for l in range(levels)]
        self.state = DeviceArray((L().nuif_color_icp_state_bytes(),), np.uint8)
        self.partials = DeviceArray((L().nuif_color_icp_partials_bytes(W, H),), np.uint8)

    @staticmethod
    def _ptrs(arrs):
        return (c_void_p * len(arrs))(*[a.ptr for a in arrs])

    def pyramid(self, frame):
        """SubSampleColors + GradientBuffers of a device BGRA frame."""
        check(L().nuif_color_pyramid(frame.ptr, self._ptrs(self.colors), self._ptrs(self.gx), self._ptrs(self.gy),
                                     self.W, self.H, self.levels, int(self.corrected), self.stream),
              "nuif_color_pyramid")

    def estimate(self, cam, vmap, colors, pose, iterations, rotation_only=None, dist_threshold=0.1):
        """Runs the levels from the device pose (updated in place when the track holds);
        returns the host view of the state."""
        its = (c_int * self.levels)(*iterations[:self.levels])
        ro = (c_int * self.levels)(*(rotation_only or [0] * self.levels)[:self.levels])
        check(L().nuif_color_icp_estimate(self.state.ptr, self.partials.ptr, C.byref(cam), self.levels, its, ro,
                                          dist_threshold, vmap.ptr, colors.ptr, self._ptrs(self.colors),
                                          self._ptrs(self.gx), self._ptrs(self.gy), pose.ptr, int(self.corrected),
                                          self.stream), "nuif_color_icp_estimate")
        r = ColorIcpResult()
        check(L().nuif_color_icp_read(self.state.ptr, C.byref(r), self.stream), "nuif_color_icp_read")
        return r

    def difference(self, cam, level, vmap, colors, pose):
        """GetColorsDifference's exact sums (sum |obs - known|^2, valid count)."""
        out = DeviceArray((2,), np.uint64)
        check(L().nuif_color_difference(vmap.ptr, colors.ptr, self.colors[level].ptr, C.byref(cam), level, pose.ptr,
                                        out.ptr, self.stream), "nuif_color_difference")
        d = out.download()
        return int(d[0]), int(d[1])


# ---- frame ingest (include/nuif_io.h) ---------------------------------------------

_PART_TYPES = [(np.uint16, 1), (np.uint8, 4), (np.float32, 2), (np.float32, 3)]


def write_frame_image(path, timestamp, buffer, compressed=True):
    """nuif_io_write_frame_image of an array [h, w(, c)]."""
    a = np.ascontiguousarray(buffer)
    h, w = a.shape[0], a.shape[1]
    kcheck(K().nuif_io_write_frame_image(os.fsencode(path), int(bool(compressed)), int(timestamp), w, h,
                                         a.ctypes.data, a.nbytes), "nuif_io_write_frame_image")


def read_frame_image(path, dtype, channels):
    """(timestamp, array[h, w, channels]) through nuif_io_read_frame_header / _buffer."""
    ts, w, h = C.c_int64(), C.c_uint32(), C.c_uint32()
    kcheck(K().nuif_io_read_frame_header(os.fsencode(path), C.byref(ts), C.byref(w), C.byref(h)),
           "nuif_io_read_frame_header")
    out = np.empty((h.value, w.value, channels), dtype)
    kcheck(K().nuif_io_read_frame_buffer(os.fsencode(path), out.ctypes.data, out.nbytes),
           "nuif_io_read_frame_buffer")
    return ts.value, out


def _part_array(ptr, w, h, k):
    dt, ch = _PART_TYPES[k]
    n = w * h * ch
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dt))), (n,)).reshape(h, w, ch)


class FrameLoader:
    """nuif_frame_loader_*: frames of a recording in id order, prefetched by `threads`
    workers into `slots` buffers (uploaded to the device when upload is set)."""

    def __init__(self, dirname, slots=4, threads=2, upload=False):
        if upload:
            init()
        self.h = c_void_p()
        kcheck(K().nuif_frame_loader_create(os.fsencode(dirname), slots, threads, int(bool(upload)),
                                            C.byref(self.h)), "nuif_frame_loader_create")

    def next_raw(self):
        """The nuif_loaded_frame of the next frame (hold it until release) or None at the end."""
        f = LoadedFrame()
        rc = kcheck(K().nuif_frame_loader_next(self.h, C.byref(f)), "nuif_frame_loader_next")
        return f if rc == 1 else None

    def next(self):
        """A dict copy of the next frame (its slot released), or None at the end."""
        f = self.next_raw()
        if f is None:
            return None
        out = {"id": f.frame_id, "parts": f.parts}
        ptrs = [f.depth, f.color, f.cmap, f.vertex]
        for k, name in enumerate(["depth", "color", "cmap", "vertex"]):
            if f.parts & (1 << k):
                out[name] = (f.timestamp[k], _part_array(ptrs[k], f.width[k], f.height[k], k).copy())
        if f.parts & FRAME_CAMERA:
            out["camera"] = (f.fx, f.fy, f.cx, f.cy)
        self.release(f.slot)
        return out

    def release(self, slot):
        kcheck(K().nuif_frame_loader_release(self.h, slot), "nuif_frame_loader_release")

    def reset(self):
        kcheck(K().nuif_frame_loader_reset(self.h), "nuif_frame_loader_reset")

    def close(self):
        if self.h:
            kcheck(K().nuif_frame_loader_destroy(self.h), "nuif_frame_loader_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
