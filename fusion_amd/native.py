"""ctypes binding of the C ABI in include/fusion_amd.h (libfusion_amd.so) and of the scene
generator in include/fusion_scene.h (libfu_scene.so).

Loading fails loudly when the HIP library has not been built: there is no fallback path.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libfusion_amd.so")
SCENE_PATH = os.path.join(LIB_DIR, "libfu_scene.so")

FU_OK = 0
FU_ERR_BUDGET = -5
FU_TRACK_RUNNING, FU_TRACK_OK, FU_TRACK_FAILED, FU_TRACK_REJECTED, FU_TRACK_FIRST = 0, 1, 2, 3, 4
FU_TRACK_LOST, FU_TRACK_RELOC = 5, 6
FU_VOXEL_F32_W8, FU_VOXEL_F32_W16, FU_VOXEL_F16_W8 = 0, 1, 2
FLAGS = {1: "heap_full", 2: "excess_full", 4: "newset_full", 8: "coord_range", 16: "render_full", 32: "tri_full"}


class FusionError(RuntimeError):
    pass


class MapParams(C.Structure):
    _fields_ = [("num_buckets", C.c_uint32), ("num_excess", C.c_uint32), ("num_blocks", C.c_uint32),
                ("voxel_size", C.c_float), ("trunc_dist", C.c_float), ("depth_min", C.c_float),
                ("depth_max", C.c_float), ("voxel_format", C.c_uint32), ("max_weight", C.c_uint32),
                ("max_render_blocks", C.c_uint32), ("max_new_blocks", C.c_uint32)]


class MapStats(C.Structure):
    _fields_ = [("n_allocated", C.c_uint32), ("n_excess", C.c_uint32), ("n_visible", C.c_uint32),
                ("n_new", C.c_uint32), ("n_fragments", C.c_uint32), ("n_triangles", C.c_uint32),
                ("flags", C.c_uint32), ("capacity", C.c_uint32)]


class PoseF(C.Structure):
    _fields_ = [("R", C.c_float * 9), ("Rinv", C.c_float * 9), ("t", C.c_float * 3)]


class TrackResult(C.Structure):
    _fields_ = [("pose", C.c_double * 16), ("last_pose", C.c_double * 16), ("x", C.c_double * 6),
                ("A", C.c_double * 36), ("b", C.c_double * 6), ("icp_error", C.c_float), ("count", C.c_int32),
                ("status", C.c_int32), ("iterations", C.c_int32), ("rgb_error", C.c_float),
                ("rgb_count", C.c_int32)]


class SystemParams(C.Structure):
    _fields_ = [("cols", C.c_int32), ("rows", C.c_int32), ("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float),
                ("cy", C.c_float), ("depth_scale", C.c_float), ("depth_cutoff", C.c_float),
                ("iters", C.c_int32 * 3), ("min_count", C.c_int32 * 3), ("icp_thresh", C.c_float),
                ("plain", C.c_int32), ("use_graph", C.c_int32), ("use_rgb", C.c_int32), ("map", MapParams)]


class KeyMapParams(C.Structure):
    _fields_ = [("grid_size", C.c_float), ("max_keys", C.c_uint32), ("n_buckets", C.c_uint32)]


class FrameResult(C.Structure):
    _fields_ = [("pose", C.c_double * 16), ("track_status", C.c_int32), ("fused", C.c_int32),
                ("icp_error", C.c_float), ("icp_count", C.c_int32), ("n_visible", C.c_uint32),
                ("n_allocated", C.c_uint32), ("n_fragments", C.c_uint32), ("flags", C.c_uint32),
                ("frame_index", C.c_uint32), ("rgb_count", C.c_int32)]


class SystemBuffers(C.Structure):
    _fields_ = [("depth_in", C.c_void_p), ("rgb_in", C.c_void_p), ("raw", C.c_void_p),
                ("depth", C.c_void_p * 3), ("vmap", C.c_void_p * 3), ("nmap", C.c_void_p * 3),
                ("model_vmap", C.c_void_p * 3), ("model_nmap", C.c_void_p * 3), ("zrange_min", C.c_void_p),
                ("zrange_max", C.c_void_p), ("zcols", C.c_int32), ("zrows", C.c_int32),
                ("image", C.c_void_p * 3), ("last_image", C.c_void_p * 3), ("dIdx", C.c_void_p * 3),
                ("dIdy", C.c_void_p * 3)]


class StageTimes(C.Structure):
    _fields_ = [("ms", C.c_double * 8), ("frames", C.c_uint32), ("n_visible", C.c_double), ("n_new", C.c_double),
                ("icp_iterations", C.c_double), ("icp_pixels", C.c_double), ("steps", C.c_double)]


STAGES = ["resize", "pyramid", "icp", "alloc", "visible", "integrate", "zrange", "raycast"]


class Tuning(C.Structure):
    _fields_ = [("raycast_wave", C.c_int32), ("xcd_order", C.c_int32), ("integrate_groups", C.c_int32),
                ("visible_groups", C.c_int32), ("alloc_groups", C.c_int32)]

P = C.c_void_p
I = C.c_int
U = C.c_uint32
F = C.c_float
FP = C.POINTER(C.c_float)
DP = C.POINTER(C.c_double)
UP = C.POINTER(C.c_uint32)
IP = C.POINTER(C.c_int32)

# name -> (restype, argtypes); every function declared in include/fusion_amd.h
class ServerParams(C.Structure):
    _fields_ = [("max_batch", C.c_int32), ("max_inflight", C.c_int32), ("lanes", C.c_int32),
                ("block_budget", C.c_uint64)]


class Completion(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("session", C.c_int32), ("track_status", C.c_int32),
                ("pose", C.c_double * 16), ("icp_error", C.c_float), ("n_visible", C.c_uint32),
                ("n_allocated", C.c_uint32), ("flags", C.c_uint32), ("t_submit_ms", C.c_double),
                ("t_done_ms", C.c_double), ("kind", C.c_int32), ("n_triangles", C.c_uint32),
                ("n_triangles_total", C.c_uint32), ("n_occupied", C.c_uint32), ("rgb_count", C.c_int32)]


class TraceFrame(C.Structure):
    _fields_ = [("session", C.c_int32), ("depth", C.c_void_p), ("rgb", C.c_void_p), ("arrival_ms", C.c_double),
                ("tag", C.c_uint64), ("kind", C.c_int32), ("max_triangles", C.c_uint32), ("vertex", C.c_void_p),
                ("normal", C.c_void_p), ("color", C.c_void_p), ("host", C.c_int32)]


SIGNATURES = {
    "fu_last_error": (C.c_char_p, []),
    "fu_version": (C.c_char_p, []),
    "fu_mc_tables": (I, [IP, IP, IP]),
    "fu_map_params_default": (None, [C.POINTER(MapParams)]),
    "fu_map_bytes": (C.c_size_t, [C.POINTER(MapParams)]),
    "fu_map_create": (I, [C.POINTER(MapParams), C.POINTER(P)]),
    "fu_map_destroy": (I, [P]),
    "fu_map_get_params": (I, [P, C.POINTER(MapParams)]),
    "fu_reset_map": (I, [P, P]),
    "fu_map_stats_get": (I, [P, C.POINTER(MapStats), P]),
    "fu_fuse_map_color": (I, [P, P, P, P, I, I, C.POINTER(PoseF), F, F, F, F, F, F, UP, P]),
    "fu_check_block_visibility": (I, [P, C.POINTER(PoseF), I, I, F, F, F, F, F, F, UP, P]),
    "fu_create_rendering_blocks": (I, [P, P, P, I, I, F, F, C.POINTER(PoseF), U, F, F, F, F, UP, IP, P]),
    "fu_raycast": (I, [P, P, P, I, I, P, P, I, I, C.POINTER(PoseF), F, F, F, F, P]),
    "fu_mesh_scene": (I, [P, P, P, P, P, U, UP, UP, P]),
    "fu_map_export_blocks": (I, [P, P, U, UP, P]),
    "fu_map_export_visible": (I, [P, P, U, UP, P]),
    "fu_map_read_blocks": (I, [P, P, U, P, P, P, P]),
    "fu_filter_depth": (I, [P, P, P, I, I, F, F, P]),
    "fu_pyrdown_gauss": (I, [P, I, I, P, P]),
    "fu_compute_vmap": (I, [P, P, I, I, F, F, F, F, F, P]),
    "fu_compute_nmap": (I, [P, P, I, I, P]),
    "fu_resize_map": (I, [P, P, I, I, P, P, P]),
    "fu_icp_step": (I, [P, P, P, P, I, I, FP, FP, FP, FP, FP, F, F, F, F, FP, DP, DP, P]),
    "fu_icp_step_noalloc": (I, [P, P, P, P, I, I, FP, FP, FP, FP, FP, F, F, F, F, FP, DP, DP, P]),
    "fu_tracker_create": (I, [I, I, C.POINTER(P)]),
    "fu_tracker_destroy": (I, [P]),
    "fu_tracker_set_last_pose": (I, [P, DP, P]),
    "fu_compute_se3": (I, [P, C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P), FP, IP, IP, F, P]),
    "fu_tracker_result": (I, [P, C.POINTER(TrackResult), P]),
    "fu_image_to_intensity": (I, [P, P, I, I, P]),
    "fu_selftest_arith": (I, [C.c_uint64, C.c_uint32, C.POINTER(C.c_uint32), P]),
    "fu_map_save": (I, [P, C.c_char_p, P]),
    "fu_map_load": (I, [P, C.c_char_p, P]),
    "fu_map_check_file": (I, [C.POINTER(MapParams), C.c_char_p]),
    "fu_system_save": (I, [P, C.c_char_p, P]),
    "fu_system_load": (I, [P, C.c_char_p, P]),
    "fu_mesh_write_ply": (I, [C.c_char_p, P, P, P, C.c_uint32, I, P]),
    "fu_pyrdown_gauss_u8": (I, [P, I, I, P, P]),
    "fu_compute_derivative_image": (I, [P, P, P, I, I, P]),
    "fu_rgb_step": (I, [P, P, P, P, P, P, I, I, FP, FP, FP, FP, FP, FP, F, F, F, F, FP, DP, DP,
                        C.POINTER(C.c_int32), P]),
    "fu_compute_se3_rgbd": (I, [P] + [C.POINTER(P)] * 8 + [FP, IP, IP, F, P]),
    "fu_system_params_default": (None, [C.POINTER(SystemParams)]),
    "fu_system_create": (I, [C.POINTER(SystemParams), C.POINTER(P)]),
    "fu_system_destroy": (I, [P]),
    "fu_system_reset": (I, [P, P]),
    "fu_system_get_buffers": (I, [P, C.POINTER(SystemBuffers)]),
    "fu_system_map": (P, [P]),
    "fu_system_grab_image": (I, [P, P]),
    "fu_system_grab_image_from": (I, [P, P, P, P]),
    "fu_system_grab_image_host": (I, [P, P, P, P]),
    "fu_system_graph_counts": (I, [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "fu_system_relocalise": (I, [P, C.POINTER(C.c_double), P]),
    "fu_system_set_mapping": (I, [P, I, P]),
    "fu_system_result": (I, [P, C.POINTER(FrameResult), P]),
    "fu_system_set_profiling": (I, [P, I]),
    "fu_system_get_stage_times": (I, [P, C.POINTER(StageTimes)]),
    "fu_icp_batch_create": (I, [I, C.POINTER(P)]),
    "fu_icp_batch_destroy": (I, [P]),
    "fu_compute_se3_batch": (I, [P, I, C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P), FP,
                                 C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_float, P]),
    "fu_batch_create": (I, [I, C.POINTER(P)]),
    "fu_batch_destroy": (I, [P]),
    "fu_batch_attach": (I, [P, P]),
    "fu_batch_detach": (I, [P, P]),
    "fu_batch_grab": (I, [P, I, C.POINTER(P), C.POINTER(P), C.POINTER(P), P]),
    "fu_batch_set_profiling": (I, [P, I]),
    "fu_batch_get_stage_times": (I, [P, C.POINTER(StageTimes)]),
    "fu_batch_set_graph": (I, [P, I]),
    "fu_batch_graph_counts": (I, [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "fu_tuning_default": (None, [C.POINTER(Tuning)]),
    "fu_set_tuning": (I, [C.POINTER(Tuning)]),
    "fu_get_tuning": (I, [C.POINTER(Tuning)]),
    "fu_server_set_profiling": (I, [P, I]),
    "fu_server_get_stage_times": (I, [P, C.POINTER(StageTimes)]),
    "fu_server_set_graph": (I, [P, I]),
    "fu_server_graph_counts": (I, [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "fu_server_params_default": (None, [C.POINTER(ServerParams)]),
    "fu_server_create": (I, [C.POINTER(ServerParams), C.POINTER(P)]),
    "fu_server_destroy": (I, [P]),
    "fu_server_open": (I, [P, C.POINTER(SystemParams), C.POINTER(I)]),
    "fu_server_close": (I, [P, I]),
    "fu_server_system": (P, [P, I]),
    "fu_server_now_ms": (C.c_double, [P]),
    "fu_server_submit": (I, [P, I, P, P, C.c_uint64, C.c_double]),
    "fu_keymap_params_default": (None, [C.POINTER(KeyMapParams)]),
    "fu_keymap_create": (I, [C.POINTER(KeyMapParams), C.POINTER(P)]),
    "fu_keymap_destroy": (I, [P]),
    "fu_reset_key_points": (I, [P, P]),
    "fu_insert_key_points": (I, [P, P, P, C.c_uint32, P]),
    "fu_collect_key_points": (I, [P, P, C.c_uint32, C.POINTER(C.c_uint32), P]),
    "fu_keymap_read": (I, [P, P, P]),
    "fu_build_adjacency_matrix": (I, [P, P, P, I, P, P]),
    "fu_lift_keypoints": (I, [P, I, P, P, I, I, F, F, F, F, F, P, P, P, P]),
    "fu_match_knn2": (I, [P, I, P, I, P, P, P]),
    "fu_pose_estimate": (I, [DP, DP, I, I, C.c_uint64, DP, P, C.POINTER(I)]),
    "fu_relocalise_keys": (I, [P, P, DP, I, I, C.c_uint64, DP, C.POINTER(I), C.POINTER(I), P]),
    "fu_select_subgraphs": (I, [P, I, I, I, F, P, P, P, P, P, P]),
    "fu_relocalise_graph_candidates": (I, [P, P, P, P, I, I, C.c_uint64, DP, C.POINTER(I), C.POINTER(I), P]),
    "fu_system_validate_poses": (I, [P, DP, I, C.POINTER(I), F, P, P, C.POINTER(I), P]),
    "fu_system_relocalise_graph": (I, [P, P, P, P, P, I, I, C.c_uint64, C.POINTER(I), C.POINTER(I), P]),
    "fu_server_submit_host": (I, [P, I, P, P, C.c_uint64, C.c_double]),
    "fu_server_submit_mesh": (I, [P, I, P, P, P, C.c_uint32, C.c_uint64, C.c_double]),
    "fu_server_poll": (I, [P, C.POINTER(Completion), I, C.POINTER(I)]),
    "fu_server_pending": (I, [P, C.POINTER(I)]),
    "fu_server_run_trace": (I, [P, C.POINTER(TraceFrame), I, I, C.POINTER(Completion)]),
}

SCENE_SIGNATURES = {
    "fu_scene_create": (I, [I, F, F, F, C.c_uint64, C.POINTER(P)]),
    "fu_scene_destroy": (None, [P]),
    "fu_scene_num_primitives": (I, [P]),
    "fu_scene_pose": (I, [P, I, DP]),
    "fu_scene_render_pose": (I, [P, DP, I, I, F, F, F, F, F, I, C.c_uint64, P, P]),
    "fu_scene_render": (I, [P, I, I, I, F, F, F, F, F, I, P, P]),
    "fu_scene_sdf": (F, [P, F, F, F]),
}

_lib = None
_scene = None


def _bind(path, sigs):
    handle = C.CDLL(path)
    for name, (res, args) in sigs.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    return handle


def lib():
    """The HIP library; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FusionError("libfusion_amd.so not built (%s); run __graft_entry__.build()" % LIB_PATH)
        _lib = _bind(LIB_PATH, SIGNATURES)
    return _lib


def scene_lib():
    global _scene
    if _scene is None:
        if not os.path.exists(SCENE_PATH):
            raise FusionError("libfu_scene.so not built (%s)" % SCENE_PATH)
        _scene = _bind(SCENE_PATH, SCENE_SIGNATURES)
    return _scene


def check(status, what=""):
    if status != FU_OK:
        msg = lib().fu_last_error().decode(errors="replace")
        raise FusionError("%s failed (%d): %s" % (what, status, msg))
