"""ctypes bindings of the built libraries.

libcublox_hip.so      HIP kernels + C ABI (include/cublox/capi.h)
libcublox.so          drop-in C++ OccupancyGrid + flat C binding (og_capi.cpp)
libcublox_seq.so      native replay of a driver's scan loop (seq_replay.h)
libcublox_scangen.so  synthetic LiDAR workload generator

Nothing here falls back to a CPU path: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))

f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int)
u32p = C.POINTER(C.c_uint)
u8p = C.POINTER(C.c_ubyte)
vp = C.c_void_p


class Float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Int3(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int)]


class RaycastCfg(C.Structure):
    _fields_ = [
        ("origin", Float3), ("origin_i", Int3), ("map_size_i", Int3),
        ("half_map_size_i", Int3), ("resolution", C.c_float),
        ("inv_resolution", C.c_float), ("max_range", C.c_float),
        ("map_vox_num", C.c_int), ("origin_at_center", C.c_int),
    ]


class RecenterCfg(C.Structure):
    _fields_ = [
        ("pos", Float3), ("origin_f", Float3), ("origin_i", Int3),
        ("map_size_i", Int3), ("half_map_size_i", Int3),
        ("resolution", C.c_float), ("inv_resolution", C.c_float),
        ("recenter_threshold_sq", C.c_float), ("origin_at_center", C.c_int),
        ("map_vox_num", C.c_int),
    ]


class RecenterResult(C.Structure):
    _fields_ = [
        ("action", C.c_int), ("new_origin_i", Int3), ("n_slices", C.c_int * 3),
        ("zeroed_voxels", C.c_longlong),
    ]


class FetchCfg(C.Structure):
    _fields_ = [
        ("center_g", Int3), ("origin_i", Int3), ("map_size_i", Int3),
        ("half_map_size_i", Int3), ("gz_min", C.c_int), ("gz_max", C.c_int),
        ("disk_radius_vox", C.c_int), ("center", Float3),
        ("radius_sq", C.c_float), ("resolution", C.c_float),
        ("l_free", C.c_float), ("l_occupied", C.c_float),
        ("l_unknown", C.c_float), ("origin_at_center", C.c_int),
        ("state_filter", C.c_int),
    ]


class SceneCfg(C.Structure):
    _fields_ = [
        ("n_boxes", C.c_int), ("box_extent", C.c_float),
        ("box_min_side", C.c_float), ("box_max_side", C.c_float),
        ("corridor_half", C.c_float), ("speed", C.c_float),
        ("yaw_rate_deg", C.c_float), ("sensor_height", C.c_float),
        ("noise_sigma", C.c_float), ("miss_extra", C.c_float),
        ("seed", C.c_ulonglong),
    ]


def _load(name: str) -> C.CDLL:
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is not built; run `python -c 'import __graft_entry__ as g; "
            "g.build()'` (or `make -C cublox_amd/csrc`) first")
    return C.CDLL(path, mode=C.RTLD_GLOBAL)


def _sig(lib, name, restype, *argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = list(argtypes)
    return fn


_hip = None
_host = None
_gen = None


def hip() -> C.CDLL:
    """The HIP kernel library with its C ABI signatures declared."""
    global _hip
    if _hip is not None:
        return _hip
    L = _load("libcublox_hip.so")
    S = C.c_size_t
    i = C.c_int
    _sig(L, "cublox_last_error", C.c_char_p)
    _sig(L, "cublox_device_count", i, i32p)
    _sig(L, "cublox_set_device", i, i)
    _sig(L, "cublox_get_device", i, i32p)
    _sig(L, "cublox_device_name", i, C.c_char_p, i)
    _sig(L, "cublox_mem_info", i, C.POINTER(S), C.POINTER(S))
    _sig(L, "cublox_malloc", i, C.POINTER(vp), S)
    _sig(L, "cublox_free", i, vp)
    _sig(L, "cublox_host_alloc", i, C.POINTER(vp), S)
    _sig(L, "cublox_host_free", i, vp)
    _sig(L, "cublox_memset_async", i, vp, i, S, vp)
    _sig(L, "cublox_memcpy_h2d_async", i, vp, vp, S, vp)
    _sig(L, "cublox_memcpy_d2h_async", i, vp, vp, S, vp)
    _sig(L, "cublox_memcpy_d2d_async", i, vp, vp, S, vp)
    _sig(L, "cublox_stream_create", i, C.POINTER(vp))
    _sig(L, "cublox_stream_destroy", i, vp)
    _sig(L, "cublox_stream_synchronize", i, vp)
    _sig(L, "cublox_device_synchronize", i)
    _sig(L, "cublox_event_create", i, C.POINTER(vp))
    _sig(L, "cublox_event_destroy", i, vp)
    _sig(L, "cublox_event_record", i, vp, vp)
    _sig(L, "cublox_event_synchronize", i, vp)
    _sig(L, "cublox_event_elapsed_ms", i, vp, vp, f32p)
    _sig(L, "cublox_stream_wait_event", i, vp, vp)
    _sig(L, "cublox_fill_zero_async", i, vp, S, vp)
    _sig(L, "cublox_launch_ray_cast_update", i, C.POINTER(RaycastCfg), vp, vp,
         vp, i, vp, vp, vp)
    _sig(L, "cublox_launch_ray_cast_update_sensor", i, C.POINTER(RaycastCfg),
         f32p, vp, vp, vp, i, vp, vp, vp)
    _sig(L, "cublox_launch_apply_update", i, vp, vp, vp, i, C.c_float,
         C.c_float, C.c_float, C.c_float, C.c_float, vp, vp, vp, vp, C.c_uint)
    _sig(L, "cublox_launch_clear_voxels_by_index", i, vp, vp, i, i, vp)
    _sig(L, "cublox_plan_recenter", i, C.POINTER(RecenterCfg), i32p, i32p,
         i32p, C.POINTER(RecenterResult))
    _sig(L, "cublox_launch_recenter", i, vp, vp, vp, C.POINTER(RecenterCfg),
         vp, C.POINTER(RecenterResult))
    _sig(L, "cublox_launch_clear_recenter_slabs_for_axis", i, vp, vp, vp, Int3,
         Int3, vp, i, i, vp)
    _sig(L, "cublox_fetch_plan", i, Int3, Int3, Int3, C.c_float, C.c_float, i,
         C.c_float, C.c_float, C.c_float, Float3, C.c_float, i, i,
         C.POINTER(FetchCfg), C.POINTER(C.c_longlong), u32p)
    _sig(L, "cublox_launch_fetch_occupancy_around", i, C.POINTER(FetchCfg), vp,
         vp, vp, vp, C.c_uint, C.c_uint, vp)
    _sig(L, "cublox_launch_fetch_occupancy_batch", i, vp, i, i, i, vp, vp, vp,
         vp, C.c_uint, vp)
    _sig(L, "cublox_counter_census", i, vp, vp, i, vp, vp)
    _sig(L, "cublox_launch_index_chain", i, vp, i, vp, vp)
    _sig(L, "cublox_launch_integrate_noalloc", i, C.POINTER(RaycastCfg), f32p, vp, vp,
         vp, i, vp, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, i, vp,
         vp, vp)
    _sig(L, "cublox_launch_fetch_gather", i, vp, vp, i, C.c_uint, C.c_uint, vp,
         vp, vp, vp, vp)
    _sig(L, "cublox_integrate_workspace_bytes", S, C.POINTER(RaycastCfg), i, i)
    _sig(L, "cublox_integrate_workspace_init", i, vp, S, vp)
    _sig(L, "cublox_launch_integrate", i, C.POINTER(RaycastCfg), f32p, vp, vp,
         vp, i, vp, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, vp,
         S, i, i, vp, vp, vp)
    _sig(L, "cublox_integrate_segment_bound", i, C.POINTER(RaycastCfg))
    _sig(L, "cublox_integrate_tile_count", i, C.POINTER(RaycastCfg))
    _sig(L, "cublox_integrate_batch_workspace_bytes", S, vp)
    _sig(L, "cublox_integrate_batch_staging_bytes", S, vp)
    _sig(L, "cublox_integrate_batch_stage", i, vp, i, vp, i, vp)
    _sig(L, "cublox_integrate_batch_stage_tuned", i, vp, i, vp, vp, vp)
    _sig(L, "cublox_launch_integrate_batch_staged", i, vp, vp, S, vp, vp)
    _sig(L, "cublox_launch_integrate_batch", i, vp, i, vp, vp, vp, S, vp)
    _sig(L, "cublox_integrate_batch_counters", i, vp, S, vp, vp, vp)
    _sig(L, "cublox_integrate_counters", i, vp, S, C.POINTER(RaycastCfg), i, i,
         vp, vp)
    _sig(L, "cublox_recenter_batch_staging_bytes", S, i)
    _sig(L, "cublox_recenter_batch_stage", i, vp, i, i, vp, vp)
    _sig(L, "cublox_launch_recenter_batch_staged", i, vp, vp, i, vp)
    _sig(L, "cublox_graph_begin_capture", i, vp)
    _sig(L, "cublox_graph_end_capture", i, vp, C.POINTER(vp))
    _sig(L, "cublox_graph_launch", i, vp, vp)
    _sig(L, "cublox_graph_destroy", i, vp)
    _hip = L
    return L


def host() -> C.CDLL:
    """The drop-in OccupancyGrid library (loads libcublox_hip.so first)."""
    global _host
    if _host is not None:
        return _host
    hip()
    L = _load("libcublox.so")
    i = C.c_int
    F = C.c_float
    _sig(L, "cublox_og_last_error", C.c_char_p)
    _sig(L, "cublox_og_create", vp, i, i, i, F, i, i, C.c_double, f32p)
    _sig(L, "cublox_og_destroy", None, vp)
    _sig(L, "cublox_og_set_max_range", i, vp, F)
    _sig(L, "cublox_og_set_log_odds", i, vp, F, F, F, F, F, F)
    _sig(L, "cublox_og_update", i, vp, f32p, f32p, f32p, i, f32p)
    _sig(L, "cublox_og_update_sensor", i, vp, f32p, f32p, f32p, i, f32p)
    _sig(L, "cublox_og_update_sensor_device", i, vp, vp, vp, vp, i, f32p)
    _sig(L, "cublox_og_update_device", i, vp, vp, vp, vp, i, f32p)
    _sig(L, "cublox_og_cast_rays_device", i, vp, vp, vp, vp, i, f32p)
    _sig(L, "cublox_og_apply_counters", i, vp)
    _sig(L, "cublox_og_cast_rays_sensor_device", i, vp, vp, vp, vp, i, f32p)
    _sig(L, "cublox_og_census", i, vp, C.POINTER(C.c_ulonglong))
    _sig(L, "cublox_og_set_update_path", i, vp, i)
    _sig(L, "cublox_og_set_near_steps", i, vp, i)
    _sig(L, "cublox_og_set_changed_capacity", i, vp, C.c_uint)
    _sig(L, "cublox_og_save_snapshot", i, vp, C.POINTER(vp), C.POINTER(C.c_size_t))
    _sig(L, "cublox_og_load_snapshot", i, vp, vp, C.c_size_t)
    _sig(L, "cublox_og_free_buffer", None, vp)
    _sig(L, "cublox_og_save_snapshot_file", i, vp, C.c_char_p)
    _sig(L, "cublox_og_load_snapshot_file", i, vp, C.c_char_p)
    _sig(L, "cublox_og_viz_cloud", i, vp, f32p, C.c_float, i, i, f32p, C.c_uint,
         u32p, C.POINTER(C.c_ulonglong))
    _sig(L, "cublox_og_update_cloud2", i, vp, vp, i, C.c_ulonglong, u32p, f32p, i32p)
    _sig(L, "cublox_og_update_depth", i, vp, vp, i, C.c_uint, C.c_uint,
         C.c_uint, i, f32p, f32p, i32p)
    _sig(L, "cublox_og_changed_voxels", i, vp, i32p, f32p, C.c_uint,
         C.POINTER(C.c_ulonglong), u32p)
    _sig(L, "cublox_og_cast_rays_fused_sensor_device", i, vp, vp, vp, vp, i, f32p)
    _sig(L, "cublox_og_recenter", i, vp, f32p, i32p)
    _sig(L, "cublox_og_fetch", i, vp, f32p, F, i, i, f32p, u8p, i, i32p)
    _sig(L, "cublox_og_fetch_batch", i, vp, f32p, i, F, i, i, u32p, f32p, u8p)
    _sig(L, "cublox_og_occ_to_host", i, vp, f32p)
    _sig(L, "cublox_og_counters_to_host", i, vp, i32p, i32p)
    _sig(L, "cublox_og_load_occ", i, vp, f32p)
    _sig(L, "cublox_og_reset", i, vp)
    _sig(L, "cublox_og_reset_voxel", i, vp, i)
    _sig(L, "cublox_og_geometry", i, vp, i32p, i32p, i32p, f32p, i32p, i32p,
         i32p, f32p, f32p)
    _sig(L, "cublox_og_query", i, vp, i, i, f32p, i32p, i32p)
    _sig(L, "cublox_og_max_horizontal_radius", i, vp, f32p, f32p)
    _sig(L, "cublox_og_enable_timing", i, vp, i)
    _sig(L, "cublox_og_kernel_times", i, vp, C.POINTER(C.c_double),
         C.POINTER(C.c_longlong))
    _sig(L, "cublox_og_reset_kernel_times", i, vp)
    _sig(L, "cublox_og_device_occ", vp, vp)
    _sig(L, "cublox_og_stream", vp, vp)
    pp = C.POINTER(vp)
    _sig(L, "cublox_gb_create", vp, i32p)
    _sig(L, "cublox_gb_destroy", None, vp)
    _sig(L, "cublox_gb_segment_bound", i, vp)
    _sig(L, "cublox_gb_tile_count", i, vp)
    _sig(L, "cublox_gb_tick", i, vp, i, pp, pp, pp, pp, i32p, f32p, i32p, f32p,
         i, pp, f32p)
    _sig(L, "cublox_gb_submit", i, vp, i, pp, pp, pp, pp, i32p, f32p, i32p, f32p,
         i, pp, f32p)
    _sig(L, "cublox_gb_wait", i, vp)
    _sig(L, "cublox_gb_set_pipelined", i, vp, i)
    _sig(L, "cublox_gb_set_graph_mode", i, vp, i)
    _sig(L, "cublox_gb_set_tuning", i, vp, i, i32p, i)
    _sig(L, "cublox_gb_tuning", i, vp, i32p)
    _sig(L, "cublox_gb_setting", i, vp, i32p)
    _sig(L, "cublox_gb_set_setting", i, vp, i, i, i, i)
    _sig(L, "cublox_gb_tune_results", i, vp, C.POINTER(C.c_double), i)
    _sig(L, "cublox_gb_last", i, vp, C.POINTER(C.c_double))
    _sig(L, "cublox_gb_counters", i, vp, C.POINTER(C.c_int))
    ll = C.c_longlong
    llp = C.POINTER(ll)
    dp = C.POINTER(C.c_double)
    _sig(L, "cublox_ms_create", vp, llp)
    _sig(L, "cublox_ms_destroy", None, vp)
    _sig(L, "cublox_ms_session_bytes", ll, i32p)
    _sig(L, "cublox_ms_open", i, vp, i32p, F, i, i, C.c_double, F, f32p, i32p)
    _sig(L, "cublox_ms_close", i, vp, i)
    _sig(L, "cublox_ms_admitted", i, vp, i, i32p)
    _sig(L, "cublox_ms_submit_scan", i, vp, i, vp, vp, vp, i, f32p, f32p, llp)
    _sig(L, "cublox_ms_submit_fetch", i, vp, i, f32p, F, i, i, llp)
    _sig(L, "cublox_ms_drain", i, vp)
    _sig(L, "cublox_ms_pause", i, vp)
    _sig(L, "cublox_ms_resume", i, vp)
    _sig(L, "cublox_ms_take_completions", i, vp, llp, i32p, i32p, dp, llp, llp,
         i, i32p)
    _sig(L, "cublox_ms_keep_fetch_results", i, vp, i)
    _sig(L, "cublox_ms_fetch_result", i, vp, ll, f32p, u8p, i, i32p)
    _sig(L, "cublox_ms_copy_occupancy", i, vp, i, f32p)
    _sig(L, "cublox_ms_origin", i, vp, i, i32p)
    _sig(L, "cublox_ms_stats", i, vp, dp)
    _sig(L, "cublox_ms_evict", i, vp, i)
    _sig(L, "cublox_ms_load_run", i, vp, i32p, i, vp, vp, vp, i32p, f32p, f32p, vp,
         llp, dp)
    _sig(L, "cublox_utils_pos_to_global", i, F, F, i)
    _sig(L, "cublox_utils_global_to_pos", F, i, F, i)
    _sig(L, "cublox_utils_global_to_local", i, i, i, i)
    _sig(L, "cublox_utils_local_to_global", i, i, i, i, i)
    _sig(L, "cublox_utils_local_to_hash", i, i32p, i32p, i32p)
    _sig(L, "cublox_utils_hash_to_local", None, i, i32p, i32p, i32p)
    _sig(L, "cublox_utils_hash_to_pos", None, i, i32p, i32p, i32p, F, i, f32p)
    _sig(L, "cublox_utils_pos_to_hash", i, f32p, i32p, i32p, F, i)
    _sig(L, "cublox_logit", F, F)
    _host = L
    return L


class SeqStep(C.Structure):
    """cublox_seq_step (include/cublox/seq_replay.h)."""
    _fields_ = [
        ("grid", vp), ("xs", vp), ("ys", vp), ("zs", vp), ("n", C.c_int),
        ("T", C.c_float * 12), ("base", C.c_float * 3),
    ]


_seq = None


def seq() -> C.CDLL:
    """The native scan-loop replay."""
    global _seq
    if _seq is not None:
        return _seq
    host()
    L = _load("libcublox_seq.so")
    _sig(L, "cublox_seq_replay", C.c_int, C.POINTER(SeqStep), C.c_int, C.c_int,
         C.POINTER(C.c_double), i32p)
    _seq = L
    return L


def gen() -> C.CDLL:
    """The synthetic workload generator."""
    global _gen
    if _gen is not None:
        return _gen
    L = _load("libcublox_scangen.so")
    i = C.c_int
    F = C.c_float
    _sig(L, "cublox_scene_create", vp, C.POINTER(SceneCfg))
    _sig(L, "cublox_scene_destroy", None, vp)
    _sig(L, "cublox_scene_box_count", i, vp)
    _sig(L, "cublox_scene_boxes", None, vp, C.POINTER(C.c_double))
    _sig(L, "cublox_scene_pose", None, vp, i, f32p, f32p)
    _sig(L, "cublox_scene_scan", i, vp, i, i, i, F, F, F, f32p, f32p, f32p,
         f32p, f32p, i32p, i)
    _sig(L, "cublox_scene_queries", None, vp, i, i, F, f32p)
    _sig(L, "cublox_rng_uniform", C.c_double, C.c_ulonglong, C.c_ulonglong,
         C.c_ulonglong)
    _gen = L
    return L


def check(status: int, lib=None) -> None:
    """Raise RuntimeError with the library's message on a non-zero status."""
    if status == 0:
        return
    msg = ""
    try:
        if lib is _host and _host is not None:
            msg = _host.cublox_og_last_error().decode()
        else:
            msg = hip().cublox_last_error().decode()
    except Exception:  # pragma: no cover - message lookup is best effort
        pass
    raise RuntimeError(f"cublox status {status}: {msg}")


def fptr(a):
    return a.ctypes.data_as(f32p)


def iptr(a):
    return a.ctypes.data_as(i32p)


def uptr(a):
    return a.ctypes.data_as(u32p)


def bptr(a):
    return a.ctypes.data_as(u8p)
