"""Argument validation of the C ABI that runs on the host (no GPU work):
batch staging of the new per-grid options (point stride, range drop,
changed-voxel stream) and the early checks of the ingest and snapshot
launchers. Each case fails before any device call, with the error code and
message the caller sees."""
import ctypes as C

import pytest

OK, INVALID, RUNTIME, CAPACITY = 0, 1, 2, 3


class Limits(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("max_grids", "max_points", "max_tiles",
                                       "max_segments_per_ray", "max_near_steps",
                                       "chunk", "max_slabs", "delta")]


def _grid_struct(cb):
    from cublox_amd import _ffi
    fp = C.c_void_p

    class Grid(C.Structure):
        _fields_ = [
            ("cfg", _ffi.RaycastCfg), ("T", C.c_float * 12), ("use_T", C.c_int),
            ("d_xs", fp), ("d_ys", fp), ("d_zs", fp), ("n", C.c_int),
            ("d_occ", fp), ("l_hit", C.c_float), ("l_miss", C.c_float),
            ("l_min", C.c_float), ("l_max", C.c_float), ("l_unknown", C.c_float),
            ("near_steps", C.c_int), ("d_export_op", fp), ("d_export_hit", fp),
            ("d_delta_count", fp), ("d_delta_idx", fp), ("d_delta_val", fp),
            ("delta_capacity", C.c_uint), ("point_stride", C.c_int),
            ("range_drop_sq", C.c_float),
        ]
    return Grid


def _setup(cb, delta_limit=0):
    from cublox_amd import _ffi
    L = _ffi.hip()
    Grid = _grid_struct(cb)
    g = Grid()
    half = (20, 20, 10)
    c = g.cfg
    c.map_size_i.x, c.map_size_i.y, c.map_size_i.z = (2 * h + 1 for h in half)
    c.half_map_size_i.x, c.half_map_size_i.y, c.half_map_size_i.z = half
    c.resolution, c.inv_resolution, c.max_range = 0.1, 10.0, 5.0
    c.map_vox_num = 41 * 41 * 21
    g.T[0] = g.T[5] = g.T[10] = 1.0
    g.use_T = 1
    fake = 0x10000  # never dereferenced: staging only writes descriptors
    g.d_xs, g.d_ys, g.d_zs, g.d_occ = fake, fake + 64, fake + 128, fake + 256
    g.n = 1000
    g.near_steps = 8
    lim = Limits(max_grids=4, max_points=4096,
                 max_tiles=L.cublox_integrate_tile_count(C.byref(g.cfg)) * 4,
                 max_segments_per_ray=L.cublox_integrate_segment_bound(C.byref(g.cfg)),
                 max_near_steps=12, chunk=0, max_slabs=0, delta=delta_limit)
    stage = (C.c_char * L.cublox_integrate_batch_staging_bytes(C.byref(lim)))()
    return L, g, lim, stage


def _stage(L, g, lim, stage):
    return L.cublox_integrate_batch_stage(C.byref(g), 1, C.byref(lim), 0, stage)


def _err(L):
    return L.cublox_last_error().decode()


def test_batch_stage_accepts_new_options(cb):
    L, g, lim, stage = _setup(cb, delta_limit=1)
    g.point_stride = 4
    g.range_drop_sq = 25.0
    g.d_delta_count, g.d_delta_idx, g.d_delta_val = 0x20000, 0x30000, 0x40000
    g.delta_capacity = 100
    assert _stage(L, g, lim, stage) == OK


def test_batch_stage_rejects(cb):
    L, g, lim, stage = _setup(cb)
    g.d_delta_count = 0x20000
    assert _stage(L, g, lim, stage) == INVALID and "limits.delta" in _err(L)
    L, g, lim, stage = _setup(cb, delta_limit=1)
    g.d_delta_count, g.delta_capacity = 0x20000, 10
    assert _stage(L, g, lim, stage) == INVALID and "both arrays" in _err(L)
    L, g, lim, stage = _setup(cb, delta_limit=1)
    g.d_delta_count = 0x20000
    g.d_export_op, g.d_export_hit = 0x50000, 0x60000
    assert _stage(L, g, lim, stage) == INVALID and "count export" in _err(L)
    L, g, lim, stage = _setup(cb)
    g.point_stride = -1
    assert _stage(L, g, lim, stage) == INVALID and "stride" in _err(L)
    L, g, lim, stage = _setup(cb)
    g.use_T = 0
    g.range_drop_sq = 4.0
    assert _stage(L, g, lim, stage) == INVALID and "sensor-frame" in _err(L)


def test_ingest_and_snapshot_launchers_reject_early(cb):
    from cublox_amd import _ffi
    L = _ffi.hip()
    fn = L.cublox_launch_depth_to_points
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_size_t, C.c_int,
                   C.POINTER(C.c_float), C.c_int, C.c_void_p, C.c_void_p,
                   C.c_void_p, C.c_void_p]
    K = (C.c_float * 4)(525.0, 525.0, 319.5, 239.5)
    fake = 0x10000
    assert fn(fake, 640, 480, 1280, 2, K, 10, fake, fake, fake, None) == INVALID
    assert "encoding" in _err(L)
    assert fn(fake, 640, 480, 1000, 0, K, 10, fake, fake, fake, None) == INVALID
    assert "row step" in _err(L)
    assert fn(None, 640, 480, 1280, 0, K, 10, fake, fake, fake, None) == INVALID
    # Nothing to do is not an error.
    assert fn(fake, 0, 480, 1280, 0, K, 10, fake, fake, fake, None) == OK
    sc = L.cublox_launch_sparse_compact
    sc.restype = C.c_int
    sc.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p,
                   C.c_void_p, C.c_uint, C.c_void_p, C.c_void_p]
    assert sc(fake, 1000, fake, 8, None, None, 0, fake, None) == CAPACITY
    assert "scratch" in _err(L)
    assert sc(fake, 1000, fake, 1 << 20, None, None, 5, fake, None) == INVALID
    assert sc(fake, 1000, fake, 1 << 20, None, None, 0, None, None) == INVALID
    sz = L.cublox_sparse_scratch_bytes
    sz.restype = C.c_size_t
    sz.argtypes = [C.c_int]
    assert sz(8192) == 4 + 256 and sz(8193) == 8 + 256
