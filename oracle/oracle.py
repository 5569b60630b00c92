"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by the test suite, the smoke check and
the CPU-baseline leg of bench.py, always as the checker or the baseline,
never as the measured product path.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int)
u8p = C.POINTER(C.c_ubyte)


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        raise ImportError(f"{path} not built (make -C oracle)")
    L = C.CDLL(path)
    i, F, vp = C.c_int, C.c_float, C.c_void_p

    def sig(name, res, *args):
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = list(args)

    sig("orc_pos_to_global", i, F, F, i)
    sig("orc_global_to_pos", F, i, F, i)
    sig("orc_global_to_local", i, i, i, i)
    sig("orc_local_to_global", i, i, i, i, i)
    sig("orc_local_to_hash", i, i32p, i32p, i32p)
    sig("orc_hash_to_local", None, i, i32p, i32p, i32p)
    sig("orc_global_to_hash", i, i32p, i32p, i32p)
    sig("orc_logit", F, F)
    sig("orc_trace_ray", i, f32p, f32p, i32p, i32p, i32p, F, F, F, i, i, i32p,
        i, i32p, i32p)
    sig("orc_plan_recenter", i, f32p, f32p, i32p, i32p, i32p, F, F, i, i32p,
        i32p, i32p, i32p, i32p)
    sig("orc_grid_create", vp, i, i, i, F, i, i, C.c_double, F, F, F)
    sig("orc_grid_destroy", None, vp)
    sig("orc_grid_set_threads", None, vp, i)
    sig("orc_grid_set_max_range", None, vp, F)
    sig("orc_grid_set_log_odds", None, vp, F, F, F, F, F, F)
    sig("orc_grid_update", None, vp, f32p, f32p, f32p, i, f32p)
    sig("orc_grid_update_sensor", None, vp, f32p, f32p, f32p, i, f32p)
    sig("orc_grid_raycast_world", None, vp, f32p, f32p, f32p, i, f32p)
    sig("orc_grid_apply", None, vp)
    sig("orc_grid_recenter", i, vp, f32p)
    sig("orc_grid_fetch", i, vp, f32p, F, i, i, f32p, u8p, i)
    sig("orc_grid_reset", None, vp)
    sig("orc_grid_reset_voxel", None, vp, i)
    sig("orc_grid_occ", f32p, vp)
    sig("orc_grid_op", i32p, vp)
    sig("orc_grid_hit", i32p, vp)
    sig("orc_grid_voxel_num", i, vp)
    sig("orc_grid_origin", None, vp, i32p, f32p)
    sig("orc_grid_load_occ", None, vp, f32p)
    sig("orc_transform_points", None, f32p, f32p, f32p, f32p, i, f32p, f32p,
        f32p)
    sig("orc_build_cloud", i, u8p, C.c_longlong, i, i, i, i, f32p, F, f32p,
        f32p, f32p)
    sig("orc_depth_to_cloud", None, u8p, i, i, C.c_longlong, i, f32p, f32p,
        f32p, f32p)
    sig("orc_viz_points", i, f32p, i, f32p, F, F, i, i, i, f32p)
    _lib = L
    return L


def _f(a):
    return a.ctypes.data_as(f32p)


def _i(a):
    return a.ctypes.data_as(i32p)


def _a3(v, dt):
    return np.ascontiguousarray(np.asarray(v, dt).reshape(3))


class OracleGrid:
    """CPU twin of OccupancyGrid (the reference host logic + kernels)."""

    def __init__(self, half: Sequence[int], resolution: float,
                 origin_at_center: bool = False,
                 recenter_threshold: Optional[float] = None,
                 origin=(0.0, 0.0, 0.0), threads: int = 0):
        L = lib()
        o = _a3(origin, np.float32)
        self._h = L.orc_grid_create(int(half[0]), int(half[1]), int(half[2]),
                                    float(resolution), 1 if origin_at_center else 0,
                                    0 if recenter_threshold is None else 1,
                                    float(recenter_threshold or 0.0),
                                    float(o[0]), float(o[1]), float(o[2]))
        if not self._h:
            raise RuntimeError("voxel count exceeds int32")
        self._L = L
        if threads:
            L.orc_grid_set_threads(self._h, int(threads))

    def __del__(self):
        try:
            if self._h:
                self._L.orc_grid_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def voxel_num(self) -> int:
        return self._L.orc_grid_voxel_num(self._h)

    def set_threads(self, n: int) -> None:
        self._L.orc_grid_set_threads(self._h, int(n))

    def set_max_raycast_range(self, r: float) -> None:
        self._max_range = float(r)
        self._L.orc_grid_set_max_range(self._h, float(r))

    def update_cloud2(self, data, n_records, point_step, offsets, T) -> bool:
        """The driver's run() step for a PointCloud2 payload: build the
        odom-frame cloud (main.cpp:640-693) with the grid's max range, and
        update only when a row was kept (main.cpp:450-453). Returns that."""
        r = np.float32(getattr(self, "_max_range", 25.0))
        xw, yw, zw = build_cloud(data, n_records, point_step, offsets, T, r * r)
        if len(xw) == 0:
            return False
        Tm = np.asarray(T, np.float32).reshape(12)
        self.update((xw, yw, zw), (Tm[3], Tm[7], Tm[11]))
        return True

    def update_depth(self, img, width, height, step, encoding, K, T) -> bool:
        """depthCallback (main.cpp:558-638), then update_cloud2 on the cloud
        it publishes."""
        xs, ys, zs = depth_to_cloud(img, width, height, step, encoding, K)
        packed = np.ascontiguousarray(np.stack([xs, ys, zs], 1), np.float32)
        return self.update_cloud2(packed.view(np.uint8).reshape(-1), len(xs), 12,
                                  (0, 4, 8), T)

    def set_log_odds_params(self, *vals) -> None:
        self._L.orc_grid_set_log_odds(self._h, *[float(v) for v in vals])

    @staticmethod
    def _cols(points):
        if isinstance(points, (tuple, list)) and len(points) == 3:
            return [np.ascontiguousarray(np.asarray(c, np.float32)) for c in points]
        p = np.asarray(points, np.float32)
        return [np.ascontiguousarray(p[:, k]) for k in range(3)]

    def update(self, points, sensor_origin) -> None:
        xs, ys, zs = self._cols(points)
        o = _a3(sensor_origin, np.float32)
        self._L.orc_grid_update(self._h, _f(xs), _f(ys), _f(zs), len(xs), _f(o))

    def update_sensor_frame(self, points, T) -> None:
        xs, ys, zs = self._cols(points)
        Tm = np.ascontiguousarray(np.asarray(T, np.float32).reshape(12))
        self._L.orc_grid_update_sensor(self._h, _f(xs), _f(ys), _f(zs), len(xs),
                                       _f(Tm))

    def raycast(self, points, sensor_origin) -> None:
        xs, ys, zs = self._cols(points)
        o = _a3(sensor_origin, np.float32)
        self._L.orc_grid_raycast_world(self._h, _f(xs), _f(ys), _f(zs), len(xs),
                                       _f(o))

    def apply(self) -> None:
        self._L.orc_grid_apply(self._h)

    def recenter(self, pos) -> int:
        p = _a3(pos, np.float32)
        return self._L.orc_grid_recenter(self._h, _f(p))

    def fetch(self, center, radius, state_filter=None, max_results=0, cap=None):
        c = _a3(center, np.float32)
        if cap is None:
            cap = 500000 if max_results <= 0 else max_results
        pos = np.empty((max(cap, 1), 3), np.float32)
        st = np.empty(max(cap, 1), np.uint8)
        n = self._L.orc_grid_fetch(self._h, _f(c), float(radius),
                                   -1 if state_filter is None else int(state_filter),
                                   int(max_results), _f(pos),
                                   st.ctypes.data_as(u8p), int(cap))
        k = min(n, cap)
        return pos[:k].copy(), st[:k].copy(), n

    def reset(self) -> None:
        self._L.orc_grid_reset(self._h)

    def reset_voxel(self, h: int) -> None:
        self._L.orc_grid_reset_voxel(self._h, int(h))

    def occupancy(self) -> np.ndarray:
        V = self.voxel_num
        return np.ctypeslib.as_array(self._L.orc_grid_occ(self._h), (V,)).copy()

    def counters(self):
        V = self.voxel_num
        op = np.ctypeslib.as_array(self._L.orc_grid_op(self._h), (V,)).copy()
        hit = np.ctypeslib.as_array(self._L.orc_grid_hit(self._h), (V,)).copy()
        return op, hit

    def load_occupancy(self, occ) -> None:
        occ = np.ascontiguousarray(occ, np.float32)
        self._L.orc_grid_load_occ(self._h, _f(occ))

    def origin(self):
        oi = np.zeros(3, np.int32)
        of = np.zeros(3, np.float32)
        self._L.orc_grid_origin(self._h, _i(oi), _f(of))
        return tuple(oi.tolist()), tuple(of.tolist())


def trace_ray(origin, endpoint, origin_i, map_size, half, res, inv_res,
              max_range, center=False, cap=100000):
    """Hash ids incremented by one ray, in order (endpoint last when counted).

    Returns (ids, clipped, endpoint_counted) or None when the ray is skipped.
    """
    L = lib()
    o = _a3(origin, np.float32)
    e = _a3(endpoint, np.float32)
    oi = _a3(origin_i, np.int32)
    m = _a3(map_size, np.int32)
    h = _a3(half, np.int32)
    out = np.empty(cap, np.int32)
    clipped = C.c_int(0)
    ep = C.c_int(0)
    V = int(m[0]) * int(m[1]) * int(m[2])
    n = L.orc_trace_ray(_f(o), _f(e), _i(oi), _i(m), _i(h), float(res),
                        float(inv_res), float(max_range), V, 1 if center else 0,
                        _i(out), cap, C.byref(clipped), C.byref(ep))
    if n < 0:
        return None
    return out[:min(n, cap)].copy(), bool(clipped.value), bool(ep.value)


def plan_recenter(pos, origin_f, origin_i, map_size, half, inv_res, thr_sq,
                  center=False):
    L = lib()
    m = _a3(map_size, np.int32)
    cap = 2 * int(max(m)) + 2
    bufs = [np.zeros(cap, np.int32) for a in range(3)]
    no = np.zeros(3, np.int32)
    ns = np.zeros(3, np.int32)
    action = L.orc_plan_recenter(_f(_a3(pos, np.float32)),
                                 _f(_a3(origin_f, np.float32)),
                                 _i(_a3(origin_i, np.int32)), _i(m),
                                 _i(_a3(half, np.int32)), float(inv_res),
                                 float(thr_sq), 1 if center else 0, _i(no),
                                 _i(bufs[0]), _i(bufs[1]), _i(bufs[2]), _i(ns))
    return action, tuple(no.tolist()), [bufs[a][:ns[a]].copy() for a in range(3)]


def transform_points(T, xs, ys, zs):
    L = lib()
    Tm = np.ascontiguousarray(np.asarray(T, np.float32).reshape(12))
    xs, ys, zs = (np.ascontiguousarray(np.asarray(c, np.float32)) for c in (xs, ys, zs))
    n = len(xs)
    out = [np.empty(n, np.float32) for _ in range(3)]
    L.orc_transform_points(_f(Tm), _f(xs), _f(ys), _f(zs), n, _f(out[0]),
                           _f(out[1]), _f(out[2]))
    return out


def build_cloud(data, n_records, point_step, offsets, T, max_range_sq):
    """orc_build_cloud: (xw, yw, zw) of the kept, transformed points."""
    L = lib()
    buf = np.ascontiguousarray(np.frombuffer(memoryview(data).cast("B"), np.uint8))
    need = int(n_records) * int(point_step)
    assert buf.size >= min(need, 1000000 * int(point_step)), "payload too short"
    Tm = np.ascontiguousarray(np.asarray(T, np.float32).reshape(12))
    cap = max(1, min(int(n_records), 1000000))
    out = [np.empty(cap, np.float32) for _ in range(3)]
    rows = L.orc_build_cloud(buf.ctypes.data_as(u8p), int(n_records), int(point_step),
                             int(offsets[0]), int(offsets[1]), int(offsets[2]),
                             _f(Tm), float(max_range_sq), _f(out[0]), _f(out[1]),
                             _f(out[2]))
    return [o[:rows].copy() for o in out]


def depth_to_cloud(img, width, height, step, encoding, K):
    """orc_depth_to_cloud: width*height sensor-frame points (NaN invalid)."""
    L = lib()
    buf = np.ascontiguousarray(np.frombuffer(memoryview(img).cast("B"), np.uint8))
    assert buf.size >= int(step) * int(height), "image too short"
    Km = np.ascontiguousarray(np.asarray(K, np.float32).reshape(4))
    n = int(width) * int(height)
    out = [np.empty(max(n, 1), np.float32) for _ in range(3)]
    L.orc_depth_to_cloud(buf.ctypes.data_as(u8p), int(width), int(height), int(step),
                         int(encoding), _f(Km), _f(out[0]), _f(out[1]), _f(out[2]))
    return [o[:n].copy() for o in out]


def viz_points(pos, robot, d_max, inv_res, center, step, max_pts):
    """orc_viz_points: (k, 4) float32 records (x, y, z, rgb) in sample order."""
    L = lib()
    pos = np.ascontiguousarray(np.asarray(pos, np.float32).reshape(-1, 3))
    r = _a3(robot, np.float32)
    n = len(pos)
    out = np.empty((max(min(n, int(max_pts)), 1), 4), np.float32)
    k = L.orc_viz_points(_f(pos), n, _f(r), float(d_max), float(inv_res),
                         1 if center else 0, int(step), int(max_pts), _f(out))
    return out[:k].copy()


def logit(p: float) -> float:
    return float(lib().orc_logit(float(p)))


def pos_to_global(p, inv_res, center=False) -> int:
    return lib().orc_pos_to_global(float(p), float(inv_res), 1 if center else 0)


def global_to_pos(g, res, center=False) -> float:
    return float(lib().orc_global_to_pos(int(g), float(res), 1 if center else 0))


def global_to_local(g, m, h) -> int:
    return lib().orc_global_to_local(int(g), int(m), int(h))


def local_to_global(l, m, h, o) -> int:
    return lib().orc_local_to_global(int(l), int(m), int(h), int(o))


def local_to_hash(l, m, h) -> int:
    return lib().orc_local_to_hash(_i(_a3(l, np.int32)), _i(_a3(m, np.int32)),
                                   _i(_a3(h, np.int32)))


def hash_to_local(hid, m, h):
    out = np.zeros(3, np.int32)
    lib().orc_hash_to_local(int(hid), _i(_a3(m, np.int32)), _i(_a3(h, np.int32)),
                            _i(out))
    return tuple(out.tolist())


def global_to_hash(g, m, h) -> int:
    return lib().orc_global_to_hash(_i(_a3(g, np.int32)), _i(_a3(m, np.int32)),
                                    _i(_a3(h, np.int32)))
