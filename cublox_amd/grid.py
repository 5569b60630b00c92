"""Python mirror of the drop-in C++ `cublox::OccupancyGrid`.

Each method calls the C++ class through its flat C binding (og_capi.cpp),
which calls the HIP kernels through include/cublox/capi.h. Method names
follow the reference API (include/cublox/OccupancyGrid.hpp:29-169 of the
reference) in snake_case.
"""
from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _ffi

UNKNOWN, FREE, OCCUPIED = 0, 1, 2


def logit(p: float) -> float:
    """Probability -> log-odds in float32 (utils.hpp:185-188 of the reference)."""
    return float(_ffi.host().cublox_logit(float(p)))


@dataclass
class Geometry:
    map_size: Tuple[int, int, int]
    half: Tuple[int, int, int]
    origin_i: Tuple[int, int, int]
    origin_f: Tuple[float, float, float]
    bound_min_i: Tuple[int, int, int]
    bound_max_i: Tuple[int, int, int]
    voxel_num: int
    resolution: float
    resolution_inv: float


def _f3(v) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(3))


class OccupancyGrid:
    def __init__(self, half_map_size: Sequence[int], resolution: float,
                 origin_at_center: bool = False,
                 recenter_threshold: Optional[float] = None,
                 origin: Sequence[float] = (0.0, 0.0, 0.0)):
        L = _ffi.host()
        o = _f3(origin)
        h = L.cublox_og_create(int(half_map_size[0]), int(half_map_size[1]),
                               int(half_map_size[2]), float(resolution),
                               1 if origin_at_center else 0,
                               0 if recenter_threshold is None else 1,
                               float(recenter_threshold or 0.0), _ffi.fptr(o))
        if not h:
            raise RuntimeError(L.cublox_og_last_error().decode())
        self._h = h
        self._L = L

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.cublox_og_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st: int) -> None:
        if st != 0:
            raise RuntimeError(self._L.cublox_og_last_error().decode())

    # ---- configuration --------------------------------------------------
    def set_max_raycast_range(self, r: float) -> None:
        self._check(self._L.cublox_og_set_max_range(self._h, float(r)))

    def set_log_odds_params(self, l_hit, l_miss, l_min, l_max, l_free,
                            l_occupied) -> None:
        self._check(self._L.cublox_og_set_log_odds(
            self._h, float(l_hit), float(l_miss), float(l_min), float(l_max),
            float(l_free), float(l_occupied)))

    @property
    def geometry(self) -> Geometry:
        m, hf, oi, bmn, bmx = (np.zeros(3, np.int32) for _ in range(5))
        of = np.zeros(3, np.float32)
        V = C.c_int(0)
        r = C.c_float(0)
        ri = C.c_float(0)
        self._check(self._L.cublox_og_geometry(
            self._h, _ffi.iptr(m), _ffi.iptr(hf), _ffi.iptr(oi), _ffi.fptr(of),
            _ffi.iptr(bmn), _ffi.iptr(bmx), C.byref(V), C.byref(r), C.byref(ri)))
        return Geometry(tuple(m.tolist()), tuple(hf.tolist()), tuple(oi.tolist()),
                        tuple(of.tolist()), tuple(bmn.tolist()),
                        tuple(bmx.tolist()), V.value, r.value, ri.value)

    @property
    def voxel_num(self) -> int:
        return self.geometry.voxel_num

    # ---- update ---------------------------------------------------------
    @staticmethod
    def _columns(points) -> Tuple[np.ndarray, np.ndarray, np.ndarray, int]:
        if isinstance(points, (tuple, list)) and len(points) == 3:
            xs, ys, zs = (np.ascontiguousarray(np.asarray(c, np.float32))
                          for c in points)
        else:
            p = np.asarray(points, dtype=np.float32)
            if p.ndim != 2 or p.shape[1] != 3:
                raise ValueError("points must be (N, 3) or a tuple of 3 columns")
            xs, ys, zs = (np.ascontiguousarray(p[:, k]) for k in range(3))
        if not (len(xs) == len(ys) == len(zs)):
            raise ValueError("column lengths differ")
        return xs, ys, zs, len(xs)

    def update(self, points, sensor_origin) -> None:
        """World-frame cloud from `sensor_origin` (OccupancyGrid::update)."""
        xs, ys, zs, n = self._columns(points)
        o = _f3(sensor_origin)
        self._check(self._L.cublox_og_update(self._h, _ffi.fptr(xs), _ffi.fptr(ys),
                                             _ffi.fptr(zs), n, _ffi.fptr(o)))

    def update_sensor_frame(self, points, T) -> None:
        """Sensor-frame cloud with pose T (3x4 [R|t]); transform fused."""
        xs, ys, zs, n = self._columns(points)
        Tm = np.ascontiguousarray(np.asarray(T, np.float32).reshape(12))
        self._check(self._L.cublox_og_update_sensor(
            self._h, _ffi.fptr(xs), _ffi.fptr(ys), _ffi.fptr(zs), n, _ffi.fptr(Tm)))

    def update_sensor_frame_device(self, d_xs: int, d_ys: int, d_zs: int,
                                   n: int, T) -> None:
        Tm = np.ascontiguousarray(np.asarray(T, np.float32).reshape(12))
        self._check(self._L.cublox_og_update_sensor_device(
            self._h, d_xs, d_ys, d_zs, int(n), _ffi.fptr(Tm)))

    def update_device(self, d_x: int, d_y: int, d_z: int, n: int,
                      sensor_origin) -> None:
        o = _f3(sensor_origin)
        self._check(self._L.cublox_og_update_device(self._h, d_x, d_y, d_z,
                                                    int(n), _ffi.fptr(o)))

    def cast_rays_device(self, d_x: int, d_y: int, d_z: int, n: int,
                         sensor_origin) -> None:
        o = _f3(sensor_origin)
        self._check(self._L.cublox_og_cast_rays_device(self._h, d_x, d_y, d_z,
                                                       int(n), _ffi.fptr(o)))

    def cast_rays_sensor_frame_device(self, d_xs: int, d_ys: int, d_zs: int,
                                      n: int, T) -> None:
        Tm = np.ascontiguousarray(np.asarray(T, np.float32).reshape(12))
        self._check(self._L.cublox_og_cast_rays_sensor_device(
            self._h, d_xs, d_ys, d_zs, int(n), _ffi.fptr(Tm)))

    FUSED_TILES, PER_VOXEL_ATOMICS = 0, 1

    def set_update_path(self, path: int) -> None:
        self._check(self._L.cublox_og_set_update_path(self._h, int(path)))

    def set_near_steps(self, k: int) -> None:
        self._check(self._L.cublox_og_set_near_steps(self._h, int(k)))

    # ---- driver ingest (reference main.cpp:558-693, on the device) ------
    def update_point_cloud2(self, data, width: int, height: int,
                            point_step: int, offsets=(0, 4, 8), T=None,
                            device: bool = False,
                            nbytes: Optional[int] = None) -> bool:
        """A PointCloud2 xyz payload (host bytes, or a device pointer with
        device=True) in the sensor frame: the first 1,000,000 records,
        finite and within the max raycast range, transformed by T and
        integrated. False (nothing updated) when no point is kept.

        The payload must hold min(width * height, 1,000,000) * point_step
        bytes (ValueError otherwise). A host payload's length is its own;
        a device payload's is `nbytes`, or taken as exactly that many bytes
        when `nbytes` is None."""
        lay = (C.c_uint * 6)(int(width), int(height), int(point_step),
                             int(offsets[0]), int(offsets[1]), int(offsets[2]))
        need = min(int(width) * int(height), 1_000_000) * int(point_step)
        Tm = np.ascontiguousarray(np.asarray(T, np.float32).reshape(12))
        upd = C.c_int(0)
        if device:
            ptr, keep = C.c_void_p(int(data)), None
            size = need if nbytes is None else int(nbytes)
        else:
            keep = np.ascontiguousarray(np.frombuffer(memoryview(data).cast("B"), np.uint8))
            ptr = C.c_void_p(keep.ctypes.data)
            size = keep.nbytes
        if size < need:
            raise ValueError(f"PointCloud2 payload of {size} bytes is shorter than the "
                             f"{need} bytes its width, height and point step describe")
        self._check(self._L.cublox_og_update_cloud2(
            self._h, ptr, 1 if device else 0, size, lay, _ffi.fptr(Tm), C.byref(upd)))
        return bool(upd.value)

    def update_depth_image(self, image, K, T, *, width: int = 0, height: int = 0,
                           step: int = 0, encoding: Optional[int] = None,
                           device: bool = False) -> bool:
        """A depth image (uint16 millimetres or float32 metres; a 2-D numpy
        array, or a device pointer with width/height/step/encoding given)
        back-projected with K = (fx, fy, cx, cy) as the driver does, then
        update_point_cloud2 of the result."""
        if device:
            ptr, keep = C.c_void_p(int(image)), None
            if encoding is None or not width or not height or not step:
                raise ValueError("a device image needs width, height, step and encoding")
        else:
            keep = np.ascontiguousarray(image)
            if keep.dtype == np.uint16:
                enc = 0
            elif keep.dtype == np.float32:
                enc = 1
            else:
                raise ValueError("depth images are uint16 (mm) or float32 (m)")
            encoding = enc if encoding is None else encoding
            height, width = keep.shape
            step = keep.strides[0]
            ptr = C.c_void_p(keep.ctypes.data)
        Km = np.ascontiguousarray(np.asarray(K, np.float32).reshape(4))
        Tm = np.ascontiguousarray(np.asarray(T, np.float32).reshape(12))
        upd = C.c_int(0)
        self._check(self._L.cublox_og_update_depth(
            self._h, ptr, 1 if device else 0, int(width), int(height), int(step),
            int(encoding), _ffi.fptr(Km), _ffi.fptr(Tm), C.byref(upd)))
        return bool(upd.value)

    def occupancy_viz_cloud(self, robot_pos, max_viz_radius: float,
                            subsample: int = 1, max_points: int = 100000):
        """The driver's viz cloud: ((k, 4) float32 records x, y, z, rgb in
        unspecified order, the driver's fetched sample count)."""
        cap = max(int(max_points), 1) if max_points > 0 else 1 << 22
        out = np.empty((cap, 4), np.float32)
        wr = C.c_uint(0)
        fetched = C.c_ulonglong(0)
        self._check(self._L.cublox_og_viz_cloud(
            self._h, _ffi.fptr(_f3(robot_pos)), float(max_viz_radius), int(subsample),
            int(max_points), _ffi.fptr(out), cap, C.byref(wr), C.byref(fetched)))
        return out[:wr.value].copy(), int(fetched.value)

    # ---- map snapshots (format: include/cublox/snapshot.h) ---------------
    def save_snapshot(self) -> np.ndarray:
        """The snapshot as a uint8 array viewing the library's buffer (freed
        with the array)."""
        out = C.c_void_p(None)
        n = C.c_size_t(0)
        self._check(self._L.cublox_og_save_snapshot(self._h, C.byref(out), C.byref(n)))
        if not n.value:
            self._L.cublox_og_free_buffer(out)
            return np.zeros(0, np.uint8)
        arr = np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_uint8)), (n.value,))
        weakref.finalize(arr, self._L.cublox_og_free_buffer, C.c_void_p(out.value))
        return arr

    def load_snapshot(self, data: bytes) -> None:
        buf = np.frombuffer(data, np.uint8)
        self._check(self._L.cublox_og_load_snapshot(
            self._h, C.c_void_p(buf.ctypes.data), buf.size))

    def save_snapshot_file(self, path: str) -> None:
        self._check(self._L.cublox_og_save_snapshot_file(self._h, str(path).encode()))

    def load_snapshot_file(self, path: str) -> None:
        self._check(self._L.cublox_og_load_snapshot_file(self._h, str(path).encode()))

    def set_changed_capacity(self, capacity: int) -> None:
        """Record (hash id, new log-odds) of every voxel whose value an
        update changes, at most `capacity` per update (0: off)."""
        self._check(self._L.cublox_og_set_changed_capacity(self._h, int(capacity)))
        self._changed_cap = int(capacity)

    def changed_voxels(self):
        """(ids int32, values float32, count) of the last update; count is
        how many voxels changed (more than len(ids) when truncated)."""
        cap = getattr(self, "_changed_cap", 0)
        ids = np.zeros(max(cap, 1), np.int32)
        vals = np.zeros(max(cap, 1), np.float32)
        cnt = C.c_ulonglong(0)
        wr = C.c_uint(0)
        self._check(self._L.cublox_og_changed_voxels(
            self._h, ids.ctypes.data_as(_ffi.i32p), vals.ctypes.data_as(_ffi.f32p),
            cap, C.byref(cnt), C.byref(wr)))
        return ids[:wr.value].copy(), vals[:wr.value].copy(), int(cnt.value)

    def cast_rays_fused_sensor_frame_device(self, d_xs: int, d_ys: int,
                                            d_zs: int, n: int, T) -> None:
        """Fused kernels, counts exported to the counter buffers."""
        Tm = np.ascontiguousarray(np.asarray(T, np.float32).reshape(12))
        self._check(self._L.cublox_og_cast_rays_fused_sensor_device(
            self._h, d_xs, d_ys, d_zs, int(n), _ffi.fptr(Tm)))

    def counter_census(self) -> dict:
        """Pending-counter work: unique voxels, visits, hits, hit voxels."""
        out = (C.c_ulonglong * 4)()
        self._check(self._L.cublox_og_census(self._h, out))
        return dict(touched=int(out[0]), visits=int(out[1]), hits=int(out[2]),
                    hit_voxels=int(out[3]))

    def apply_counters(self) -> None:
        self._check(self._L.cublox_og_apply_counters(self._h))

    # ---- recenter / reset -------------------------------------------------
    def recenter(self, pos) -> bool:
        p = _f3(pos)
        moved = C.c_int(0)
        self._check(self._L.cublox_og_recenter(self._h, _ffi.fptr(p), C.byref(moved)))
        return bool(moved.value)

    def reset(self) -> None:
        self._check(self._L.cublox_og_reset(self._h))

    def reset_voxel(self, hash_id: int) -> None:
        self._check(self._L.cublox_og_reset_voxel(self._h, int(hash_id)))

    # ---- reads ------------------------------------------------------------
    def occupancy(self) -> np.ndarray:
        out = np.empty(self.voxel_num, np.float32)
        self._check(self._L.cublox_og_occ_to_host(self._h, _ffi.fptr(out)))
        return out

    def counters(self) -> Tuple[np.ndarray, np.ndarray]:
        V = self.voxel_num
        op = np.empty(V, np.int32)
        hit = np.empty(V, np.int32)
        self._check(self._L.cublox_og_counters_to_host(self._h, _ffi.iptr(op),
                                                       _ffi.iptr(hit)))
        return op, hit

    def load_occupancy(self, occ: np.ndarray) -> None:
        occ = np.ascontiguousarray(occ, np.float32)
        assert occ.size == self.voxel_num
        self._check(self._L.cublox_og_load_occ(self._h, _ffi.fptr(occ)))

    def fetch_occupancy_around(self, center, radius: float,
                               state_filter: Optional[int] = None,
                               max_results: int = 0,
                               cap: Optional[int] = None):
        """Returns (positions (n,3) float32, states (n,) uint8)."""
        c = _f3(center)
        if cap is None:
            r_vox = int(np.ceil(radius * self.geometry.resolution_inv)) + 1
            est = (2 * r_vox + 1) ** 2 * (2 * r_vox + 1)
            cap = min(est, max_results if max_results > 0 else 500000)
        cap = max(int(cap), 1)
        pos = np.empty((cap, 3), np.float32)
        st = np.empty(cap, np.uint8)
        n = C.c_int(0)
        self._check(self._L.cublox_og_fetch(
            self._h, _ffi.fptr(c), float(radius),
            -1 if state_filter is None else int(state_filter), int(max_results),
            _ffi.fptr(pos), _ffi.bptr(st), cap, C.byref(n)))
        k = min(n.value, cap)
        return pos[:k].copy(), st[:k].copy()

    def fetch_occupancy_batch(self, centers, radius: float,
                              state_filter: Optional[int], per_query_cap: int):
        """Returns (counts (q,), positions (total, 3), states (total,)) with
        the samples of query 0 first, then query 1, ... (offsets = cumsum)."""
        cs = np.ascontiguousarray(np.asarray(centers, np.float32).reshape(-1, 3))
        q = cs.shape[0]
        counts = np.zeros(q, np.uint32)
        # The binding writes at most q * per_query_cap samples: output
        # arrays of that size are kept between calls (they only grow), and
        # the filled part is returned as copies.
        need = max(q * int(per_query_cap), 1)
        buf = getattr(self, "_fetch_out", None)
        if buf is None or buf[1].shape[0] < need:
            buf = (np.empty((need, 3), np.float32), np.empty(need, np.uint8))
            self._fetch_out = buf
        pos, st = buf
        self._check(self._L.cublox_og_fetch_batch(
            self._h, _ffi.fptr(cs), q, float(radius),
            -1 if state_filter is None else int(state_filter), int(per_query_cap),
            _ffi.uptr(counts), _ffi.fptr(pos), _ffi.bptr(st)))
        total = int(counts.sum())
        return counts, pos[:total].copy(), st[:total].copy()

    def _query(self, kind: int, by: int, q) -> bool:
        qf = np.zeros(3, np.float32)
        qi = np.zeros(3, np.int32)
        if by == 0:
            qf[:] = np.asarray(q, np.float32).reshape(3)
        elif by == 1:
            qi[:] = np.asarray(q, np.int32).reshape(3)
        else:
            qi[0] = int(q)
        r = C.c_int(0)
        self._check(self._L.cublox_og_query(self._h, kind, by, _ffi.fptr(qf),
                                            _ffi.iptr(qi), C.byref(r)))
        return bool(r.value)

    def is_occupied(self, pos=None, *, index=None, hash_id=None) -> bool:
        return self._query(0, *self._which(pos, index, hash_id))

    def is_free(self, pos=None, *, index=None, hash_id=None) -> bool:
        return self._query(1, *self._which(pos, index, hash_id))

    def is_unknown(self, pos=None, *, index=None, hash_id=None) -> bool:
        return self._query(2, *self._which(pos, index, hash_id))

    @staticmethod
    def _which(pos, index, hash_id):
        if pos is not None:
            return 0, pos
        if index is not None:
            return 1, index
        return 2, hash_id

    def max_horizontal_in_window_radius(self, center) -> float:
        c = _f3(center)
        r = C.c_float(0)
        self._check(self._L.cublox_og_max_horizontal_radius(self._h, _ffi.fptr(c),
                                                            C.byref(r)))
        return float(r.value)

    def enable_kernel_timing(self, on: bool = True) -> None:
        self._check(self._L.cublox_og_enable_timing(self._h, 1 if on else 0))

    def reset_kernel_times(self) -> None:
        self._check(self._L.cublox_og_reset_kernel_times(self._h))

    def kernel_times(self) -> dict:
        """Device ms per kernel (HIP events on this grid's stream)."""
        ms = (C.c_double * 4)()
        n = (C.c_longlong * 7)()
        self._check(self._L.cublox_og_kernel_times(self._h, ms, n))
        return dict(update_ms=ms[0], apply_ms=ms[1], recenter_ms=ms[2],
                    fetch_ms=ms[3], update_launches=n[0], apply_launches=n[1],
                    recenter_launches=n[2], recenter_zeroed_voxels=n[3],
                    fetch_launches=n[4], fetch_tested_voxels=n[5], fetch_samples=n[6])

    @property
    def device_occupancy_ptr(self) -> int:
        return int(self._L.cublox_og_device_occ(self._h) or 0)

    @property
    def stream(self):
        return self._L.cublox_og_stream(self._h)


class Replay:
    """A prepared sequence of driver steps for cublox_seq_replay
    (include/cublox/seq_replay.h): per step the blocking update of `grid`
    with n sensor-frame points (pose T) and the recenter at `base`, run in
    native code. Host columns (on_host) must stay alive while the replay
    runs; this object keeps the arrays it was given."""

    def __init__(self, steps, on_host: bool = False):
        self._keep = []
        arr = (_ffi.SeqStep * max(len(steps), 1))()
        for k, (grid, xs, ys, zs, n, T, base) in enumerate(steps):
            st = arr[k]
            st.grid = grid._h
            if on_host:
                cols = [np.ascontiguousarray(np.asarray(c, np.float32)) for c in (xs, ys, zs)]
                self._keep.append(cols)
                st.xs, st.ys, st.zs = (c.ctypes.data for c in cols)
            else:
                st.xs, st.ys, st.zs = int(xs), int(ys), int(zs)
            st.n = int(n)
            st.T[:] = [float(v) for v in np.asarray(T, np.float32).reshape(12)]
            st.base[:] = [float(v) for v in _f3(base)]
        self._arr = arr
        self._count = len(steps)
        self._on_host = 1 if on_host else 0
        self.step_ms = np.zeros(max(self._count, 1), np.float64)
        self.moved = np.zeros(max(self._count, 1), np.int32)

    def run(self):
        """Runs the steps; returns (per-step wall ms, per-step moved flags)."""
        st = _ffi.seq().cublox_seq_replay(
            self._arr, self._count, self._on_host,
            self.step_ms.ctypes.data_as(C.POINTER(C.c_double)), _ffi.iptr(self.moved))
        _ffi.check(st, _ffi.host())
        return list(self.step_ms[:self._count]), [bool(m) for m in self.moved[:self._count]]


def replay(steps, on_host: bool = False):
    """Runs driver steps (grid, xs, ys, zs, n, T, base) natively; returns
    (per-step wall ms, moved flags). Prepare a Replay once to keep the
    argument packing out of a timed region."""
    return Replay(steps, on_host).run()
