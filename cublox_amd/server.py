"""Occupancy map server (the C++ cublox::MapServer).

Many sessions -- each an occupancy grid of its own shape -- share the GPU.
A session is admitted while its device bytes fit the HBM budget (else it
waits, FIFO); a worker thread forms each tick from the requests ready at
that moment (one scan per session, continuous batching) and runs it as one
batched update + recenter, copying the next tick's points while the current
one runs. Fetches run after every earlier scan of their session. Each
session's grid evolves exactly as the same requests issued alone would
leave it.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from . import _ffi

SCAN, FETCH = 0, 1


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


class MapServer:
    def __init__(self, *, hbm_budget_bytes: int = 200 << 30, max_sessions: int = 256,
                 max_points_per_scan: int = 131072, max_points_per_tick: int = 1 << 20,
                 max_grids_per_tick: int = 256, max_tiles_per_tick: int = 1 << 21,
                 max_segments_per_ray: int = 96, graph: bool = True,
                 point_arena_bytes: int = 0, evict_idle: bool = False,
                 pipelined: bool = False, chunk: int = 0):
        L = _ffi.host()
        lim = (C.c_longlong * 12)(int(hbm_budget_bytes), int(max_sessions),
                                  int(max_points_per_scan), int(max_points_per_tick),
                                  int(max_grids_per_tick), int(max_tiles_per_tick),
                                  int(max_segments_per_ray), int(bool(graph)),
                                  int(point_arena_bytes), int(bool(evict_idle)),
                                  int(bool(pipelined)), int(chunk))
        h = L.cublox_ms_create(lim)
        if not h:
            raise RuntimeError(L.cublox_og_last_error().decode())
        self._h = h
        self._L = L
        self._shapes: Dict[int, Tuple[int, int, int]] = {}

    @staticmethod
    def session_bytes(half: Sequence[int]) -> int:
        return int(_ffi.host().cublox_ms_session_bytes((C.c_int * 3)(*map(int, half))))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.cublox_ms_destroy(self._h)
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

    def open_session(self, half: Sequence[int], resolution: float,
                     origin_at_center: bool = False, threshold: Optional[float] = None,
                     max_range: float = 25.0,
                     log_odds: Optional[Sequence[float]] = None) -> int:
        sid = C.c_int()
        lo = None if log_odds is None else (C.c_float * 6)(*map(float, log_odds))
        self._check(self._L.cublox_ms_open(
            self._h, (C.c_int * 3)(*map(int, half)), float(resolution),
            int(bool(origin_at_center)), int(threshold is not None),
            float(threshold or 0.0), float(max_range), lo, C.byref(sid)))
        self._shapes[sid.value] = tuple(2 * int(h) + 1 for h in half)
        return sid.value

    def close_session(self, sid: int) -> None:
        self._check(self._L.cublox_ms_close(self._h, int(sid)))

    def evict_session(self, sid: int) -> None:
        """Move the session's map to a host snapshot (restored on readmission)."""
        self._check(self._L.cublox_ms_evict(self._h, int(sid)))

    def admitted(self, sid: int) -> bool:
        out = C.c_int()
        self._check(self._L.cublox_ms_admitted(self._h, int(sid), C.byref(out)))
        return bool(out.value)

    def submit_scan(self, sid: int, xs, ys, zs, T, base) -> int:
        """Sensor-frame columns (copied), T 3x4 sensor->world; the grid then
        recenters toward `base`."""
        xs, ys, zs = _f32(xs), _f32(ys), _f32(zs)
        Tm = _f32(np.asarray(T).reshape(12))
        b = _f32(np.asarray(base).reshape(3))
        rid = C.c_longlong()
        self._check(self._L.cublox_ms_submit_scan(
            self._h, int(sid), xs.ctypes.data, ys.ctypes.data, zs.ctypes.data, int(xs.size),
            Tm.ctypes.data_as(_ffi.f32p), b.ctypes.data_as(_ffi.f32p), C.byref(rid)))
        return rid.value

    def submit_fetch(self, sid: int, center, radius: float,
                     state_filter: Optional[int] = None, max_results: int = 0) -> int:
        c = _f32(np.asarray(center).reshape(3))
        rid = C.c_longlong()
        self._check(self._L.cublox_ms_submit_fetch(
            self._h, int(sid), c.ctypes.data_as(_ffi.f32p), float(radius),
            -1 if state_filter is None else int(state_filter), int(max_results),
            C.byref(rid)))
        return rid.value

    def drain(self) -> None:
        self._check(self._L.cublox_ms_drain(self._h))

    def pause(self) -> None:
        """No new tick is formed until resume(); requests keep queueing."""
        self._check(self._L.cublox_ms_pause(self._h))

    def resume(self) -> None:
        self._check(self._L.cublox_ms_resume(self._h))

    def completions(self, cap: int = 1 << 16) -> np.ndarray:
        """All completions so far as a structured array: id, session, kind,
        submit_ms, start_ms, done_ms, tick, result."""
        parts = []
        while True:
            ids = np.empty(cap, np.int64)
            ses = np.empty(cap, np.int32)
            kinds = np.empty(cap, np.int32)
            times = np.empty(3 * cap, np.float64)
            ticks = np.empty(cap, np.int64)
            res = np.empty(cap, np.int64)
            n = C.c_int()
            self._check(self._L.cublox_ms_take_completions(
                self._h, ids.ctypes.data_as(C.POINTER(C.c_longlong)),
                ses.ctypes.data_as(_ffi.i32p), kinds.ctypes.data_as(_ffi.i32p),
                times.ctypes.data_as(C.POINTER(C.c_double)),
                ticks.ctypes.data_as(C.POINTER(C.c_longlong)),
                res.ctypes.data_as(C.POINTER(C.c_longlong)), cap, C.byref(n)))
            m = n.value
            a = np.empty(m, dtype=[("id", np.int64), ("session", np.int32),
                                   ("kind", np.int32), ("submit_ms", np.float64),
                                   ("start_ms", np.float64), ("done_ms", np.float64),
                                   ("tick", np.int64), ("result", np.int64)])
            a["id"], a["session"], a["kind"] = ids[:m], ses[:m], kinds[:m]
            t = times[:3 * m].reshape(m, 3)
            a["submit_ms"], a["start_ms"], a["done_ms"] = t[:, 0], t[:, 1], t[:, 2]
            a["tick"], a["result"] = ticks[:m], res[:m]
            parts.append(a)
            if m < cap:
                break
        return np.concatenate(parts)

    def run_load(self, sids: Sequence[int], templates: Sequence[Tuple], *,
                 scan_rate_hz: float, duration_s: float, fetch_ratio: float = 0.2,
                 fetch_radius: float = 1.5, fetch_radius_max: float = 0.0,
                 speed: float = 0.5, seed: int = 1, clients: int = 8) -> Tuple[int, float]:
        """Open-loop Poisson load from a native client thread (this call):
        templates[k] = (xs, ys, zs, T 3x4, base) of session sids[k], replayed
        with the pose moved `speed` m along x per scan; fetch radii uniform in
        [fetch_radius, fetch_radius_max]; `clients` submitting threads share
        the sessions. Returns (requests submitted, largest submission delay
        behind schedule in ms); drain() waits for the requests."""
        k = len(sids)
        cols = [tuple(_f32(c) for c in t[:3]) for t in templates]
        ptr = C.c_void_p * k
        xs = ptr(*[c[0].ctypes.data for c in cols])
        ys = ptr(*[c[1].ctypes.data for c in cols])
        zs = ptr(*[c[2].ctypes.data for c in cols])
        n = (C.c_int * k)(*[c[0].size for c in cols])
        T = _f32(np.concatenate([np.asarray(t[3], np.float32).reshape(12) for t in templates]))
        base = _f32(np.concatenate([np.asarray(t[4], np.float32).reshape(3) for t in templates]))

        class Cfg(C.Structure):
            _fields_ = [("scan_rate_hz", C.c_double), ("fetch_ratio", C.c_double),
                        ("duration_s", C.c_double), ("fetch_radius", C.c_float),
                        ("speed", C.c_float), ("seed", C.c_ulonglong),
                        ("fetch_radius_max", C.c_float), ("clients", C.c_int)]

        cfg = Cfg(float(scan_rate_hz), float(fetch_ratio), float(duration_s),
                  float(fetch_radius), float(speed), int(seed), float(fetch_radius_max),
                  int(clients))
        out = C.c_longlong()
        lag = C.c_double()
        self._check(self._L.cublox_ms_load_run(
            self._h, (C.c_int * k)(*map(int, sids)), k, C.cast(xs, C.c_void_p),
            C.cast(ys, C.c_void_p), C.cast(zs, C.c_void_p), n,
            T.ctypes.data_as(_ffi.f32p), base.ctypes.data_as(_ffi.f32p), C.byref(cfg),
            C.byref(out), C.byref(lag)))
        return out.value, lag.value

    def keep_fetch_results(self, on: bool) -> None:
        self._check(self._L.cublox_ms_keep_fetch_results(self._h, int(bool(on))))

    def fetch_result(self, rid: int, cap: int = 500000) -> Tuple[np.ndarray, np.ndarray]:
        """A kept fetch's samples: positions (n, 3) and states (n,)."""
        pos = np.empty((cap, 3), np.float32)
        st = np.empty(cap, np.uint8)
        n = C.c_int()
        self._check(self._L.cublox_ms_fetch_result(
            self._h, int(rid), pos.ctypes.data_as(_ffi.f32p), st.ctypes.data_as(_ffi.u8p),
            cap, C.byref(n)))
        m = min(n.value, cap)
        return pos[:m].copy(), st[:m].copy()

    def occupancy(self, sid: int) -> np.ndarray:
        """The session grid's occupancy in hash order (after drain)."""
        mx, my, mz = self._shapes[sid]
        out = np.empty(mx * my * mz, np.float32)
        self._check(self._L.cublox_ms_copy_occupancy(self._h, int(sid),
                                                     out.ctypes.data_as(_ffi.f32p)))
        return out

    def origin(self, sid: int) -> Tuple[int, int, int]:
        o = (C.c_int * 3)()
        self._check(self._L.cublox_ms_origin(self._h, int(sid), o))
        return o[0], o[1], o[2]

    def stats(self) -> dict:
        out = (C.c_double * 14)()
        self._check(self._L.cublox_ms_stats(self._h, out))
        keys = ("ticks", "scans", "fetches", "points", "device_ms", "resident_bytes",
                "workspace_bytes", "admitted", "waiting", "now_ms", "launch_ms",
                "evictions", "restores", "snapshot_bytes")
        return {k: (out[i] if k in ("device_ms", "now_ms", "launch_ms") else int(out[i]))
                for i, k in enumerate(keys)}
