"""Batched updates of many grids (the C++ cublox::GridBatch).

One tick runs the fused update of every listed scan on its grid and then
the listed recenters, as one launch sequence each for all grids, with an
optional HIP-graph replay and an online autotuner. Each grid ends a tick
exactly as its own update + recenter would leave it.
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import _ffi


class PreparedTick:
    """A tick's packed arguments (GridBatch.prepare)."""

    def __init__(self, args: tuple, keep: tuple):
        self.args = args
        self._keep = keep


class GridBatch:
    def __init__(self, *, max_grids: int, max_points: int, max_tiles: int,
                 max_segments_per_ray: int, max_near_steps: int = 12,
                 min_chunk: int = 256, max_slabs: int = 65536,
                 max_regions: int = 0, changed_voxels: bool = False):
        L = _ffi.host()
        lim = (C.c_int * 9)(int(max_grids), int(max_points), int(max_tiles),
                            int(max_segments_per_ray), int(max_near_steps),
                            int(min_chunk), int(max_slabs), int(max_regions),
                            int(bool(changed_voxels)))
        h = L.cublox_gb_create(lim)
        if not h:
            raise RuntimeError(L.cublox_og_last_error().decode())
        self._h = h
        self._L = L
        self.max_grids = int(max_grids)

    @staticmethod
    def segment_bound(grid) -> int:
        return _ffi.host().cublox_gb_segment_bound(grid._h)

    @staticmethod
    def tile_count(grid) -> int:
        return _ffi.host().cublox_gb_tile_count(grid._h)

    @classmethod
    def for_grids(cls, grids: Sequence, max_points: int, **kw) -> "GridBatch":
        """Limits that fit one tick over `grids` with max_points in total."""
        kw.setdefault("max_tiles", sum(cls.tile_count(g) for g in grids))
        kw.setdefault("max_segments_per_ray",
                      max(cls.segment_bound(g) for g in grids))
        return cls(max_grids=len(grids), max_points=max_points, **kw)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.cublox_gb_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int) -> None:
        if st != 0:
            raise RuntimeError(self._L.cublox_og_last_error().decode())

    def tick(self, scans: Sequence[Tuple], recenters: Iterable[Tuple] = ()) -> None:
        """scans: (grid, d_x, d_y, d_z, n, T 3x4 sensor->world[, point
        stride[, range drop squared]]) with device buffers (objects with
        .ptr or ints); recenters: (grid, pos). Blocks until done."""
        self._tick(self._L.cublox_gb_tick, scans, recenters)

    def submit(self, scans: Sequence[Tuple], recenters: Iterable[Tuple] = ()) -> None:
        """tick() without the wait: returns once the tick is launched (when
        pipelined, after the tick submitted two before it has finished).
        The scans' device buffers must live until wait()."""
        self._tick(self._L.cublox_gb_submit, scans, recenters)

    def wait(self) -> None:
        """Blocks until every submitted tick has finished."""
        self._check(self._L.cublox_gb_wait(self._h))

    def set_pipelined(self, on: bool) -> None:
        """Overlap consecutive ticks (a second stream and workspace)."""
        self._check(self._L.cublox_gb_set_pipelined(self._h, 1 if on else 0))
        self._pipelined = bool(on)

    @property
    def pipelined(self) -> bool:
        return getattr(self, "_pipelined", False)

    def prepare(self, scans: Sequence[Tuple], recenters: Iterable[Tuple] = ()) -> "PreparedTick":
        """The arguments of a tick packed once, for tick_prepared /
        submit_prepared (a caller that replays known ticks keeps the packing
        out of its loop). The scans' device buffers must outlive its use."""
        return PreparedTick(*self._pack(scans, recenters))

    def tick_prepared(self, p: "PreparedTick") -> None:
        """tick() with arguments packed by prepare()."""
        self._check(self._L.cublox_gb_tick(self._h, *p.args))

    def submit_prepared(self, p: "PreparedTick") -> None:
        """submit() with arguments packed by prepare()."""
        self._check(self._L.cublox_gb_submit(self._h, *p.args))

    def _tick(self, fn, scans: Sequence[Tuple], recenters: Iterable[Tuple]) -> None:
        args, _keep = self._pack(scans, recenters)
        self._check(fn(self._h, *args))

    @staticmethod
    def _pack(scans: Sequence[Tuple], recenters: Iterable[Tuple]) -> tuple:
        scans = list(scans)
        recenters = list(recenters)
        k = len(scans)
        P = C.c_void_p * max(k, 1)
        ptr = lambda b: b if isinstance(b, int) else b.ptr
        grids = P(*[s[0]._h for s in scans])
        xs = P(*[ptr(s[1]) for s in scans])
        ys = P(*[ptr(s[2]) for s in scans])
        zs = P(*[ptr(s[3]) for s in scans])
        n = np.array([int(s[4]) for s in scans] or [0], np.int32)
        T = np.ascontiguousarray(
            np.array([np.asarray(s[5], np.float32).reshape(12) for s in scans]
                     or [np.zeros(12, np.float32)], np.float32))
        stride = np.array([int(s[6]) if len(s) > 6 else 0 for s in scans] or [0], np.int32)
        drop = np.array([float(s[7]) if len(s) > 7 else 0.0 for s in scans] or [0.0],
                        np.float32)
        R = C.c_void_p * max(len(recenters), 1)
        rg = R(*[r[0]._h for r in recenters])
        rp = np.ascontiguousarray(
            np.array([np.asarray(r[1], np.float32).reshape(3) for r in recenters]
                     or [np.zeros(3, np.float32)], np.float32))
        args = (k, grids, xs, ys, zs, _ffi.iptr(n), _ffi.fptr(T), _ffi.iptr(stride),
                _ffi.fptr(drop), len(recenters), rg, _ffi.fptr(rp))
        # The arrays behind the pointers, kept alive beside the arguments.
        return args, (n, T, stride, drop, rp)

    def set_graph_mode(self, on: bool) -> None:
        self._check(self._L.cublox_gb_set_graph_mode(self._h, 1 if on else 0))

    @staticmethod
    def _setting4(c) -> Tuple[int, int, int, int]:
        """(near_steps, chunk[, lanes per ray[, hot-tile mode]]), the last
        two 0 (automatic lanes, chunk slabs) when left out."""
        c = tuple(int(v) for v in c)
        return (c + (0, 0))[:4] if len(c) >= 2 else (8, 1024, 0, 0)

    def set_tuning(self, candidates: Sequence[Tuple[int, ...]], ticks_each: int) -> None:
        arr = np.array([v for c in candidates for v in self._setting4(c)] or [0], np.int32)
        self._check(self._L.cublox_gb_set_tuning(self._h, len(candidates),
                                                 _ffi.iptr(arr), int(ticks_each)))

    @property
    def tuning(self) -> bool:
        v = C.c_int(0)
        self._check(self._L.cublox_gb_tuning(self._h, C.byref(v)))
        return bool(v.value)

    @property
    def setting(self) -> Tuple[int, int, int, int]:
        """(near_steps, chunk, lanes per ray, hot-tile mode)."""
        out = np.zeros(4, np.int32)
        self._check(self._L.cublox_gb_setting(self._h, _ffi.iptr(out)))
        return tuple(int(v) for v in out)

    def set_setting(self, near_steps: int, chunk: int, lanes: int = 0,
                    hot_mode: int = 0) -> None:
        self._check(self._L.cublox_gb_set_setting(self._h, int(near_steps), int(chunk),
                                                  int(lanes), int(hot_mode)))

    def tune_results(self, n: int) -> List[float]:
        out = (C.c_double * max(n, 1))()
        self._check(self._L.cublox_gb_tune_results(self._h, out, n))
        return [out[i] for i in range(n)]

    @property
    def last(self) -> Tuple[float, float, int]:
        """(update ms, recenter ms, zeroed voxels) of the last tick; in graph
        mode the whole tick counts as update."""
        out = (C.c_double * 3)()
        self._check(self._L.cublox_gb_last(self._h, out))
        return out[0], out[1], int(out[2])

    def counters(self) -> dict:
        """Counters of the last tick's update: work items, tile records,
        overflow flag, split and unsplit hot tiles."""
        out = (C.c_int * 5)()
        self._check(self._L.cublox_gb_counters(self._h, out))
        return dict(items=out[0], records=out[1], overflow=out[2],
                    split_tiles=out[3], unsplit_hot=out[4])
