"""Seeded synthetic LiDAR workload (wraps libcublox_scangen.so)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _ffi


@dataclass
class Scan:
    xs: np.ndarray  # sensor frame, float32, azimuth-major
    ys: np.ndarray
    zs: np.ndarray
    T: np.ndarray  # (3, 4) float32 sensor -> world
    base: np.ndarray  # (3,) float32, robot base position (recenter target)
    n_miss: int

    @property
    def n(self) -> int:
        return len(self.xs)

    @property
    def sensor_origin(self) -> np.ndarray:
        return self.T[:, 3].copy()

    def world(self) -> np.ndarray:
        """Points in the world frame, (N, 3) float32, evaluated as
        ((R0*x + R1*y) + R2*z) + t per row in float32 (the fused-transform
        order)."""
        T = self.T
        out = np.empty((self.n, 3), np.float32)
        for r in range(3):
            a = T[r, 0] * self.xs
            a = a + T[r, 1] * self.ys
            a = a + T[r, 2] * self.zs
            out[:, r] = a + T[r, 3]
        return out


class Scene:
    def __init__(self, *, seed: int, n_boxes: int, box_extent: float,
                 speed: float, yaw_rate_deg: float = 2.0,
                 box_min_side: float = 0.5, box_max_side: float = 4.0,
                 corridor_half: float = 2.5, sensor_height: float = 1.8,
                 noise_sigma: float = 0.02, miss_extra: float = 1.0):
        L = _ffi.gen()
        cfg = _ffi.SceneCfg(int(n_boxes), float(box_extent), float(box_min_side),
                            float(box_max_side), float(corridor_half),
                            float(speed), float(yaw_rate_deg),
                            float(sensor_height), float(noise_sigma),
                            float(miss_extra), int(seed))
        self._h = L.cublox_scene_create(C.byref(cfg))
        if not self._h:
            raise RuntimeError("scene creation failed")
        self._L = L

    def __del__(self):
        try:
            if self._h:
                self._L.cublox_scene_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def boxes(self) -> np.ndarray:
        n = self._L.cublox_scene_box_count(self._h)
        out = np.zeros((n, 6), np.float64)
        if n:
            self._L.cublox_scene_boxes(self._h,
                                       out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def pose(self, k: int):
        T = np.zeros(12, np.float32)
        base = np.zeros(3, np.float32)
        self._L.cublox_scene_pose(self._h, int(k), _ffi.fptr(T), _ffi.fptr(base))
        return T.reshape(3, 4), base

    def scan(self, k: int, beams: int, azimuths: int, el_lo_deg: float,
             el_hi_deg: float, max_range: float, threads: int = 0) -> Scan:
        n = beams * azimuths
        xs = np.empty(n, np.float32)
        ys = np.empty(n, np.float32)
        zs = np.empty(n, np.float32)
        T = np.zeros(12, np.float32)
        base = np.zeros(3, np.float32)
        miss = C.c_int(0)
        got = self._L.cublox_scene_scan(self._h, int(k), int(beams), int(azimuths),
                                        float(el_lo_deg), float(el_hi_deg),
                                        float(max_range), _ffi.fptr(xs),
                                        _ffi.fptr(ys), _ffi.fptr(zs), _ffi.fptr(T),
                                        _ffi.fptr(base), C.byref(miss),
                                        int(threads))
        assert got == n
        return Scan(xs, ys, zs, T.reshape(3, 4), base, miss.value)

    def queries(self, k: int, n: int, spread: float) -> np.ndarray:
        out = np.zeros((n, 3), np.float32)
        self._L.cublox_scene_queries(self._h, int(k), int(n), float(spread),
                                     _ffi.fptr(out))
        return out


def uniform(seed: int, tag: int, j: int) -> float:
    """Counter-based uniform in (0, 1) shared with the C++ generator."""
    return float(_ffi.gen().cublox_rng_uniform(int(seed), int(tag), int(j)))
