"""cublox_amd — MI355X-native occupancy-update path.

Hand-written HIP kernels for gfx950 behind a C ABI (include/cublox/capi.h),
a drop-in C++ OccupancyGrid/Grid API (include/cublox/*.hpp), and this thin
Python layer over both. Importing the package loads the built libraries and
raises if they are missing: there is no CPU fallback.
"""
from . import _ffi
from .batch import GridBatch
from .device import DeviceBuffer, Event, Stream, device_count, device_name, synchronize
from .grid import FREE, OCCUPIED, UNKNOWN, OccupancyGrid, logit, replay
from .server import MapServer
from .scene import Scan, Scene
from .workloads import ALL as WORKLOADS
from .workloads import log_odds

# Fail at import time when the native libraries are not built.
_ffi.hip()
_ffi.host()
_ffi.seq()
_ffi.gen()

__all__ = [
    "OccupancyGrid", "GridBatch", "MapServer", "Scene", "Scan", "DeviceBuffer", "Stream", "Event",
    "device_count", "device_name", "synchronize", "logit", "log_odds", "replay",
    "WORKLOADS", "UNKNOWN", "FREE", "OCCUPIED",
]
