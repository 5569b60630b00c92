"""The fused update that allocates nothing (cublox_launch_integrate_noalloc),
called through the C ABI, against the same update through
cublox_launch_integrate with a caller-owned workspace: bit-identical
occupancy on two calls in a row (the second with a smaller scan, so the
library's scratch array holds the first call's contents), and a clear
capacity error for a scan whose workspace exceeds the array.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cfg(ffi, w, origin_i, sensor):
    c = ffi.RaycastCfg()
    c.origin = ffi.Float3(*[float(v) for v in sensor])
    c.origin_i = ffi.Int3(*origin_i)
    m = w.map_size
    c.map_size_i = ffi.Int3(*m)
    c.half_map_size_i = ffi.Int3(*w.half)
    c.resolution = w.resolution
    c.inv_resolution = float(np.float32(1.0) / np.float32(w.resolution))
    c.max_range = w.max_range
    c.map_vox_num = m[0] * m[1] * m[2]
    c.origin_at_center = int(w.origin_at_center)
    return c


def test_noalloc_update_matches_workspace_update(cb, gpu_ready):
    from cublox_amd import _ffi
    L = _ffi.hip()
    w = cb.WORKLOADS["readme"]
    lo = cb.log_odds()
    scene = w.scene()
    scans = [scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
             for k in range(2)]
    T0 = np.asarray(scans[0].T, np.float32).reshape(3, 4)
    inv = np.float32(1.0) / np.float32(w.resolution)
    origin_i = [int(np.floor(np.float32(T0[a, 3]) * inv)) for a in range(3)]
    V = int(np.prod(w.map_size))
    occ_ref = cb.DeviceBuffer(4 * V)
    occ_na = cb.DeviceBuffer(4 * V)
    occ_ref.zero()
    occ_na.zero()
    n_max = scans[0].n
    cfg0 = _cfg(_ffi, w, origin_i, T0[:, 3])
    ws_bytes = L.cublox_integrate_workspace_bytes(C.byref(cfg0), n_max, 8)
    ws = cb.DeviceBuffer(ws_bytes)
    _ffi.check(L.cublox_integrate_workspace_init(ws.ptr, ws_bytes, None))
    for k, n in ((0, n_max), (1, 20000)):
        s = scans[k]
        T = np.ascontiguousarray(np.asarray(s.T, np.float32).reshape(12))
        cfg = _cfg(_ffi, w, origin_i, T.reshape(3, 4)[:, 3])
        cols = [cb.DeviceBuffer.from_array(np.ascontiguousarray(c[:n]))
                for c in (s.xs, s.ys, s.zs)]
        print(f"[noalloc] call {k}: n={n}", flush=True)
        _ffi.check(L.cublox_launch_integrate(
            C.byref(cfg), _ffi.fptr(T), cols[0].ptr, cols[1].ptr, cols[2].ptr, n,
            occ_ref.ptr, *lo[:4], 0.0, ws.ptr, ws_bytes, n_max, 8, None, None, None))
        _ffi.check(L.cublox_launch_integrate_noalloc(
            C.byref(cfg), _ffi.fptr(T), cols[0].ptr, cols[1].ptr, cols[2].ptr, n,
            occ_na.ptr, *lo[:4], 0.0, 8, None, None, None))
        a = occ_ref.download(np.uint32, V)
        b = occ_na.download(np.uint32, V)
        diff = int(np.count_nonzero(a != b))
        assert diff == 0, f"call {k}: {diff} voxels differ"
        assert np.count_nonzero(a) > 1_000_000

    # A scan whose workspace exceeds the scratch array: refused up front,
    # nothing enqueued (1401x1401x151 at 0.05 m, 60 m rays, 262,144 points).
    ws_ = cb.WORKLOADS["stress"]
    big = _cfg(_ffi, ws_, [0, 0, 0], (0.0, 0.0, 0.0))
    n_big = ws_.points
    dummy = cb.DeviceBuffer(4 * 1024)
    st = L.cublox_launch_integrate_noalloc(
        C.byref(big), None, dummy.ptr, dummy.ptr, dummy.ptr, n_big, dummy.ptr,
        *lo[:4], 0.0, 8, None, None, None)
    assert st == 3, st  # CUBLOX_ERR_CAPACITY
    assert b"scratch" in L.cublox_last_error()
