"""GPU parity of the grid lifecycle calls against the oracle:
  * resetVoxel (reference OccupancyGrid.cpp:128-139) zeroes one voxel and
    its counters; ids outside [0, V) are no-ops;
  * reset (OccupancyGrid.cpp:116-126) zeroes the grid and makes the next
    update derive the window origin from its sensor again, so a reset grid
    equals a fresh one;
  * the by-index clear (kernels.cu:273-296) zeroes the listed valid voxels
    and ignores the others, including duplicates and negative ids.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import assert_occ_bitexact, make_pair

pytestmark = pytest.mark.gpu


def _scan(w, scene, k):
    return scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)


def _run(g, o, w, scene, ks):
    for k in ks:
        s = _scan(w, scene, k)
        print(f"[lifecycle] scan {k} n={s.n}", flush=True)
        for x in (g, o):
            x.update(s.world(), s.sensor_origin)
            x.recenter(s.base)
        assert g.geometry.origin_i == o.origin()[0]
        assert_occ_bitexact(g.occupancy(), o.occupancy(), f"scan {k}")


def test_reset_voxel_matches_oracle(cb, orc, gpu_ready):
    w = cb.WORKLOADS["small"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w)
    _run(g, o, w, scene, range(3))
    occ = g.occupancy()
    V = occ.size
    hit = np.nonzero(occ > 0)[0]
    miss = np.nonzero(occ < 0)[0]
    assert hit.size and miss.size
    ids = [int(hit[0]), int(hit[-1]), int(miss[0]), int(miss[miss.size // 2]),
           0, V - 1, -1, V, V + 1000, -(2 ** 31)]
    for h in ids:
        g.reset_voxel(h)
        o.reset_voxel(h)
    got = g.occupancy()
    assert_occ_bitexact(got, o.occupancy(), "after reset_voxel")
    for h in ids[:6]:
        assert got[h] == 0.0
    op, hc = g.counters()
    assert not op.any() and not hc.any()
    # The grid keeps integrating from the edited state.
    _run(g, o, w, scene, range(3, 5))


def test_reset_matches_fresh_grid(cb, orc, gpu_ready):
    w = cb.WORKLOADS["small"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w)
    _run(g, o, w, scene, range(4))
    g.reset()
    o.reset()
    assert not g.occupancy().any()
    op, hc = g.counters()
    assert not op.any() and not hc.any()
    # The next update takes its origin from scan 9's sensor, as a fresh
    # grid's first update does.
    _run(g, o, w, scene, [9, 10])
    f, fo = make_pair(cb, orc, w)
    _run(f, fo, w, scene, [9, 10])
    assert f.geometry.origin_i == g.geometry.origin_i
    assert_occ_bitexact(g.occupancy(), f.occupancy(), "reset vs fresh")


def test_clear_voxels_by_index(cb, orc, gpu_ready):
    from cublox_amd import _ffi
    w = cb.WORKLOADS["small"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w)
    _run(g, o, w, scene, range(3))
    occ = o.occupancy()
    V = occ.size
    rng = np.random.default_rng(7)
    touched = np.nonzero(occ != 0)[0]
    valid = np.concatenate([rng.choice(touched, 500, replace=False),
                            rng.integers(0, V, 500), [0, V - 1]])
    idx = np.concatenate([valid, valid[:50],  # duplicates
                          [-1, -7, V, V + 3, 2 ** 31 - 1, -(2 ** 31)]]).astype(np.int32)
    rng.shuffle(idx)
    expect = occ.copy()
    expect[valid] = 0.0
    L = _ffi.hip()
    d_idx = cb.DeviceBuffer.from_array(idx)
    st = L.cublox_launch_clear_voxels_by_index(C.c_void_p(g.device_occupancy_ptr),
                                               C.c_void_p(d_idx.ptr), int(idx.size),
                                               int(V), g.stream)
    assert st == 0
    assert L.cublox_stream_synchronize(g.stream) == 0
    assert_occ_bitexact(g.occupancy(), expect, "by-index clear")
    # Nothing to do is not an error and changes nothing (kernels.cu:288-290).
    assert L.cublox_launch_clear_voxels_by_index(
        C.c_void_p(g.device_occupancy_ptr), C.c_void_p(d_idx.ptr), 0, int(V), g.stream) == 0
    assert L.cublox_launch_clear_voxels_by_index(
        C.c_void_p(g.device_occupancy_ptr), None, int(idx.size), int(V), g.stream) == 0
    assert L.cublox_stream_synchronize(g.stream) == 0
    assert_occ_bitexact(g.occupancy(), expect, "no-op clears")
