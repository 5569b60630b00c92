"""A recenter returns before its slab clear has run (the clear is planned
on the host and rides in the next update's walk launch, whose extra
workgroups clear the slabs beside the walk; any other call launches it
first). Every later call must still see the cleared window, bit for bit as
the oracle (reference OccupancyGrid.cpp:141-179 then the next call):
  * update right after a recenter (the folds come after the clear);
  * two recenters in a row, then reads (fetch, point query, occupancy);
  * a recenter, then a batched tick of the same grid on another stream;
  * the per-voxel path and the stage-wise calls after a recenter.
"""
import numpy as np
import pytest

from helpers import assert_occ_bitexact, make_pair, sample_multiset

pytestmark = pytest.mark.gpu


def _scans(cb, w, k0, k1):
    scene = w.scene()
    return [scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
            for k in range(k0, k1)]


def test_update_after_recenter_readme(cb, orc, gpu_ready):
    """The readme shape: 10 scans with recenters."""
    w = cb.WORKLOADS["readme"]
    g, o = make_pair(cb, orc, w, threads=16)
    moves = 0
    for k, s in enumerate(_scans(cb, w, 0, 10)):
        print(f"[overlap] readme scan {k}", flush=True)
        g.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        o.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        moves += int(g.recenter(s.base))
        o.recenter(s.base)
    assert moves >= 3
    assert g.geometry.origin_i == o.origin()[0]
    assert_occ_bitexact(g.occupancy(), o.occupancy(), "readme after 10 scans")
    g.close()


def test_back_to_back_recenters_then_reads(cb, orc, gpu_ready):
    w = cb.WORKLOADS["small"]
    g, o = make_pair(cb, orc, w)
    scans = _scans(cb, w, 0, 3)
    for s in scans:
        g.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        o.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
    base = np.asarray(scans[-1].base, np.float32)
    for dx in (2.5, 5.0, 7.5):  # three clears queued without an update between
        p = base + np.array([dx, 0.3 * dx, 0.0], np.float32)
        assert g.recenter(p)
        assert o.recenter(p) != 0
    c = base + np.array([6.0, 1.0, 0.2], np.float32)
    gp, gs = g.fetch_occupancy_around(c, 2.0)
    op_, os_, on = o.fetch(c, 2.0)
    assert len(gs) == on
    assert np.array_equal(sample_multiset(gp, gs), sample_multiset(op_, os_))
    assert_occ_bitexact(g.occupancy(), o.occupancy(), "after three recenters")
    g.close()


def test_recenter_then_batch_tick(cb, orc, gpu_ready):
    w = cb.WORKLOADS["small"]
    g, o = make_pair(cb, orc, w)
    scans = _scans(cb, w, 0, 7)
    for s in scans[:5]:
        g.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        o.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        g.recenter(s.base)
        o.recenter(s.base)
    batch = cb.GridBatch.for_grids([g], 1 << 15, max_near_steps=12)
    for s in scans[5:]:
        bufs = [cb.DeviceBuffer.from_array(np.ascontiguousarray(c)) for c in (s.xs, s.ys, s.zs)]
        moved = g.recenter(np.asarray(s.base, np.float32) + np.float32(2.5))
        o.recenter(np.asarray(s.base, np.float32) + np.float32(2.5))
        assert moved
        batch.tick([(g, bufs[0], bufs[1], bufs[2], s.n, s.T)], [(g, s.base)])
        o.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        o.recenter(s.base)
        assert g.geometry.origin_i == o.origin()[0]
        assert_occ_bitexact(g.occupancy(), o.occupancy(), "batch after recenter")
    batch.close()
    g.close()


def test_per_voxel_and_stagewise_after_recenter(cb, orc, gpu_ready):
    w = cb.WORKLOADS["small"]
    g, o = make_pair(cb, orc, w)
    g.set_update_path(cb.OccupancyGrid.PER_VOXEL_ATOMICS)
    scans = _scans(cb, w, 0, 12)
    for k, s in enumerate(scans):
        if k == 6:
            g.set_update_path(cb.OccupancyGrid.FUSED_TILES)
        if k >= 9:  # stage-wise: cast, recenter, apply
            pts = s.world()
            bufs = [cb.DeviceBuffer.from_array(np.ascontiguousarray(pts[:, a]))
                    for a in range(3)]
            g.cast_rays_device(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, s.n, s.sensor_origin)
            o.raycast(pts, s.sensor_origin)
            g.recenter(s.base)
            o.recenter(s.base)
            g.apply_counters()
            o.apply()
        else:
            g.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
            o.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
            g.recenter(s.base)
            o.recenter(s.base)
        assert_occ_bitexact(g.occupancy(), o.occupancy(), f"scan {k}")
    g.close()
