"""GPU parity of the tile-binned fused update (the grid's default path).

The fused kernels never materialise per-voxel counters in normal operation;
their count-export mode writes the counts they fold, so the traversed voxel
sets can be compared bit-exactly with the oracle's ray cast. The per-voxel
atomics path (reference shape) is checked end to end as well.
"""
import numpy as np
import pytest

from helpers import assert_counters_equal, assert_occ_bitexact, make_pair

pytestmark = pytest.mark.gpu


def _upload(cb, scan):
    return [cb.DeviceBuffer.from_array(np.ascontiguousarray(c))
            for c in (scan.xs, scan.ys, scan.zs)]


@pytest.mark.parametrize("near_steps", [1, 4, 8, 12])
def test_fused_counts_small(cb, orc, gpu_ready, near_steps):
    w = cb.WORKLOADS["small"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w)
    g.set_near_steps(near_steps)
    for k in range(w.scans):
        s = scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
        bx, by, bz = _upload(cb, s)
        g.cast_rays_fused_sensor_frame_device(bx.ptr, by.ptr, bz.ptr, s.n, s.T)
        o.raycast(s.world(), s.sensor_origin)
        g_op, g_hit = g.counters()
        o_op, o_hit = o.counters()
        assert_counters_equal(g_op, g_hit, o_op, o_hit, f"K={near_steps} scan {k}")
        g.apply_counters()
        o.apply()
        g.recenter(s.base)
        o.recenter(s.base)
        assert_occ_bitexact(g.occupancy(), o.occupancy(), f"K={near_steps} scan {k}")


def test_fused_counts_readme(cb, orc, gpu_ready):
    w = cb.WORKLOADS["readme"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w, threads=16)
    for k in range(3):
        s = scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
        print(f"[readme fused] scan {k}", flush=True)
        bx, by, bz = _upload(cb, s)
        g.cast_rays_fused_sensor_frame_device(bx.ptr, by.ptr, bz.ptr, s.n, s.T)
        o.raycast(s.world(), s.sensor_origin)
        g_op, g_hit = g.counters()
        o_op, o_hit = o.counters()
        assert_counters_equal(g_op, g_hit, o_op, o_hit, f"readme scan {k}")
        g.apply_counters()
        o.apply()
        g.recenter(s.base)
        o.recenter(s.base)


@pytest.mark.parametrize("path", [0, 1])
def test_update_paths_end_to_end_readme(cb, orc, gpu_ready, path):
    """Both update paths: occupancy bit-exact after each update+recenter."""
    w = cb.WORKLOADS["readme"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w, threads=16)
    g.set_update_path(path)
    for k in range(6):
        s = scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
        g.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        o.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        g.recenter(s.base)
        o.recenter(s.base)
        assert_occ_bitexact(g.occupancy(), o.occupancy(), f"path {path} scan {k}")


def test_fused_centre_binning_and_long_range(cb, orc, gpu_ready):
    """Centre binning, a range longer than the window (rays leave the
    window), on a small grid with partial tiles on every axis."""
    w = cb.WORKLOADS["small"]
    scene = w.scene()
    g = cb.OccupancyGrid((37, 21, 9), 0.1, True, 1.0)
    o = orc.OracleGrid((37, 21, 9), 0.1, True, 1.0)
    for x in (g, o):
        x.set_max_raycast_range(25.0)
        x.set_log_odds_params(*cb.log_odds())
    for k in range(8):
        s = scene.scan(k, 16, 512, -15.0, 15.0, 25.0)
        g.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        o.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        g.recenter(s.base)
        o.recenter(s.base)
        assert_occ_bitexact(g.occupancy(), o.occupancy(), f"centre scan {k}")


def test_fused_sensor_outside_window(cb, orc, gpu_ready):
    """No recentering while the robot drives away: after the first scans the
    sensor voxel lies outside the window and rays enter it from outside."""
    from cublox_amd.scene import Scene
    scene = Scene(seed=77, n_boxes=40, box_extent=10.0, speed=1.0)
    g = cb.OccupancyGrid((15, 10, 6), 0.1, False, 0.5)
    o = orc.OracleGrid((15, 10, 6), 0.1, False, 0.5)
    for x in (g, o):
        x.set_max_raycast_range(8.0)
        x.set_log_odds_params(*cb.log_odds())
    outside = 0
    for k in range(8):
        s = scene.scan(k, 16, 1024, -15.0, 15.0, 8.0)
        bx, by, bz = _upload(cb, s)
        g.cast_rays_fused_sensor_frame_device(bx.ptr, by.ptr, bz.ptr, s.n, s.T)
        o.raycast(s.world(), s.sensor_origin)
        g_op, g_hit = g.counters()
        o_op, o_hit = o.counters()
        assert_counters_equal(g_op, g_hit, o_op, o_hit, f"outside scan {k}")
        g.apply_counters()
        o.apply()
        assert_occ_bitexact(g.occupancy(), o.occupancy(), f"outside scan {k}")
        oi = np.asarray(o.origin()[0])
        sv = np.floor(np.asarray(s.sensor_origin, np.float64) / 0.1).astype(int)
        if np.any(np.abs(sv - oi) > [15, 10, 6]):
            outside += 1
    assert outside >= 3
