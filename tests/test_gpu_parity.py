"""GPU parity of the drop-in OccupancyGrid against the CPU oracle.

Every case runs the HIP kernels through the C ABI (libcublox_hip.so) via
the C++ OccupancyGrid (libcublox.so) and compares with the oracle on the
same seeded inputs:
  * op/hit counters after the ray cast: identical (traversed voxel sets);
  * log-odds after apply: |d| <= 1e-4 (expected bit-exact, also asserted
    bit-exact where the arithmetic is pure IEEE FP32);
  * occupancy after recenter: bit-exact, and the same window origin;
  * sphere fetch: same count and the same multiset of (pos bits, state).
"""
import numpy as np
import pytest

from helpers import (assert_counters_equal, assert_occ_bitexact,
                     assert_occ_close, make_pair, sample_multiset)

pytestmark = pytest.mark.gpu


def _upload_world(cb, scan):
    pts = scan.world()
    bufs = [cb.DeviceBuffer.from_array(np.ascontiguousarray(pts[:, k]))
            for k in range(3)]
    return bufs


def test_small_full_run_stagewise(cb, orc, gpu_ready):
    """20 scans of the small workload: counters, apply, recenter, fetch."""
    w = cb.WORKLOADS["small"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w)
    l_free, l_occ = cb.log_odds()[4], cb.log_odds()[5]
    moves = 0
    for k in range(w.scans):
        s = scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
        print(f"[small] scan {k} n={s.n} miss={s.n_miss}", flush=True)
        bx, by, bz = _upload_world(cb, s)
        g.cast_rays_device(bx.ptr, by.ptr, bz.ptr, s.n, s.sensor_origin)
        o.raycast(s.world(), s.sensor_origin)
        g_op, g_hit = g.counters()
        o_op, o_hit = o.counters()
        assert_counters_equal(g_op, g_hit, o_op, o_hit, f"scan {k}")
        g.apply_counters()
        o.apply()
        go, oo = g.occupancy(), o.occupancy()
        assert_occ_close(go, oo, l_free, l_occ, f"apply scan {k}")
        assert_occ_bitexact(go, oo, f"apply scan {k}")
        moved = g.recenter(s.base)
        act = o.recenter(s.base)
        assert moved == (act != 0)
        moves += int(moved)
        assert g.geometry.origin_i == o.origin()[0]
        assert_occ_bitexact(g.occupancy(), o.occupancy(), f"recenter scan {k}")
        # Counters are zero between updates.
        g_op, g_hit = g.counters()
        assert not g_op.any() and not g_hit.any()
    assert moves >= 3, "the trajectory must exercise recentering"
    # Fetch around a few centres, all states and occupied only.
    qs = scene.queries(w.scans - 1, 8, 6.0)
    for qi, c in enumerate(qs):
        for filt in (None, cb.OCCUPIED, cb.FREE, cb.UNKNOWN):
            gp, gs = g.fetch_occupancy_around(c, 1.5, filt)
            op_, os_, on = o.fetch(c, 1.5, filt)
            assert len(gs) == on, f"query {qi} filter {filt}: count"
            assert np.array_equal(sample_multiset(gp, gs), sample_multiset(op_, os_))


def test_small_fused_transform_end_to_end(cb, orc, gpu_ready):
    """Sensor-frame input through the fused transform equals the oracle's
    host transform + update, scan by scan, including recenters."""
    w = cb.WORKLOADS["small"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w)
    for k in range(w.scans):
        s = scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
        g.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        o.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        g.recenter(s.base)
        o.recenter(s.base)
        assert_occ_bitexact(g.occupancy(), o.occupancy(), f"fused scan {k}")


def test_world_frame_update_matches(cb, orc, gpu_ready):
    """The reference-shaped update(cloud, origin) with host world points."""
    w = cb.WORKLOADS["small"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w)
    for k in range(5):
        s = scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
        pts = s.world()
        g.update(pts, s.sensor_origin)
        o.update(pts, s.sensor_origin)
        assert_occ_bitexact(g.occupancy(), o.occupancy(), f"world scan {k}")


def test_readme_shape_first_scans(cb, orc, gpu_ready):
    """The 701x701x151 grid: first 4 scans stage-wise, 16 threads of oracle."""
    w = cb.WORKLOADS["readme"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w, threads=16)
    for k in range(4):
        s = scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
        print(f"[readme] scan {k}", flush=True)
        bx, by, bz = _upload_world(cb, s)
        g.cast_rays_device(bx.ptr, by.ptr, bz.ptr, s.n, s.sensor_origin)
        o.raycast(s.world(), s.sensor_origin)
        g_op, g_hit = g.counters()
        o_op, o_hit = o.counters()
        assert_counters_equal(g_op, g_hit, o_op, o_hit, f"readme scan {k}")
        g.apply_counters()
        o.apply()
        assert_occ_bitexact(g.occupancy(), o.occupancy(), f"readme apply {k}")
        g.recenter(s.base)
        o.recenter(s.base)
        assert_occ_bitexact(g.occupancy(), o.occupancy(), f"readme recenter {k}")
    qs = scene.queries(3, 16, 20.0)
    for c in qs:
        gp, gs = g.fetch_occupancy_around(c, w.fetch_radius, cb.OCCUPIED)
        op_, os_, on = o.fetch(c, w.fetch_radius, cb.OCCUPIED)
        assert len(gs) == on
        assert np.array_equal(sample_multiset(gp, gs), sample_multiset(op_, os_))


def test_batched_fetch_beyond_one_launch(cb, orc, gpu_ready):
    """70,000 queries: the batch runs in launches of 65,535 queries; every
    query's count equals the oracle's, and queries of the second launch
    give the oracle's samples."""
    from helpers import sample_multiset
    w = cb.WORKLOADS["small"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w)
    for k in range(3):
        s = scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
        g.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        o.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
    rng = np.random.default_rng(8)
    base = np.asarray(s.base, np.float32)
    q = 70_000
    cs = (base + rng.uniform([-8, -8, -1], [8, 8, 2], (q, 3))).astype(np.float32)
    cap = 512
    counts, pos, st = g.fetch_occupancy_batch(cs, 0.35, None, cap)
    offs = np.zeros(q + 1, np.int64)
    offs[1:] = np.cumsum(np.minimum(counts, cap).astype(np.int64))
    want = np.array([o.fetch(c, 0.35, None, cap, cap=cap)[2] for c in cs])
    assert np.array_equal(counts, np.minimum(want, cap)), "per-query counts"
    for j in list(range(65_530, 65_545)) + list(rng.integers(65_535, q, 40)):
        op_, os_, on = o.fetch(cs[j], 0.35, None, cap, cap=cap)
        a, b = offs[j], offs[j + 1]
        assert np.array_equal(sample_multiset(pos[a:b], st[a:b]),
                              sample_multiset(op_, os_)), j


def test_batched_fetch_dense_queries(cb, orc, gpu_ready):
    """Queries that keep hundreds of samples per wave (no state filter,
    1.5 m spheres, some reaching past the window): each wave's staged
    samples flush several times; counts and samples equal the oracle's.
    With a cap below the counts, counts stay exact and every returned
    sample is one of the oracle's, without repeats."""
    from helpers import sample_multiset
    w = cb.WORKLOADS["small"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w)
    for k in range(3):
        s = scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
        g.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        o.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
    rng = np.random.default_rng(21)
    base = np.asarray(s.base, np.float32)
    cs = (base + rng.uniform([-11, -11, -3], [11, 11, 3], (24, 3))).astype(np.float32)
    radius = 1.5
    full = 40_000
    counts, pos, st = g.fetch_occupancy_batch(cs, radius, None, full)
    offs = np.zeros(len(cs) + 1, np.int64)
    offs[1:] = np.cumsum(counts.astype(np.int64))
    assert counts.max() > 5_000, "queries should keep thousands of samples"
    for j, c in enumerate(cs):
        op_, os_, on = o.fetch(c, radius, None, full, cap=full)
        assert counts[j] == on, j
        a, b = offs[j], offs[j + 1]
        assert np.array_equal(sample_multiset(pos[a:b], st[a:b]),
                              sample_multiset(op_, os_)), j
    cap = 1000
    counts_c, pos_c, st_c = g.fetch_occupancy_batch(cs, radius, None, cap)
    assert np.array_equal(counts_c, np.minimum(counts, cap))
    offs_c = np.zeros(len(cs) + 1, np.int64)
    offs_c[1:] = np.cumsum(counts_c.astype(np.int64))
    for j, c in enumerate(cs):
        op_, os_, _ = o.fetch(c, radius, None, full, cap=full)
        every = set(map(tuple, sample_multiset(op_, os_).tolist()))
        got = sample_multiset(pos_c[offs_c[j]:offs_c[j + 1]], st_c[offs_c[j]:offs_c[j + 1]])
        rows = list(map(tuple, got.tolist()))
        assert len(set(rows)) == len(rows) and set(rows) <= every, j
