"""GPU parity at the stress shape: 1401x1401x151 @ 0.05 m (296 M voxels),
262,144-point scans of 128 beams x 2048 azimuths, 60 m range (rays of up
to ~1,200 voxels), 2,000 boxes, a recenter on every scan (0.5 m threshold).

The fused update (one lane per ray at this size) and the recenter are
compared bit-exactly with the oracle (16 threads) after every scan, a set
of single sphere fetches (r = 1 m) as sample multisets, and the shape's
whole fetch workload (100,000 queries in one batched launch) query by
query.
"""
import numpy as np
import pytest

from helpers import assert_occ_bitexact, make_pair, sample_multiset

pytestmark = pytest.mark.gpu


def test_stress_shape_scans(cb, orc, gpu_ready):
    w = cb.WORKLOADS["stress"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w, threads=16)
    moves = 0
    for k in range(3):
        s = scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
        print(f"[stress] scan {k} n={s.n}", flush=True)
        bufs = [cb.DeviceBuffer.from_array(np.ascontiguousarray(c)) for c in (s.xs, s.ys, s.zs)]
        g.update_sensor_frame_device(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, s.n, s.T)
        o.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        assert_occ_bitexact(g.occupancy(), o.occupancy(), f"stress update {k}")
        moved = g.recenter(s.base)
        act = o.recenter(s.base)
        assert moved == (act != 0)
        moves += int(moved)
        assert g.geometry.origin_i == o.origin()[0]
        assert_occ_bitexact(g.occupancy(), o.occupancy(), f"stress recenter {k}")
    assert moves >= 2, "the 0.5 m threshold recenters on every scan"
    qs = scene.queries(2, 64, w.fetch_spread)
    for c in qs:
        gp, gs = g.fetch_occupancy_around(c, w.fetch_radius)
        op_, os_, on = o.fetch(c, w.fetch_radius)
        assert len(gs) == on
        assert np.array_equal(sample_multiset(gp, gs), sample_multiset(op_, os_))
    g.close()


def _mix(pos, st):
    """Order-independent per-sample hash terms (64-bit, wrapping)."""
    b = np.ascontiguousarray(pos, np.float32).view(np.uint32).astype(np.uint64)
    return (b[:, 0] * np.uint64(0x9E3779B97F4A7C15) ^ b[:, 1] * np.uint64(0xC2B2AE3D27D4EB4F)
            ^ b[:, 2] * np.uint64(0x165667B19E3779F9) ^ st.astype(np.uint64) * np.uint64(0x27D4EB2F165667C5))


def test_stress_batched_fetch_100k(cb, orc, gpu_ready):
    """The stress shape's fetch workload in full: 100,000 OCCUPIED queries
    (r = 1 m, 68,921 voxels each) in one batched launch after three scans,
    against the oracle query by query: every count, a checksum of each
    query's samples (order-independent), and the full sample multisets of
    500 queries."""
    from concurrent.futures import ThreadPoolExecutor
    w = cb.WORKLOADS["stress"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w, threads=16)
    for k in range(3):
        s = scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
        bufs = [cb.DeviceBuffer.from_array(np.ascontiguousarray(c)) for c in (s.xs, s.ys, s.zs)]
        g.update_sensor_frame_device(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, s.n, s.T)
        o.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        g.recenter(s.base)
        o.recenter(s.base)
    assert_occ_bitexact(g.occupancy(), o.occupancy(), "stress before the fetch batch")
    qs = np.asarray(scene.queries(3, w.fetch_queries, w.fetch_spread), np.float32)
    assert len(qs) == 100000
    r_vox = int(np.ceil(w.fetch_radius / w.resolution))
    cap = min((2 * r_vox + 1) ** 3, 16384)
    print(f"[stress] batched fetch of {len(qs)} queries", flush=True)
    counts, pos, st = g.fetch_occupancy_batch(qs, w.fetch_radius, cb.OCCUPIED, cap)
    assert int(counts.max()) < cap, "no query may hit the output cap"
    offs = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])
    mix = _mix(pos, st)
    # Sums over each query's samples (a zero appended, so that every start
    # index, the empty queries' at the end included, is in range).
    gsum = np.add.reduceat(np.append(mix, np.uint64(0)), offs[:-1])
    gsum[counts == 0] = 0

    def oracle_range(lo, hi):
        n = np.zeros(hi - lo, np.int64)
        h = np.zeros(hi - lo, np.uint64)
        for i in range(lo, hi):
            op_, os_, on = o.fetch(qs[i], w.fetch_radius, state_filter=cb.OCCUPIED, cap=cap)
            n[i - lo] = on
            h[i - lo] = _mix(op_, os_).sum(dtype=np.uint64) if on else 0
        return n, h

    step = len(qs) // 16
    with ThreadPoolExecutor(16) as ex:
        parts = list(ex.map(lambda t: oracle_range(t * step, len(qs) if t == 15 else (t + 1) * step),
                            range(16)))
    on_all = np.concatenate([p[0] for p in parts])
    oh_all = np.concatenate([p[1] for p in parts])
    assert np.array_equal(counts.astype(np.int64), on_all), "per-query sample counts"
    bad = np.nonzero(gsum != oh_all)[0]
    assert len(bad) == 0, "per-query sample checksums"
    rng = np.random.default_rng(7)
    for i in rng.choice(len(qs), 500, replace=False):
        op_, os_, on = o.fetch(qs[i], w.fetch_radius, state_filter=cb.OCCUPIED, cap=cap)
        a, b = offs[i], offs[i + 1]
        assert np.array_equal(sample_multiset(pos[a:b], st[a:b]), sample_multiset(op_, os_)), i
    g.close()
