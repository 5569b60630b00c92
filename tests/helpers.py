"""Helpers shared by the parity tests: paired product/oracle grids and the
comparison rules (bit-exact counters and recentered grids, log-odds within
1e-4, labels equal away from the thresholds, fetch as multisets)."""
from __future__ import annotations

import numpy as np

LOGODDS_TOL = 1e-4  # absolute; float atomic order is not involved here


def make_pair(cb, orc, w, *, center=None, threshold="cfg", threads=0,
              log_odds=None, max_range=None):
    center = w.origin_at_center if center is None else center
    thr = w.threshold if threshold == "cfg" else threshold
    g = cb.OccupancyGrid(w.half, w.resolution, center, thr)
    o = orc.OracleGrid(w.half, w.resolution, center, thr, threads=threads)
    lo = log_odds or cb.log_odds()
    mr = w.max_range if max_range is None else max_range
    for x in (g, o):
        x.set_max_raycast_range(mr)
        x.set_log_odds_params(*lo)
    return g, o


def assert_counters_equal(g_op, g_hit, o_op, o_hit, label=""):
    if not np.array_equal(g_op, o_op):
        d = np.nonzero(g_op != o_op)[0]
        raise AssertionError(
            f"{label}: op counters differ at {d.size} voxels, e.g. h={d[:5]} "
            f"gpu={g_op[d[:5]]} oracle={o_op[d[:5]]}")
    if not np.array_equal(g_hit, o_hit):
        d = np.nonzero(g_hit != o_hit)[0]
        raise AssertionError(
            f"{label}: hit counters differ at {d.size} voxels, e.g. h={d[:5]} "
            f"gpu={g_hit[d[:5]]} oracle={o_hit[d[:5]]}")


def assert_occ_close(g_occ, o_occ, l_free, l_occ, label=""):
    diff = np.abs(g_occ.astype(np.float64) - o_occ.astype(np.float64))
    worst = float(diff.max()) if diff.size else 0.0
    assert worst <= LOGODDS_TOL, f"{label}: max |dlog-odds| {worst} > {LOGODDS_TOL}"
    # Labels: identical except within the tolerance of a threshold.
    def labels(a):
        return np.where(a >= l_occ, 2, np.where(a < l_free, 1, 0))
    near = (np.abs(o_occ - l_occ) <= LOGODDS_TOL) | (np.abs(o_occ - l_free) <= LOGODDS_TOL)
    bad = (labels(g_occ) != labels(o_occ)) & ~near
    assert not bad.any(), f"{label}: {int(bad.sum())} labels differ"


def assert_occ_bitexact(g_occ, o_occ, label=""):
    gb = g_occ.view(np.uint32)
    ob = o_occ.view(np.uint32)
    if not np.array_equal(gb, ob):
        d = np.nonzero(gb != ob)[0]
        raise AssertionError(
            f"{label}: occupancy differs bitwise at {d.size} voxels, e.g. "
            f"h={d[:5]} gpu={g_occ[d[:5]]} oracle={o_occ[d[:5]]}")


def sample_multiset(pos, st):
    """Rows of (x bits, y bits, z bits, state), sorted: order-free compare."""
    if len(st) == 0:
        return np.zeros((0, 4), np.uint64)
    bits = np.ascontiguousarray(pos, np.float32).view(np.uint32).astype(np.uint64)
    rows = np.concatenate([bits.reshape(-1, 3), st.astype(np.uint64)[:, None]], 1)
    order = np.lexsort(rows.T[::-1])
    return rows[order]
