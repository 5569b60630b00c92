"""GPU checks of the point queries and of a recenter between the stage-wise
ray cast and apply.

Point queries (reference OccupancyGrid.hpp:90-115, OccupancyGrid.cpp:
358-429): isOccupied / isFree / isUnknown by position, global index and hash
id, on a grid that scans have filled and recenters have moved, against
labels computed here from the oracle grid's log-odds with the reference's
rules written out independently of the product:
  * occupied: l >= l_occ; free: l < l_free; unknown: neither;
  * a position or global index outside the window: unknown true, the
    others false; a hash id outside [0, V): the stored value reads as 0.
The threshold edges are planted exactly (l_occ, l_free and their float
neighbours) through loadOccupancyFromHost.

Stage-wise recenter: castRays leaves the per-voxel counters pending; a
recenter before applyCounters must clear the counts in its slabs together
with occ, as the reference's launchRecenter clears occ, op and hit
(kernels.cu:404-499, 652-663).
"""
import numpy as np
import pytest

from helpers import assert_counters_equal, assert_occ_bitexact, make_pair

pytestmark = pytest.mark.gpu


def _labels(l, l_free, l_occ):
    """(occupied, free, unknown) of stored log-odds l (l_unknown = 0)."""
    occ = l >= l_occ
    free = l < l_free
    return occ, free, (l >= l_free) & (l < l_occ)


def _filled_pair(cb, orc, scans=6):
    w = cb.WORKLOADS["small"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w)
    moves = 0
    for k in range(scans):
        s = scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
        g.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        o.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        moves += int(g.recenter(s.base))
        o.recenter(s.base)
    assert moves >= 1, "the queries must run on a recentered window"
    assert g.geometry.origin_i == o.origin()[0]
    return w, g, o


def _plant_edges(g, o, l_free, l_occ, rng):
    """Exact threshold values and their float neighbours at random voxels
    (the same values in both grids). Returns the planted hash ids."""
    occ = o.occupancy()
    vals = np.array([l_occ, np.nextafter(l_occ, np.float32(-np.inf)),
                     np.nextafter(l_occ, np.float32(np.inf)), l_free,
                     np.nextafter(l_free, np.float32(-np.inf)),
                     np.nextafter(l_free, np.float32(np.inf)), 0.0, -0.0],
                    np.float32)
    ids = rng.choice(occ.size, size=vals.size * 16, replace=False)
    occ[ids] = np.tile(vals, 16)
    g.load_occupancy(occ)
    o.load_occupancy(occ)
    assert_occ_bitexact(g.occupancy(), o.occupancy(), "planted")
    return ids


def test_point_queries_all_overloads(cb, orc, gpu_ready):
    w, g, o = _filled_pair(cb, orc)
    lo = cb.log_odds()
    l_free, l_occ = np.float32(lo[4]), np.float32(lo[5])
    rng = np.random.default_rng(5)
    planted = _plant_edges(g, o, l_free, l_occ, rng)
    occ = o.occupancy()
    geo = g.geometry
    m = np.array(geo.map_size)
    h = np.array(geo.half)
    oi = np.array(o.origin()[0])
    V = int(m.prod())
    res, inv = np.float32(w.resolution), np.float32(1.0) / np.float32(w.resolution)

    def expect_hash(hid):
        l = occ[hid] if 0 <= hid < V else np.float32(0.0)
        return _labels(l, l_free, l_occ)

    def expect_index(gi):
        gi = np.asarray(gi)
        if np.any(np.abs(gi - oi) > h):
            return False, False, True
        return expect_hash(orc.global_to_hash(gi, m, h))

    def check(label, got, exp):
        assert tuple(bool(x) for x in got) == tuple(bool(x) for x in exp), (label, got, exp)

    # Hash ids: the planted edges, random valid ids, and ids outside [0, V).
    hids = list(planted) + list(rng.integers(0, V, 200)) + [-1, V, V + 7, -(2 ** 31), 2 ** 31 - 1]
    n_occ = n_free = 0
    for hid in hids:
        hid = int(hid)
        got = (g.is_occupied(hash_id=hid), g.is_free(hash_id=hid), g.is_unknown(hash_id=hid))
        exp = expect_hash(hid)
        check(f"hash {hid}", got, exp)
        n_occ += int(exp[0])
        n_free += int(exp[1])
    assert n_occ > 0 and n_free > 0, "labels of every kind must be exercised"

    # Global indices: every planted voxel's index, random indices inside the
    # window and indices just past each face of it.
    idx = []
    for hid in planted[:40]:
        loc = np.array(orc.hash_to_local(int(hid), m, h))
        idx.append([orc.local_to_global(int(loc[a]), int(m[a]), int(h[a]), int(oi[a]))
                    for a in range(3)])
    for _ in range(150):
        idx.append((oi + rng.integers(-h, h + 1)).tolist())
    for a in range(3):
        for sgn in (-1, 1):
            e = oi.copy()
            e[a] += sgn * h[a]
            idx.append(e.tolist())      # on the face: inside
            e2 = e.copy()
            e2[a] += sgn
            idx.append(e2.tolist())     # one past: outside
    for gi in idx:
        got = (g.is_occupied(index=gi), g.is_free(index=gi), g.is_unknown(index=gi))
        check(f"index {gi}", got, expect_index(gi))

    # Positions: voxel centres of the indices above (mapped back with the
    # oracle's index math) and random points in and around the window.
    pts = []
    for gi in idx[:120]:
        pts.append([orc.global_to_pos(int(gi[a]), float(res)) + 0.5 * float(res)
                    for a in range(3)])
    span = (h + 8) * float(res)
    of = np.array(o.origin()[1], np.float64)
    for _ in range(200):
        pts.append((of + rng.uniform(-span, span)).tolist())
    for p in pts:
        p32 = np.asarray(p, np.float32)
        gi = [orc.pos_to_global(float(p32[a]), float(inv)) for a in range(3)]
        got = (g.is_occupied(p32), g.is_free(p32), g.is_unknown(p32))
        check(f"pos {p32.tolist()}", got, expect_index(gi))
    g.close()


def test_recenter_between_cast_and_apply(cb, orc, gpu_ready):
    """castRays -> recenter -> applyCounters on both sides: the counters
    pending in the cleared slabs are cleared with occ."""
    w = cb.WORKLOADS["small"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w)
    moves = 0
    for k in range(12):  # recenters fire at scans 5 and 10
        s = scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
        pts = s.world()
        bufs = [cb.DeviceBuffer.from_array(np.ascontiguousarray(pts[:, a])) for a in range(3)]
        print(f"[stagewise] scan {k} n={s.n}", flush=True)
        g.cast_rays_device(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, s.n, s.sensor_origin)
        o.raycast(pts, s.sensor_origin)
        moved = g.recenter(s.base)
        act = o.recenter(s.base)
        assert moved == (act != 0)
        moves += int(moved)
        g_op, g_hit = g.counters()
        o_op, o_hit = o.counters()
        assert_counters_equal(g_op, g_hit, o_op, o_hit, f"after recenter, scan {k}")
        assert_occ_bitexact(g.occupancy(), o.occupancy(), f"after recenter, scan {k}")
        g.apply_counters()
        o.apply()
        assert_occ_bitexact(g.occupancy(), o.occupancy(), f"apply, scan {k}")
    assert moves >= 2
    g.close()
