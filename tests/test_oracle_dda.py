"""The oracle's ray cast (reference kernels.cu:39-151) against geometry.

Pins: on tiny grids, the voxels the walk increments are exactly the voxels
a double-precision segment passes through (voxels it crosses by more than
1e-4 m must be visited; visited voxels must be within 1e-4 m of the
segment); the walk is 6-connected and moves one unit of L1 distance per
step; axis-aligned rays are straight runs; non-finite and zero-length rays
are skipped; range-clipped rays count their endpoint without a hit.
"""
import numpy as np
import pytest

HALF = (6, 6, 6)
MAP = tuple(2 * h + 1 for h in HALF)
RES = np.float32(0.25)
INV = np.float32(1.0) / RES


def _local_to_global(o_i, hid):
    """hash -> global index for the window around o_i (independent of the
    oracle: congruence search)."""
    my, mz = MAP[1], MAP[2]
    c = (hid // (my * mz), (hid // mz) % my, hid % mz)
    out = []
    for a in range(3):
        l = c[a] - HALF[a]
        window = range(o_i[a] - HALF[a], o_i[a] + HALF[a] + 1)
        out.append([g for g in window if (g - l) % MAP[a] == 0][0])
    return tuple(out)


def _seg_len_in_box(o, e, lo, hi):
    """Length of segment o->e inside the box [lo, hi] (slab clipping)."""
    d = e - o
    t0, t1 = 0.0, 1.0
    for a in range(3):
        if abs(d[a]) < 1e-15:
            if o[a] < lo[a] or o[a] > hi[a]:
                return -1.0
            continue
        ta = (lo[a] - o[a]) / d[a]
        tb = (hi[a] - o[a]) / d[a]
        if ta > tb:
            ta, tb = tb, ta
        t0, t1 = max(t0, ta), min(t1, tb)
        if t0 > t1:
            return -1.0
    return (t1 - t0) * float(np.linalg.norm(d))


def _touches(o, e, lo, hi, eps=1e-4):
    """Whether the segment meets the box inflated by eps on every side."""
    return _seg_len_in_box(o, e, lo - eps, hi + eps) >= 0.0


def _trace(orc, o, e, o_i=(0, 0, 0), max_range=100.0, center=False):
    return orc.trace_ray(o, e, o_i, MAP, HALF, RES, INV, max_range, center)


@pytest.mark.parametrize("center", [False, True])
def test_walk_matches_segment_voxels(orc, center):
    rng = np.random.default_rng(123 + int(center))
    checked = 0
    for _ in range(300):
        o = rng.uniform(-1.4, 1.4, 3).astype(np.float32)
        e = rng.uniform(-1.4, 1.4, 3).astype(np.float32)
        res = _trace(orc, o, e, center=center)
        assert res is not None
        ids, clipped, ep = res
        assert not clipped
        visited = [_local_to_global((0, 0, 0), h) for h in ids]
        # Geometry in voxel units (double).
        od, ed = o.astype(np.float64), e.astype(np.float64)
        cells = set(visited)
        lo_off = -0.5 if center else 0.0
        # Every voxel crossed by more than 1e-4 m is visited.
        gmin = np.floor(np.minimum(od, ed) / float(RES)).astype(int) - 1
        gmax = np.floor(np.maximum(od, ed) / float(RES)).astype(int) + 1
        for gx in range(gmin[0], gmax[0] + 1):
            for gy in range(gmin[1], gmax[1] + 1):
                for gz in range(gmin[2], gmax[2] + 1):
                    lo = (np.array([gx, gy, gz]) + lo_off) * float(RES)
                    L = _seg_len_in_box(od, ed, lo, lo + float(RES))
                    if L > 1e-4:
                        assert (gx, gy, gz) in cells, (o, e, (gx, gy, gz), L)
        # Every visited voxel touches the segment.
        for g in cells:
            lo = (np.array(g) + lo_off) * float(RES)
            assert _touches(od, ed, lo, lo + float(RES)), (o, e, g)
        checked += 1
    assert checked == 300


def test_walk_is_6_connected_and_l1_monotone(orc):
    rng = np.random.default_rng(5)
    for _ in range(300):
        o = rng.uniform(-1.4, 1.4, 3).astype(np.float32)
        e = rng.uniform(-1.4, 1.4, 3).astype(np.float32)
        ids, clipped, ep = _trace(orc, o, e)
        vox = [_local_to_global((0, 0, 0), h) for h in ids]
        start = tuple(int(np.floor(o[a] * INV)) for a in range(3))
        end = tuple(int(np.floor(np.float32(e[a] * INV))) for a in range(3))
        assert ep and vox[-1] == end  # endpoint counted last
        walk = vox[:-1]
        if walk:
            assert walk[0] == start
        for k, v in enumerate(walk):
            assert sum(abs(v[a] - start[a]) for a in range(3)) == k
        for a_, b_ in zip(walk, walk[1:]):
            assert sum(abs(a_[i] - b_[i]) for i in range(3)) == 1
        if walk:
            # the walk stops right before the endpoint voxel
            assert sum(abs(walk[-1][i] - end[i]) for i in range(3)) == 1


def test_axis_aligned_runs(orc):
    for axis in range(3):
        for sign in (-1, 1):
            o = np.array([0.1, 0.1, 0.1], np.float32)
            e = o.copy()
            e[axis] += sign * 1.2
            ids, clipped, ep = _trace(orc, o, e)
            vox = [_local_to_global((0, 0, 0), h) for h in ids]
            n_exp = abs(int(np.floor(e[axis] * INV)) - int(np.floor(o[axis] * INV))) + 1
            assert len(vox) == n_exp
            for k, v in enumerate(vox):
                want = [0, 0, 0]
                want[axis] = sign * k
                assert v == tuple(want)


def test_skip_cases(orc):
    o = np.zeros(3, np.float32)
    assert _trace(orc, o, np.array([np.nan, 0, 0], np.float32)) is None
    assert _trace(orc, o, np.array([np.inf, 0, 0], np.float32)) is None
    assert _trace(orc, o, o) is None
    assert _trace(orc, o, np.array([1e-8, 0, 0], np.float32)) is None


def test_range_clip(orc):
    o = np.array([0.05, 0.05, 0.05], np.float32)
    e = np.array([1.45, 0.05, 0.05], np.float32)
    ids, clipped, ep = _trace(orc, o, e, max_range=0.6)
    assert clipped and ep
    vox = [_local_to_global((0, 0, 0), h) for h in ids]
    # clip point at x = 0.65 -> voxel 2; walk visits 0, 1 then endpoint 2
    assert vox == [(0, 0, 0), (1, 0, 0), (2, 0, 0)]
    ids2, clipped2, _ = _trace(orc, o, e, max_range=5.0)
    assert not clipped2 and len(ids2) == 6


def test_out_of_window_voxels_not_counted(orc):
    o = np.array([0.01, 0.01, 0.01], np.float32)
    e = np.array([3.0, 0.01, 0.01], np.float32)  # beyond the 13-voxel window
    ids, clipped, ep = _trace(orc, o, e)
    vox = [_local_to_global((0, 0, 0), h) for h in ids]
    assert vox == [(k, 0, 0) for k in range(7)]  # x = 0..6 in window
    assert not ep  # endpoint x = 12 is outside the window
