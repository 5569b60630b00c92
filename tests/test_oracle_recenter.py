"""Recenter plan (reference kernels.cu:313-398) and slab clearing.

Pins: the "exiting" slices equal the local indices of global cells that
leave the window (brute-force window membership); when the new origin is
congruent to 0 modulo the map size, the whole cleared set equals the local
indices of cells in exactly one of the old and new windows (the case where
the reference's origin-independent "entering" slices are exact); the
full-reset rule uses the largest map extent over all axes. The product's
host plan (cublox_plan_recenter) must return the same lists, and its
zeroed-voxel count must equal a brute-force count over the grid.
"""
import ctypes as C

import numpy as np
import pytest


def _fold(g, m, h):
    return ((g + h) % m) - h


def _exits(o, n, h, m):
    """Local indices of cells in the old window [o-h, o+h] not in the new."""
    old = set(range(o - h, o + h + 1))
    new = set(range(n - h, n + h + 1))
    return {_fold(g, m, h) for g in old - new}


def _sym(o, n, h, m):
    old = set(range(o - h, o + h + 1))
    new = set(range(n - h, n + h + 1))
    return {_fold(g, m, h) for g in old ^ new}


def _plan(orc, pos, origin_i, half, res=0.1, thr=0.5, center=False):
    m = tuple(2 * h + 1 for h in half)
    inv = np.float32(1.0) / np.float32(res)
    of = [orc.global_to_pos(origin_i[a], np.float32(res), center) for a in range(3)]
    return orc.plan_recenter(pos, of, origin_i, m, half, inv,
                             np.float32(thr) * np.float32(thr), center)


def test_exit_slices_match_window_membership(orc):
    rng = np.random.default_rng(3)
    half = (9, 7, 5)
    m = tuple(2 * h + 1 for h in half)
    for _ in range(400):
        o = tuple(int(v) for v in rng.integers(-40, 40, 3))
        shift = tuple(int(v) for v in rng.integers(-8, 9, 3))
        n = tuple(o[a] + shift[a] for a in range(3))
        pos = [np.float32((n[a] + 0.5) * 0.1) for a in range(3)]
        action, new_o, slices = _plan(orc, pos, o, half)
        if all(s == 0 for s in shift):
            continue
        if action == 0:
            continue
        assert action == 1 and tuple(new_o) == n
        for a in range(3):
            got = set(slices[a].tolist())
            if shift[a] == 0:
                assert not got
                continue
            ex = _exits(o[a], n[a], half[a], m[a])
            assert ex <= got
            assert len(slices[a]) == 2 * abs(shift[a])
            assert all(-half[a] <= v <= half[a] for v in got)


def test_congruent_origin_clears_exactly_changed_cells(orc):
    half = (9, 7, 5)
    m = tuple(2 * h + 1 for h in half)
    rng = np.random.default_rng(8)
    for _ in range(200):
        n = tuple(int(rng.integers(-3, 4)) * m[a] for a in range(3))
        shift = tuple(int(v) for v in rng.integers(-4, 5, 3))
        o = tuple(n[a] - shift[a] for a in range(3))
        pos = [np.float32((n[a] + 0.5) * 0.1) for a in range(3)]
        action, new_o, slices = _plan(orc, pos, o, half)
        if action == 0:
            continue
        for a in range(3):
            got = set(slices[a].tolist())
            assert got == (_sym(o[a], n[a], half[a], m[a]) if shift[a] else set())


def test_full_reset_uses_largest_extent(orc):
    half = (20, 20, 3)  # map 41 x 41 x 7
    o = (0, 0, 0)
    # z shift of 8 >= z map (7) but < largest extent (41): still slabs.
    action, _, slices = _plan(orc, [0.05, 0.05, np.float32(0.85)], o, half)
    assert action == 1
    assert len(slices[2]) == 16  # 2*|s| entries, more than map_z: the union
    assert set(slices[2].tolist()) == set(range(-3, 4))
    # x shift of 41 = largest extent: full reset.
    action, _, _ = _plan(orc, [np.float32(4.15), 0.05, 0.05], o, half)
    assert action == 2


def test_threshold_gate(orc):
    half = (5, 5, 5)
    a, _, _ = _plan(orc, [0.1, 0.0, 0.0], (0, 0, 0), half, thr=0.5)
    assert a == 0
    a, _, _ = _plan(orc, [0.9, 0.0, 0.0], (0, 0, 0), half, thr=0.5)
    assert a == 1


def test_product_plan_matches_oracle(cb, orc):
    from cublox_amd import _ffi
    L = _ffi.hip()
    rng = np.random.default_rng(21)
    for _ in range(300):
        half = tuple(int(v) for v in rng.integers(1, 12, 3))
        m = tuple(2 * h + 1 for h in half)
        res = float(rng.choice([0.05, 0.1, 0.25]))
        ctr = bool(rng.integers(0, 2))
        o = tuple(int(v) for v in rng.integers(-50, 50, 3))
        pos = (np.array(o) + rng.uniform(-30, 30, 3)) * res
        pos = pos.astype(np.float32)
        inv = np.float32(1.0) / np.float32(res)
        of = [orc.global_to_pos(o[a], np.float32(res), ctr) for a in range(3)]
        thr_sq = np.float32(0.3) * np.float32(0.3)
        a_o, no_o, sl_o = orc.plan_recenter(pos, of, o, m, half, inv, thr_sq, ctr)
        cfg = _ffi.RecenterCfg(
            _ffi.Float3(*pos), _ffi.Float3(*of), _ffi.Int3(*o), _ffi.Int3(*m),
            _ffi.Int3(*half), np.float32(res), inv, thr_sq, int(ctr),
            m[0] * m[1] * m[2])
        bufs = [(C.c_int * (2 * max(m) + 2))() for a in range(3)]
        r = _ffi.RecenterResult()
        assert L.cublox_plan_recenter(C.byref(cfg), bufs[0], bufs[1], bufs[2],
                                      C.byref(r)) == 0
        assert r.action == a_o
        if a_o == 0:
            assert r.zeroed_voxels == 0
            continue
        assert (r.new_origin_i.x, r.new_origin_i.y, r.new_origin_i.z) == tuple(no_o)
        V = m[0] * m[1] * m[2]
        if a_o == 2:
            assert r.zeroed_voxels == V
            continue
        for a in range(3):
            assert list(bufs[a][:r.n_slices[a]]) == sl_o[a].tolist()
        # Brute-force zeroed-voxel count.
        mask = np.zeros(m, bool)
        for a in range(3):
            for l in set(sl_o[a].tolist()):
                idx = [slice(None)] * 3
                idx[a] = l + half[a]
                mask[tuple(idx)] = True
        assert r.zeroed_voxels == int(mask.sum())


def test_oracle_grid_recenter_clears_planned_slabs(orc):
    """A full grid filled with ones: after recenter the zero pattern is the
    union of the planned slabs (x-slab = hash range, etc.)."""
    half = (6, 5, 4)
    m = tuple(2 * h + 1 for h in half)
    g = orc.OracleGrid(half, 0.1, False, 0.2)
    V = g.voxel_num
    g.load_occupancy(np.ones(V, np.float32))
    oi, of = g.origin()
    pos = np.array([0.75, -0.35, 0.25], np.float32)
    act, new_o, sl = orc.plan_recenter(pos, of, oi, m, half, np.float32(10.0),
                                       np.float32(0.04), False)
    assert act == 1
    assert g.recenter(pos) == 1
    occ = g.occupancy().reshape(m)
    mask = np.zeros(m, bool)
    for a in range(3):
        for l in set(sl[a].tolist()):
            idx = [slice(None)] * 3
            idx[a] = l + half[a]
            mask[tuple(idx)] = True
    assert np.array_equal(occ == 0, mask)
    assert g.origin()[0] == tuple(new_o)
