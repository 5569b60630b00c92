"""GPU parity at the stress shape: 1401x1401x151 @ 0.05 m (296 M voxels),
262,144-point scans of 128 beams x 2048 azimuths, 60 m range (rays of up
to ~1,200 voxels), 2,000 boxes, a recenter on every scan (0.5 m threshold).

The fused update (one lane per ray at this size) and the recenter are
compared bit-exactly with the oracle (16 threads) after every scan, and a
batch of sphere fetches (r = 1 m) as sample multisets.
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
