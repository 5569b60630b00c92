"""GPU tests of map snapshots (no reference counterpart; format in
include/cublox/snapshot.h, read here by the tests' own numpy reader).

A snapshot of a grid must decode to the oracle grid's occupancy and origin
bit-exactly; restored into a fresh grid of the same geometry it must give
the same occupancy, and updates continued from it must stay bit-identical
to the oracle (so origin, first-run flag, threshold, log-odds and range
travel too). Covered: the sparse encoding (ascending hash ids) and the
dense one, -0.0 kept, file round trips, and rejected snapshots (other
geometry, truncated, corrupted).
"""
import numpy as np
import pytest

from helpers import assert_occ_bitexact, make_pair
from snapshot_format import decode

pytestmark = pytest.mark.gpu


def test_snapshot_round_trip_and_continue(cb, orc, gpu_ready, tmp_path):
    w = cb.WORKLOADS["small"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w)
    lo = (0.9, -0.3, -1.5, 3.0, -0.01, 1.2)
    g.set_log_odds_params(*lo)
    o.set_log_odds_params(*lo)
    g.set_max_raycast_range(8.0)
    o.set_max_raycast_range(8.0)
    scans = [scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, 8.0) for k in range(8)]
    for s in scans[:4]:
        g.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        o.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        g.recenter(s.base)
        o.recenter(s.base)
    blob = g.save_snapshot()
    h, occ, ids = decode(blob)
    assert h["encoding"] == 1 and ids is not None
    assert np.all(np.diff(ids) > 0)
    assert_occ_bitexact(occ, o.occupancy(), "decoded snapshot")
    assert tuple(h["origin_i"]) == o.origin()[0]
    assert tuple(h["half_map_size"]) == tuple(w.half)
    assert h["first_run"] == 0 and h["has_threshold"] == 1
    assert np.float32(h["max_raycast_range"]) == np.float32(8.0)
    assert len(blob) == 120 + 8 + 8 * ids.size

    # Restore into a fresh grid with default settings, then continue.
    g2 = cb.OccupancyGrid(w.half, w.resolution, w.origin_at_center, 0.5)
    g2.load_snapshot(blob)
    assert_occ_bitexact(g2.occupancy(), g.occupancy(), "restored")
    assert g2.geometry.origin_i == g.geometry.origin_i
    path = tmp_path / "grid.cbsnap"
    g.save_snapshot_file(str(path))
    g3 = cb.OccupancyGrid(w.half, w.resolution, w.origin_at_center, None)
    g3.load_snapshot_file(str(path))
    for s in scans[4:]:
        for x in (g, g2, g3, o):
            x.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
            x.recenter(s.base)
        want = o.occupancy()
        for name, x in (("original", g), ("restored", g2), ("from file", g3)):
            assert_occ_bitexact(x.occupancy(), want, f"{name} after restore")
            assert x.geometry.origin_i == o.origin()[0]


def test_snapshot_dense_encoding_and_negative_zero(cb, gpu_ready):
    half = (30, 20, 10)
    g = cb.OccupancyGrid(half, 0.1, False, 1.0)
    V = g.voxel_num
    rng = np.random.default_rng(12)
    occ = rng.uniform(-2, 3, V).astype(np.float32)
    occ[rng.random(V) < 0.3] = 0.0
    occ[::97] = np.float32(-0.0)
    g.load_occupancy(occ)
    blob = g.save_snapshot()
    h, dec, ids = decode(blob)
    assert h["encoding"] == 0 and ids is None
    assert np.array_equal(dec.view(np.uint32), occ.view(np.uint32))
    g2 = cb.OccupancyGrid(half, 0.1, False, 1.0)
    g2.load_snapshot(blob)
    assert np.array_equal(g2.occupancy().view(np.uint32), occ.view(np.uint32))
    # Mostly zero with some -0.0: the sparse form keeps the -0.0 voxels.
    sparse = np.zeros(V, np.float32)
    sparse[::50] = np.float32(-0.0)
    sparse[7::50] = 1.25
    g.load_occupancy(sparse)
    blob = g.save_snapshot()
    h, dec, ids = decode(blob)
    assert h["encoding"] == 1 and ids.size == len(range(0, V, 50)) + len(range(7, V, 50))
    g2.load_snapshot(blob)
    assert np.array_equal(g2.occupancy().view(np.uint32), sparse.view(np.uint32))


def test_snapshot_rejects_mismatch_and_damage(cb, gpu_ready):
    g = cb.OccupancyGrid((30, 20, 10), 0.1, False, 1.0)
    g.load_occupancy(np.ones(g.voxel_num, np.float32))
    blob = g.save_snapshot()
    for other in (cb.OccupancyGrid((30, 20, 11), 0.1, False, 1.0),
                  cb.OccupancyGrid((30, 20, 10), 0.2, False, 1.0),
                  cb.OccupancyGrid((30, 20, 10), 0.1, True, 1.0)):
        with pytest.raises(RuntimeError, match="differs"):
            other.load_snapshot(blob)
    with pytest.raises(RuntimeError, match="truncated|payload"):
        g.load_snapshot(blob[:100])
    with pytest.raises(RuntimeError, match="payload"):
        g.load_snapshot(blob[:-4])
    bad = bytearray(blob)
    bad[0:8] = b"NOTASNAP"
    with pytest.raises(RuntimeError, match="not a cublox snapshot"):
        g.load_snapshot(bytes(bad))
    # The grid is unchanged by the rejected loads.
    assert np.array_equal(g.occupancy(), np.ones(g.voxel_num, np.float32))
