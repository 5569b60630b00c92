"""GPU parity of the occupancy viz cloud (the driver's OCCUPIED fetch,
main.cpp:484-491, and publishOccupancyCloudAsync, main.cpp:740-817, fused
into one launch).

The oracle fetches the same sphere and runs the driver's build loop over
its samples; with the point limit above the fetched count, the GPU records
must equal the oracle's as a multiset of (x, y, z, rgb) bit patterns (the
fetch order is unspecified in the reference too), and the fetched counts
must match. With the limit binding, which samples are fetched depends on
that order, so the GPU records must be a subset of the oracle's unlimited
records, no more than the limit, and the fetched count must be the limit.
"""
import numpy as np
import pytest

from helpers import make_pair

pytestmark = pytest.mark.gpu


def _rows(recs):
    b = np.ascontiguousarray(recs, np.float32).view(np.uint32).reshape(-1, 4)
    return b[np.lexsort(b.T[::-1])]


def _oracle_viz(orc, o, robot, radius, step, max_points, w, center):
    cap = max_points if max_points > 0 else 500000
    pos, _st, n = o.fetch(robot, radius, 2, max_points, cap=cap)
    recs = orc.viz_points(pos, robot, radius, 1.0 / w.resolution, center,
                          step, max(min(len(pos), cap), 1))
    return recs, min(n, cap)


@pytest.mark.parametrize("center", [False, True])
def test_viz_cloud_readme(cb, orc, gpu_ready, center):
    w = cb.WORKLOADS["readme"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w, center=center, threads=16)
    for k in range(3):
        s = scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
        g.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        o.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        g.recenter(s.base)
        o.recenter(s.base)
    robot = np.asarray(s.base, np.float32)
    for radius, step, limit in ((25.0, 1, 1_000_000), (25.0, 3, 1_000_000),
                                (12.0, 2, 1_000_000), (0.5, 1, 1_000_000)):
        print(f"[viz] center={center} r={radius} step={step}", flush=True)
        got, fetched = g.occupancy_viz_cloud(robot, radius, step, limit)
        want, n = _oracle_viz(orc, o, robot, radius, step, limit, w, center)
        assert fetched == n, (radius, step)
        assert n < limit
        assert np.array_equal(_rows(got), _rows(want)), (radius, step)
    # The limit binds: fetched == limit, records a subset of the full set.
    full, n_full = _oracle_viz(orc, o, robot, 25.0, 2, 1_000_000, w, center)
    limit = n_full // 3
    got, fetched = g.occupancy_viz_cloud(robot, 25.0, 2, limit)
    assert fetched == limit
    assert 0 < len(got) <= limit
    full_rows = {tuple(r) for r in _rows(full).tolist()}
    assert all(tuple(r) in full_rows for r in _rows(got).tolist())
