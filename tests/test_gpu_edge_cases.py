"""Edge cases of the world-frame update against the oracle (reference
OccupancyGrid.cpp:181-238 and the ray cast of kernels.cu:39-151), on both
update paths and both binnings:
  * an empty cloud is a no-op and does not pin the window origin;
  * zero-length and sub-1e-6 rays, non-finite endpoints and endpoints whose
    squared length overflows are skipped;
  * endpoints on voxel faces, axis-aligned rays, and diagonal rays with
    exactly equal crossing parameters (the tie-break of the step choice);
  * rays of exactly the maximum range (kept whole) and just beyond it
    (clipped, no hit), and rays that leave the window (a longer range).
Every coordinate is an exact binary fraction where the case needs it.
"""
import numpy as np
import pytest

from helpers import assert_occ_bitexact, make_pair

pytestmark = pytest.mark.gpu

S = np.array([0.25, -0.75, 1.5], np.float32)


def _cloud():
    inf, nan = np.float32(np.inf), np.float32(np.nan)
    rel = [
        (0, 0, 0), (0, 0, 0), (1e-7, 0, 0), (0, -1e-7, 5e-8),  # degenerate
        (nan, 0, 0), (inf, 1, 1), (1, -inf, 1), (1, 1, nan),    # non-finite
        (1e30, 0, 1), (3e19, 0, 0), (1e18, 2e18, -1e18),        # overflow / huge
        (1.75, 0, 0), (0, 4.25, 0), (0, 0, -2.0), (-3.0, 0, 0),  # axis-aligned
        (1.0, 1.0, 0), (-2.0, 2.0, 0), (1.5, -1.5, -1.5), (0.5, 0.5, 0.5),  # ties
        (-2.75, -2.25, -1.0), (3.125, -0.375, 0.625),           # on voxel faces
        (10.0, 0, 0), (0, 6.0, 8.0), (10.0001, 0, 0), (0, 6.0001, 8.0),  # range
        (-9.5, 2.0, 0.5), (0.5, 0.2, 6.0), (7.0, -7.0, 0.0),    # window edges
    ]
    pts = S + np.array(rel, np.float32)
    # Ordinary rays around them, so the tile and hot paths are exercised too.
    rng = np.random.default_rng(11)
    az = rng.uniform(0, 2 * np.pi, 3000)
    el = rng.uniform(-0.5, 0.3, 3000)
    r = rng.uniform(0.5, 14.0, 3000)
    more = S + np.stack([r * np.cos(el) * np.cos(az), r * np.cos(el) * np.sin(az),
                         r * np.sin(el)], 1).astype(np.float32)
    return np.concatenate([pts, more]).astype(np.float32)


@pytest.mark.parametrize("max_range", [10.0, 40.0])
@pytest.mark.parametrize("center", [False, True])
@pytest.mark.parametrize("path", [0, 1])
def test_update_edge_cases(cb, orc, gpu_ready, path, center, max_range):
    w = cb.WORKLOADS["small"]
    g, o = make_pair(cb, orc, w, center=center, max_range=max_range)
    g.set_update_path(path)
    print(f"[edge] path={path} center={center} max_range={max_range}", flush=True)
    empty = np.zeros((0, 3), np.float32)
    for x in (g, o):
        x.update(empty, S + np.float32(5.0))
    assert not g.occupancy().any()
    pts = _cloud()
    for k, origin in enumerate([S, S + np.array([0.5, -0.25, 0.0], np.float32)]):
        for x in (g, o):
            x.update(pts, origin)
        # The first non-empty cloud pins the origin from its own sensor.
        assert g.geometry.origin_i == o.origin()[0]
        assert_occ_bitexact(g.occupancy(), o.occupancy(), f"update {k}")
    # Only degenerate rays: nothing changes.
    before = g.occupancy()
    bad = _cloud()[:10]
    for x in (g, o):
        x.update(bad, S)
    assert_occ_bitexact(g.occupancy(), before, "degenerate rays only")
    assert_occ_bitexact(g.occupancy(), o.occupancy(), "degenerate rays only")
