"""Regression pins against the committed fixtures of tests/golden (written
by tests/golden/make_golden.py from this repository's generator and
oracle). The GPU half compares the HIP path with the same fixtures."""
import json
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(HERE, name)) as f:
        return json.load(f)


def test_scangen_fixture(cb):
    from golden.make_golden import scangen_fixture
    assert scangen_fixture() == _load("scangen.json")


def test_small_oracle_fixture(cb, orc):
    from golden.make_golden import small_oracle_fixture
    got = small_oracle_fixture()
    want = _load("small_oracle.json")
    for k, (a, b) in enumerate(zip(got["scans"], want["scans"])):
        assert a == b, f"scan {k}"
    assert got["fetches"] == want["fetches"]


@pytest.mark.gpu
def test_gpu_small_matches_fixture(cb, gpu_ready):
    """The HIP path (fused update) reproduces the oracle's digests."""
    from golden.make_golden import multiset_digest, sha
    want = _load("small_oracle.json")
    w = cb.WORKLOADS["small"]
    sc = w.scene()
    g = cb.OccupancyGrid(w.half, w.resolution, w.origin_at_center, w.threshold)
    g.set_max_raycast_range(w.max_range)
    g.set_log_odds_params(*cb.log_odds())
    for k in range(w.scans):
        s = sc.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
        g.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        assert sha(g.occupancy()) == want["scans"][k]["occ_apply"], f"apply {k}"
        g.recenter(s.base)
        assert sha(g.occupancy()) == want["scans"][k]["occ_recenter"], f"recenter {k}"
        assert list(g.geometry.origin_i) == want["scans"][k]["origin"]
    for f in want["fetches"]:
        c = np.array(f["center"], np.uint32).view(np.float32)
        pos, st = g.fetch_occupancy_around(c, 1.5, None if f["filter"] < 0 else f["filter"])
        assert len(st) == f["count"]
        assert multiset_digest(pos, st) == f["digest"]
