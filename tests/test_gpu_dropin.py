"""A C++ caller of the drop-in headers (tests/cpp/dropin_caller.cpp, built
by the library's Makefile and linked against libcublox.so) runs the
reference perf test's grid and cloud (test/test_occupancy_grid_perf.cpp:
54-100) and the driver's call pattern (src/main.cpp:437-498); every grid it
leaves, every fetch it returns and every point query it answers must match
the oracle replaying the calls it recorded.
"""
import os
import subprocess

import numpy as np
import pytest

from helpers import assert_occ_bitexact, sample_multiset

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CALLER = os.path.join(ROOT, "tests", "cpp", "dropin_caller")
# OccupancyGrid.hpp defaults (reference OccupancyGrid.hpp:162-168).
HEADER_LOG_ODDS = (0.847, -0.405, -1.992, 3.476, -0.0004, 1.7346)


def _f32(path, shape=None):
    a = np.fromfile(path, np.float32)
    return a if shape is None else a.reshape(shape)


def test_cpp_caller_matches_oracle(cb, orc, gpu_ready, tmp_path):
    if not os.path.exists(CALLER):
        pytest.fail("tests/cpp/dropin_caller is not built (run the library build)")
    out = subprocess.run([CALLER, str(tmp_path), "1"], capture_output=True, text=True,
                         timeout=300)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0, out.stderr

    # A: two updates of the 500,000-point shell from the origin, no threshold.
    cloud = _f32(tmp_path / "a_cloud.f32")
    n = cloud.size // 3
    cols = cloud.reshape(3, n)  # column-major N x 3: x column, y column, z column
    o = orc.OracleGrid((200, 200, 100), 0.05, True, None, threads=16)
    o.set_max_raycast_range(25.0)
    o.set_log_odds_params(*[float(np.float32(v)) for v in HEADER_LOG_ODDS])
    pts = np.ascontiguousarray(cols.T)
    for _ in range(2):
        o.update(pts, (0.0, 0.0, 0.0))
    assert_occ_bitexact(_f32(tmp_path / "a_occ.f32"), o.occupancy(), "perf-test grid")
    assert np.count_nonzero(o.occupancy()) > 1_000_000
    del o

    # B: the driver's loop with recenters, then fetches and queries.
    lo = [orc.logit(p) for p in (0.9, 0.49, 0.12, 0.98, 0.499, 0.85)]
    o = orc.OracleGrid((100, 100, 25), 0.1, False, 2.0)
    o.set_max_raycast_range(10.0)
    o.set_log_odds_params(*lo)
    meta = _f32(tmp_path / "b_meta.f32", (-1, 6))
    moves = 0
    for k, m in enumerate(meta):
        c = _f32(tmp_path / f"b_cloud{k}.f32")
        pts = np.ascontiguousarray(c.reshape(3, -1).T)
        o.update(pts, m[:3])
        moves += int(o.recenter(m[3:]) != 0)
    assert moves >= 1
    assert_occ_bitexact(_f32(tmp_path / "b_occ.f32"), o.occupancy(), "driver grid")
    robot = np.array([4.2, 1.4, 0.0], np.float32)
    for q, filt in enumerate((cb.OCCUPIED, None)):
        pos = _f32(tmp_path / f"b_fetch{q}.f32", (-1, 3))
        st = np.fromfile(tmp_path / f"b_fetch{q}.u8", np.uint8)
        op_, os_, on = o.fetch(robot, 3.0, filt)
        assert len(st) == on > 0, (q, len(st), on)
        assert np.array_equal(sample_multiset(pos, st), sample_multiset(op_, os_))
    occ = o.occupancy()
    geo_o = np.array(o.origin()[0])
    half = np.array((100, 100, 25))
    m3 = 2 * half + 1
    qp = _f32(tmp_path / "b_queries.f32", (-1, 3))
    qr = np.fromfile(tmp_path / "b_queries.u8", np.uint8).reshape(-1, 3)
    inv = np.float32(1.0) / np.float32(0.1)
    l_free, l_occ = np.float32(lo[4]), np.float32(lo[5])
    for p, r in zip(qp, qr):
        gi = np.array([orc.pos_to_global(float(p[a]), float(inv)) for a in range(3)])
        if np.any(np.abs(gi - geo_o) > half):
            exp = (0, 0, 1)
        else:
            v = occ[orc.global_to_hash(gi, m3, half)]
            exp = (int(v >= l_occ), int(v < l_free), int(l_free <= v < l_occ))
        assert tuple(int(x) for x in r) == exp, (p, r, exp)
