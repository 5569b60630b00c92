"""Pins of the oracle's occupancy viz cloud (reference main.cpp:71-84
subsample, 275-304 colour map and packing, 740-817 build loop).

The colour map is checked at distances where every step is exact (the
ends, the midpoint and the quarter points of the range, beyond the range,
a range below 1 m), the subsample against Python-integer arithmetic on the
global indices (two's complement for negative indices), and the build loop
for order and the point limit.
"""
import numpy as np

import oracle as orc


def _rgb_bits(recs):
    return recs[:, 3].view(np.uint32)


def test_colour_map_exact_points():
    d_max = 10.0
    ds = [0.0, 10.0, 5.0, 2.5, 7.5, 14.0]
    pos = np.array([[d, 0, 0] for d in ds], np.float32)
    recs = orc.viz_points(pos, (0, 0, 0), d_max, 10.0, False, 1, 100)
    assert np.array_equal(recs[:, :3], pos)
    want = [0x00FF00, 0xFF0000, 0x0000FF, 0x007F7F, 0x7F007F, 0xFF0000]
    assert _rgb_bits(recs).tolist() == want
    # A range below 1 m counts as 1 m: 0.5 m is then the midpoint.
    small = orc.viz_points(np.array([[0, 0.5, 0]], np.float32), (0, 0, 0), 0.3,
                           10.0, False, 1, 10)
    assert _rgb_bits(small).tolist() == [0x0000FF]
    # Distances are from the robot position, in 3-D (3-4-5 triangle).
    tri = orc.viz_points(np.array([[4, 5, 3]], np.float32), (1, 1, 3), 10.0, 10.0,
                         False, 1, 10)
    assert _rgb_bits(tri).tolist() == [0x0000FF]


def _expected_keep(g, step):
    mix = (g[:, 0] & 0xFFFFFFFF) ^ (g[:, 1] & 0xFFFFFFFF) ^ (g[:, 2] & 0xFFFFFFFF)
    return (mix % step) == 0


def test_subsample_by_global_index():
    rng = np.random.default_rng(4)
    g = rng.integers(-300, 300, (2000, 3)).astype(np.int64)
    res = 0.1
    for center, pos in ((False, (g + 0.5) * res), (True, g * res)):
        pos = pos.astype(np.float32)
        for step in (1, 2, 3, 7):
            recs = orc.viz_points(pos, (0, 0, 0), 20.0, 10.0, center, step, 10**6)
            keep = _expected_keep(g, step) if step > 1 else np.ones(len(g), bool)
            assert np.array_equal(recs[:, :3], pos[keep]), (center, step)


def test_build_loop_limit_keeps_order():
    pos = np.stack([np.arange(50, dtype=np.float32) * 0.1 + 0.05,
                    np.full(50, 0.05, np.float32), np.full(50, 0.05, np.float32)], 1)
    recs = orc.viz_points(pos, (0, 0, 0), 5.0, 10.0, False, 2, 7)
    g = np.floor(pos * 10).astype(np.int64)
    keep = np.nonzero(_expected_keep(g, 2))[0][:7]
    assert np.array_equal(recs[:, :3], pos[keep])
