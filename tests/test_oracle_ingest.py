"""Pins of the oracle's driver ingest (reference src/main.cpp:558-693):
PointCloud2 decode, the finite and range filters, the 1,000,000-record cap,
the odom-frame transform, and depth-image back-projection.

Each check uses a case whose answer is fixed without restating the
formula: exact special values (principal-point pixels, unit focal lengths,
depths that round to whole metres, permutation transforms with integer
translations), the range boundary at an exactly representable squared
norm, a sign symmetry, layout invariance (field offsets and padding must
not change the result), and what each filter leaves (ordering, counts).
"""
import numpy as np
import pytest

import oracle as orc

NAN = np.float32(np.nan)


def _records(pts, step=12, offs=(0, 4, 8), fill=0):
    pts = np.asarray(pts, np.float32).reshape(-1, 3)
    buf = np.full((len(pts), step), fill, np.uint8)
    for k, off in enumerate(offs):
        buf[:, off:off + 4] = pts[:, k:k + 1].view(np.uint8).reshape(-1, 4)
    return buf.reshape(-1)


EYE = np.eye(3, 4, dtype=np.float32)


def test_non_finite_points_dropped_in_order():
    pts = np.array([[1, 2, 3], [NAN, 0, 0], [0, np.inf, 0], [0, 0, -np.inf],
                    [4, 5, 6], [np.nan, np.nan, np.nan], [-1, -2, -3]], np.float32)
    xw, yw, zw = orc.build_cloud(_records(pts), len(pts), 12, (0, 4, 8), EYE, 1e6)
    kept = pts[[0, 4, 6]]
    assert np.array_equal(np.stack([xw, yw, zw], 1), kept)


def test_range_boundary_is_inclusive():
    r2 = np.float32(625.0)  # 25 m
    pts = np.array([[25, 0, 0], [15, 20, 0], [0, 15, 20], [np.nextafter(np.float32(25), np.float32(30)), 0, 0],
                    [15, 20, 0.01], [0, 0, -25]], np.float32)
    xw, _, _ = orc.build_cloud(_records(pts), len(pts), 12, (0, 4, 8), EYE, r2)
    # 625 == 625 kept; anything above dropped (0.01^2 survives the rounding
    # of 625 + 1e-4, a smaller z would not).
    assert len(xw) == 4
    assert np.array_equal(xw, np.array([25, 15, 0, 0], np.float32))


def test_record_cap_of_one_million():
    n = 1_000_010
    pts = np.zeros((n, 3), np.float32)
    pts[:, 0] = np.arange(n, dtype=np.float32) % 1000
    pts[:, 2] = 1.0
    xw, _, zw = orc.build_cloud(_records(pts), n, 12, (0, 4, 8), EYE, 1e12)
    assert len(xw) == 1_000_000
    assert np.array_equal(xw, pts[:1_000_000, 0])
    # A cap that cuts a filtered tail: rows counted after the cap.
    pts2 = pts.copy()
    pts2[:10, 0] = np.nan
    xw2, _, _ = orc.build_cloud(_records(pts2), n, 12, (0, 4, 8), EYE, 1e12)
    assert len(xw2) == 1_000_000 - 10


def test_field_offsets_and_padding_do_not_matter():
    rng = np.random.default_rng(3)
    pts = rng.uniform(-20, 20, (500, 3)).astype(np.float32)
    pts[::17] = np.nan
    T = np.array([[0.6, -0.8, 0, 1.5], [0.8, 0.6, 0, -2.25], [0, 0, 1, 0.75]], np.float32)
    ref = orc.build_cloud(_records(pts), len(pts), 12, (0, 4, 8), T, np.float32(400))
    # x at 16, y at 0, z at 8 in 32-byte records whose padding holds NaN bytes.
    rec = _records(pts, step=32, offs=(16, 0, 8), fill=0xFF)
    got = orc.build_cloud(rec, len(pts), 32, (16, 0, 8), T, np.float32(400))
    for a, b in zip(ref, got):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_permutation_transform_exact():
    # Rotation by +90 degrees about z plus an integer translation: every
    # product is by 0 or +-1, so x' = -y + 3, y' = x - 7, z' = z + 2 exactly.
    T = np.array([[0, -1, 0, 3], [1, 0, 0, -7], [0, 0, 1, 2]], np.float32)
    rng = np.random.default_rng(5)
    pts = rng.integers(-50, 50, (200, 3)).astype(np.float32) * np.float32(0.25)
    xw, yw, zw = orc.build_cloud(_records(pts), len(pts), 12, (0, 4, 8), T, 1e9)
    assert np.array_equal(xw, -pts[:, 1] + 3)
    assert np.array_equal(yw, pts[:, 0] - 7)
    assert np.array_equal(zw, pts[:, 2] + 2)


def test_depth_u16_special_values():
    W, H = 8, 6
    img = np.zeros((H, W), np.uint16)
    img[3, 4] = 1000          # principal point, 1 m
    img[3, 2] = 2000          # 2 px left of it, 2 m
    img[3, 6] = 2000          # mirror image
    img[1, 4] = 500           # 0.5 m, 2 rows above
    K = (1.0, 1.0, 4.0, 3.0)
    xs, ys, zs = orc.depth_to_cloud(img.view(np.uint8).reshape(-1), W, H, W * 2, 0, K)
    p = lambda u, v: (xs[v * W + u], ys[v * W + u], zs[v * W + u])
    assert p(4, 3) == (0.0, 0.0, 1.0)
    assert p(2, 3) == (-4.0, 0.0, 2.0)
    assert p(6, 3) == (4.0, 0.0, 2.0)
    assert p(4, 1) == (0.0, -1.0, 0.5)
    zero = img.reshape(-1) == 0
    assert np.isnan(xs[zero]).all() and np.isnan(ys[zero]).all() and np.isnan(zs[zero]).all()
    assert not np.isnan(zs[~zero]).any()


def test_depth_row_pitch_and_symmetry():
    W, H, pad = 10, 7, 6
    rng = np.random.default_rng(9)
    img = rng.integers(0, 9000, (H, W)).astype(np.uint16)
    K = (525.0, 525.0, 4.5, 3.0)
    a = orc.depth_to_cloud(img.view(np.uint8).reshape(-1), W, H, W * 2, 0, K)
    padded = np.zeros((H, W * 2 + pad), np.uint8)
    padded[:, :W * 2] = img.view(np.uint8).reshape(H, W * 2)
    padded[:, W * 2:] = 0xAB
    b = orc.depth_to_cloud(padded.reshape(-1), W, H, W * 2 + pad, 0, K)
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    # Mirrored columns about cx = 4.5 give x of opposite sign, same y and z.
    m = orc.depth_to_cloud(img[:, ::-1].copy().view(np.uint8).reshape(-1), W, H, W * 2, 0, K)
    xa = a[0].reshape(H, W)
    xm = m[0].reshape(H, W)[:, ::-1]
    ok = ~np.isnan(xa)
    assert np.array_equal(xa[ok], -xm[ok])


def test_depth_f32_invalid_rules():
    vals = np.array([[1.5, 0.0, -0.0, -2.0],
                     [np.nan, np.inf, -np.inf, 3.25]], np.float32)
    K = (2.0, 2.0, 0.0, 0.0)
    xs, ys, zs = orc.depth_to_cloud(vals.view(np.uint8).reshape(-1), 4, 2, 16, 1, K)
    valid = np.array([True, False, False, False, False, False, False, True])
    assert np.isnan(zs[~valid]).all()
    assert np.array_equal(zs[valid], vals.reshape(-1)[valid])
    # u = 0 column: x = 0; pixel (3, 1) with fx = 2: x = 3 * 3.25 / 2.
    assert xs[0] == 0.0 and xs[7] == np.float32(4.875) and ys[7] == np.float32(1.625)


def test_oracle_grid_skips_update_when_nothing_kept():
    o = orc.OracleGrid((20, 20, 10), 0.1, False, 1.0)
    o.set_max_raycast_range(5.0)
    far = np.array([[6, 0, 0], [0, 7, 0], [np.nan, 0, 0]], np.float32)
    T = np.array([[1, 0, 0, 3], [0, 1, 0, 4], [0, 0, 1, 1]], np.float32)
    assert not o.update_cloud2(_records(far), 3, 12, (0, 4, 8), T)
    assert o.origin()[0] == (0, 0, 0)
    assert not o.occupancy().any()
    near = np.array([[1, 0, 0]], np.float32)
    assert o.update_cloud2(_records(near), 1, 12, (0, 4, 8), T)
    assert o.origin()[0] == (30, 40, 10)
