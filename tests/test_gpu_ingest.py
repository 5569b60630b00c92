"""GPU parity of the driver ingest on the device (reference src/main.cpp):
PointCloud2 payloads decoded in the fused update's ray set-up with the
driver's filters (finite, within the max raycast range, the 1,000,000-record
cap) and transform, and depth images back-projected as depthCallback does.

Each case runs the product path and the oracle's restatement of the same
driver steps (buildCloudInOdomFrame, then update only when a point is kept)
and compares the grids bit-exactly after every update and recenter, and the
returned "updated" flags. Covered: host and device payloads, interleaved
records with permuted field offsets and NaN padding bytes, non-finite
points, misses beyond the range (dropped here, unlike the plain update), a
cloud longer than the cap, a first cloud with nothing kept (no origin pin),
16-bit and 32-bit float depth images with invalid pixels and a padded row
pitch, from host and device memory.
"""
import numpy as np
import pytest

from helpers import assert_occ_bitexact, make_pair

pytestmark = pytest.mark.gpu


def _records(cols, step, offs, fill=0xFF):
    n = len(cols[0])
    buf = np.full((n, step), fill, np.uint8)
    for c, off in zip(cols, offs):
        buf[:, off:off + 4] = np.ascontiguousarray(c, np.float32).view(np.uint8).reshape(-1, 4)
    return buf.reshape(-1)


def test_point_cloud2_small(cb, orc, gpu_ready):
    w = cb.WORKLOADS["small"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w)
    rng = np.random.default_rng(17)
    step, offs = 32, (16, 0, 8)
    for k in range(w.scans):
        s = scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
        xs, ys, zs = (np.array(c, np.float32) for c in (s.xs, s.ys, s.zs))
        bad = rng.choice(s.n, 200, replace=False)
        xs[bad[:70]] = np.nan
        ys[bad[70:140]] = np.inf
        zs[bad[140:]] = -np.inf
        rec = _records((xs, ys, zs), step, offs)
        print(f"[ingest cloud2] scan {k} n={s.n}", flush=True)
        if k % 2:
            buf = cb.DeviceBuffer.from_array(rec)
            ug = g.update_point_cloud2(buf.ptr, s.n, 1, step, offs, s.T, device=True)
        else:
            ug = g.update_point_cloud2(rec, s.n, 1, step, offs, s.T)
        uo = o.update_cloud2(rec, s.n, step, offs, s.T)
        assert ug == uo
        assert_occ_bitexact(g.occupancy(), o.occupancy(), f"cloud2 scan {k}")
        g.recenter(s.base)
        o.recenter(s.base)
        assert_occ_bitexact(g.occupancy(), o.occupancy(), f"cloud2 recenter {k}")


def test_point_cloud2_drops_misses(cb, orc, gpu_ready):
    """Misses beyond the range are free-space rays for the plain update but
    are dropped by the driver's filter: the two results must differ, and
    the ingest result must equal the plain update of the kept points."""
    w = cb.WORKLOADS["small"]
    s = w.scene().scan(0, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
    g1, _ = make_pair(cb, orc, w)
    g2, _ = make_pair(cb, orc, w)
    rec = _records((s.xs, s.ys, s.zs), 12, (0, 4, 8))
    assert g1.update_point_cloud2(rec, s.n, 1, 12, (0, 4, 8), s.T)
    r2 = np.float32(w.max_range) * np.float32(w.max_range)
    norm = (s.xs * s.xs + s.ys * s.ys) + s.zs * s.zs
    keep = norm <= r2
    assert (~keep).sum() > 100, "the scan has misses beyond the range"
    g2.update_sensor_frame((s.xs[keep], s.ys[keep], s.zs[keep]), s.T)
    assert_occ_bitexact(g1.occupancy(), g2.occupancy(), "ingest vs kept points")
    g3, _ = make_pair(cb, orc, w)
    g3.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
    assert not np.array_equal(g1.occupancy(), g3.occupancy())


def test_point_cloud2_cap_and_empty_first(cb, orc, gpu_ready):
    w = cb.WORKLOADS["small"]
    g, o = make_pair(cb, orc, w, threads=8)
    T0 = np.array([[1, 0, 0, 1.0], [0, 1, 0, 2.0], [0, 0, 1, 1.5]], np.float32)
    far = np.full((1000, 3), 2 * w.max_range, np.float32)
    far[::3] = np.nan
    rec = _records(far.T, 12, (0, 4, 8))
    assert not g.update_point_cloud2(rec, 1000, 1, 12, (0, 4, 8), T0)
    assert not o.update_cloud2(rec, 1000, 12, (0, 4, 8), T0)
    assert g.geometry.origin_i == (0, 0, 0) == o.origin()[0]
    # 1,050,000 points: the last 50,000 would mark a wall at x = +6 m.
    rng = np.random.default_rng(23)
    n, tail = 1_050_000, 50_000
    az = rng.uniform(-np.pi, np.pi, n).astype(np.float32)
    r = rng.uniform(1.0, 8.0, n).astype(np.float32)
    pts = np.stack([r * np.cos(az), r * np.sin(az), rng.uniform(-1, 1, n).astype(np.float32)], 1)
    pts[-tail:] = np.stack([np.full(tail, 6.0, np.float32),
                            rng.uniform(-3, 3, tail).astype(np.float32),
                            rng.uniform(-1, 1, tail).astype(np.float32)], 1)
    pts = pts.astype(np.float32)
    T1 = np.array([[0, -1, 0, 4.0], [1, 0, 0, -3.0], [0, 0, 1, 1.2]], np.float32)
    rec = _records(pts.T, 16, (0, 4, 8), fill=0)
    print("[ingest cap] 1.05 M points", flush=True)
    assert g.update_point_cloud2(rec, n, 1, 16, (0, 4, 8), T1)
    assert o.update_cloud2(rec, n, 16, (0, 4, 8), T1)
    assert g.geometry.origin_i == o.origin()[0] != (0, 0, 0)
    assert_occ_bitexact(g.occupancy(), o.occupancy(), "cap")
    # The same cloud without its tail gives the same grid.
    g2, _ = make_pair(cb, orc, w)
    rec2 = _records(pts[:1_000_000].T, 16, (0, 4, 8), fill=0)
    assert g2.update_point_cloud2(rec2, 1_000_000, 1, 16, (0, 4, 8), T1)
    assert np.array_equal(g.occupancy(), g2.occupancy())


def _depth_scene(kind, W, H, k, rng):
    u = np.arange(W, dtype=np.float32)[None, :]
    v = np.arange(H, dtype=np.float32)[:, None]
    depth = 1.5 + 5.5 * (v / H) + 0.8 * np.sin(u / 37.0 + k) + rng.normal(0, 0.02, (H, W))
    depth = depth.astype(np.float32)
    if kind == "u16":
        img = np.round(depth * 1000).astype(np.uint16)
        img[rng.random((H, W)) < 0.1] = 0
        return img
    img = depth.copy()
    m = rng.random((H, W))
    img[m < 0.03] = np.nan
    img[(m >= 0.03) & (m < 0.05)] = -1.0
    img[(m >= 0.05) & (m < 0.06)] = 0.0
    img[(m >= 0.06) & (m < 0.07)] = np.inf
    img[(m >= 0.07) & (m < 0.09)] = 12.0  # beyond the range: dropped
    return img


@pytest.mark.parametrize("kind", ["u16", "f32"])
def test_depth_images(cb, orc, gpu_ready, kind):
    w = cb.WORKLOADS["small"]
    g, o = make_pair(cb, orc, w, threads=8)
    W, H = 640, 480
    K = (525.0, 525.0, 319.5, 239.5)
    rng = np.random.default_rng(31 if kind == "u16" else 37)
    enc = 0 if kind == "u16" else 1
    elem = 2 if kind == "u16" else 4
    for k in range(4):
        img = _depth_scene(kind, W, H, k, rng)
        # Camera frame (x right, y down, z forward) to world (x forward,
        # y left, z up), moving along +x and turning slowly.
        c, s_ = np.cos(0.05 * k), np.sin(0.05 * k)
        Rz = np.array([[c, -s_, 0], [s_, c, 0], [0, 0, 1]])
        R = Rz @ np.array([[0, 0, 1], [-1, 0, 0], [0, -1, 0]])
        T = np.concatenate([R, [[0.4 * k], [0.1 * k], [1.3]]], 1).astype(np.float32)
        print(f"[ingest depth {kind}] frame {k}", flush=True)
        if k % 2:
            # Device image with a padded row pitch.
            pitch = W * elem + 64
            raw = np.zeros((H, pitch), np.uint8)
            raw[:, :W * elem] = img.view(np.uint8).reshape(H, W * elem)
            buf = cb.DeviceBuffer.from_array(raw.reshape(-1))
            ug = g.update_depth_image(buf.ptr, K, T, width=W, height=H, step=pitch,
                                      encoding=enc, device=True)
            uo = o.update_depth(raw.reshape(-1), W, H, pitch, enc, K, T)
        else:
            ug = g.update_depth_image(img, K, T)
            uo = o.update_depth(img.view(np.uint8).reshape(-1), W, H, W * elem, enc, K, T)
        assert ug and uo
        assert_occ_bitexact(g.occupancy(), o.occupancy(), f"depth {kind} frame {k}")
        base = T[:, 3].copy()
        base[2] = 0.0
        g.recenter(base)
        o.recenter(base)
        assert_occ_bitexact(g.occupancy(), o.occupancy(), f"depth {kind} recenter {k}")


def test_batched_ingest(cb, orc, gpu_ready):
    """GridBatch ticks with PointCloud2 payloads read in place (record
    stride, range drop) beside a plain column scan: every grid equals its
    oracle run alone."""
    specs = [  # half, resolution, max range, beams, azimuths, seed, layout
        ((40, 30, 12), 0.1, 6.0, 16, 512, 41, (32, (16, 0, 8))),
        ((37, 21, 9), 0.1, 12.0, 16, 256, 42, None),
        ((50, 50, 16), 0.05, 5.0, 16, 1024, 43, (16, (0, 4, 8))),
    ]
    lo = cb.log_odds()
    grids, oracles, scenes = [], [], []
    for half, res, rng_, _b, _a, seed, _l in specs:
        g = cb.OccupancyGrid(half, res, False, 1.0)
        o = orc.OracleGrid(half, res, False, 1.0)
        for x in (g, o):
            x.set_max_raycast_range(rng_)
            x.set_log_odds_params(*lo)
        grids.append(g)
        oracles.append(o)
        scenes.append(cb.Scene(seed=seed, n_boxes=40, box_extent=10.0, speed=0.5))
    total = sum(b * a for (_, _, _, b, a, _, _) in specs)
    batch = cb.GridBatch.for_grids(grids, total, max_near_steps=12)
    for k in range(4):
        scans, recenters, keep = [], [], []
        for i, (half, res, rng_, beams, az, seed, layout) in enumerate(specs):
            s = scenes[i].scan(k, beams, az, -15.0, 15.0, rng_)
            if layout is None:
                bufs = [cb.DeviceBuffer.from_array(np.ascontiguousarray(c))
                        for c in (s.xs, s.ys, s.zs)]
                keep.append(bufs)
                scans.append((grids[i], bufs[0], bufs[1], bufs[2], s.n, s.T))
                oracles[i].update_sensor_frame((s.xs, s.ys, s.zs), s.T)
            else:
                step, offs = layout
                rec = _records((s.xs, s.ys, s.zs), step, offs)
                buf = cb.DeviceBuffer.from_array(rec)
                keep.append(buf)
                r2 = float(np.float32(rng_) * np.float32(rng_))
                scans.append((grids[i], buf.ptr + offs[0], buf.ptr + offs[1],
                              buf.ptr + offs[2], s.n, s.T, step // 4, r2))
                assert oracles[i].update_cloud2(rec, s.n, step, offs, s.T)
            recenters.append((grids[i], s.base))
            oracles[i].recenter(s.base)
        print(f"[ingest batch] tick {k}", flush=True)
        batch.tick(scans, recenters)
        for i in range(len(specs)):
            assert_occ_bitexact(grids[i].occupancy(), oracles[i].occupancy(),
                                f"batched ingest tick {k} grid {i}")
    batch.close()
