"""GPU parity of the changed-voxel stream (the reference's modified-voxel
compaction in applyUpdateKernel, kernels.cu:202-209, which its launcher
gates off, kernels.cu:251-262; enabled here for both update paths and the
batched update).

After every update the stream must hold exactly the voxels whose stored
value differs from its value before that update, each with its new value:
the expected set is the difference of the oracle grid's occupancy before and
after the same update, compared as a set of (hash id, value bits). Recenter
zeroing is not part of the stream. Covered: the fused path on the small and
README shapes, the per-voxel path (capacity of at least the voxel count,
as the reference gates it), truncation at a small capacity (count exact,
entries a subset), the stream turned off, and batched ticks with grids with
and without a stream, eager and replayed from a HIP graph.
"""
import numpy as np
import pytest

from helpers import assert_occ_bitexact, make_pair

pytestmark = pytest.mark.gpu


def _upload(cb, s):
    return [cb.DeviceBuffer.from_array(np.ascontiguousarray(c))
            for c in (s.xs, s.ys, s.zs)]


def _expected(before, after):
    ids = np.nonzero(after != before)[0].astype(np.int32)
    return ids, after[ids]


def _check_stream(got, before, after, label, truncated_to=None):
    ids, vals, count = got
    e_ids, e_vals = _expected(before, after)
    assert count == e_ids.size, f"{label}: count {count} != {e_ids.size}"
    if truncated_to is None:
        assert ids.size == count, f"{label}: {ids.size} entries for {count}"
    else:
        assert ids.size == min(truncated_to, count), label
    assert np.unique(ids).size == ids.size, f"{label}: duplicate entries"
    pos = np.searchsorted(e_ids, ids)
    assert (pos < e_ids.size).all() and np.array_equal(e_ids[np.minimum(pos, e_ids.size - 1)], ids), \
        f"{label}: entries for voxels that did not change"
    assert np.array_equal(vals.view(np.uint32), e_vals[pos].view(np.uint32)), \
        f"{label}: stream values differ from the new occupancy"


@pytest.mark.parametrize("path", [0, 1])
def test_changed_voxels_small(cb, orc, gpu_ready, path):
    w = cb.WORKLOADS["small"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w)
    g.set_update_path(path)
    g.set_changed_capacity(g.voxel_num)
    nonempty = 0
    for k in range(w.scans):
        s = scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
        before = o.occupancy()
        g.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        o.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        after = o.occupancy()
        assert_occ_bitexact(g.occupancy(), after, f"path {path} scan {k}")
        got = g.changed_voxels()
        _check_stream(got, before, after, f"path {path} scan {k}")
        nonempty += got[2] > 0
        g.recenter(s.base)
        o.recenter(s.base)
    assert nonempty == w.scans


def test_changed_voxels_truncated_and_off(cb, orc, gpu_ready):
    w = cb.WORKLOADS["small"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w)
    cap = 1000
    g.set_changed_capacity(cap)
    for k in range(3):
        s = scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
        before = o.occupancy()
        g.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        o.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        after = o.occupancy()
        got = g.changed_voxels()
        assert got[2] > cap, "the scan changes more voxels than the capacity"
        _check_stream(got, before, after, f"truncated scan {k}", truncated_to=cap)
    # The per-voxel path keeps the reference's gate: capacity >= voxels.
    with pytest.raises(RuntimeError, match="capacity"):
        g.set_update_path(1)
    s = scene.scan(3, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
    g.set_changed_capacity(0)
    g.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
    ids, vals, count = g.changed_voxels()
    assert count == 0 and ids.size == 0


def test_changed_voxels_readme(cb, orc, gpu_ready):
    w = cb.WORKLOADS["readme"]
    scene = w.scene()
    g, o = make_pair(cb, orc, w, threads=16)
    g.set_changed_capacity(1 << 24)
    for k in range(2):
        s = scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
        print(f"[delta readme] scan {k}", flush=True)
        before = o.occupancy()
        bx, by, bz = _upload(cb, s)
        g.update_sensor_frame_device(bx.ptr, by.ptr, bz.ptr, s.n, s.T)
        o.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        after = o.occupancy()
        got = g.changed_voxels()
        assert got[2] <= 1 << 24
        _check_stream(got, before, after, f"readme scan {k}")
        g.recenter(s.base)
        o.recenter(s.base)


# half, resolution, centre binning, recenter threshold, max range, beams,
# azimuths, speed (m/scan), scene seed, stream capacity (0: none)
GRIDS = [
    ((40, 30, 12), 0.1, False, 1.0, 6.0, 16, 512, 0.5, 21, 1 << 20),
    ((37, 21, 9), 0.1, True, 1.0, 12.0, 16, 256, 0.7, 22, 0),
    ((60, 60, 20), 0.1, False, 2.0, 8.0, 32, 512, 1.0, 23, 1 << 20),
    ((50, 50, 16), 0.05, True, 0.5, 5.0, 16, 1024, 0.4, 24, 500),
]


@pytest.mark.parametrize("graph", [False, True])
def test_changed_voxels_batched(cb, orc, gpu_ready, graph):
    lo = cb.log_odds()
    grids, oracles, scenes = [], [], []
    for (half, res, ctr, thr, rng, _b, _a, speed, seed, cap) in GRIDS:
        g = cb.OccupancyGrid(half, res, ctr, thr)
        o = orc.OracleGrid(half, res, ctr, thr)
        for x in (g, o):
            x.set_max_raycast_range(rng)
            x.set_log_odds_params(*lo)
        if cap:
            g.set_changed_capacity(cap)
        grids.append(g)
        oracles.append(o)
        scenes.append(cb.Scene(seed=seed, n_boxes=40, box_extent=10.0, speed=speed))
    total = sum(b * a for (_, _, _, _, _, b, a, _, _, _) in GRIDS)
    with pytest.raises(RuntimeError, match="changed_voxels"):
        plain = cb.GridBatch.for_grids(grids, total, max_near_steps=12)
        s = scenes[0].scan(0, GRIDS[0][5], GRIDS[0][6], -15.0, 15.0, GRIDS[0][4])
        bufs = _upload(cb, s)
        plain.tick([(grids[0], bufs[0], bufs[1], bufs[2], s.n, s.T)], [])
    batch = cb.GridBatch.for_grids(grids, total, max_near_steps=12,
                                   changed_voxels=True)
    if graph:
        batch.set_graph_mode(True)
    for k in range(4):
        scans, recenters, keep, before, after = [], [], [], [], []
        for i, (half, res, ctr, thr, rng, beams, az, speed, seed, cap) in enumerate(GRIDS):
            s = scenes[i].scan(k, beams, az, -15.0, 15.0, rng)
            bufs = _upload(cb, s)
            keep.append(bufs)
            scans.append((grids[i], bufs[0], bufs[1], bufs[2], s.n, s.T))
            recenters.append((grids[i], s.base))
            before.append(oracles[i].occupancy())
            oracles[i].update_sensor_frame((s.xs, s.ys, s.zs), s.T)
            after.append(oracles[i].occupancy())
            oracles[i].recenter(s.base)
        print(f"[delta batch] tick {k} graph={graph}", flush=True)
        batch.tick(scans, recenters)
        for i, spec in enumerate(GRIDS):
            assert_occ_bitexact(grids[i].occupancy(), oracles[i].occupancy(),
                                f"tick {k} grid {i}")
            cap = spec[-1]
            if cap:
                got = grids[i].changed_voxels()
                trunc = cap if got[2] > cap else None
                _check_stream(got, before[i], after[i], f"tick {k} grid {i}",
                              truncated_to=trunc)
    batch.close()
