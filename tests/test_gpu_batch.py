"""GPU parity of batched multi-grid updates (cublox::GridBatch).

Every tick updates all grids in one launch sequence and recenters them in
another; each grid must end every tick bit-identical to the oracle grid
updated alone with the same scans. Covered: grids of different shapes,
resolutions, ranges, corner and centre binning, a grid without a recenter
threshold, a z axis shorter than one tile, HIP-graph replay, and the
autotuner switching settings between ticks.
"""
import numpy as np
import pytest

from helpers import assert_occ_bitexact

pytestmark = pytest.mark.gpu

# half, resolution, centre binning, recenter threshold, max range, beams,
# azimuths, speed (m/scan), scene seed
GRIDS = [
    ((40, 30, 12), 0.1, False, 1.0, 6.0, 16, 512, 0.5, 11),
    ((37, 21, 9), 0.1, True, 1.0, 12.0, 16, 256, 0.7, 12),
    ((60, 60, 20), 0.1, False, 2.0, 8.0, 32, 512, 1.0, 13),
    ((25, 45, 10), 0.2, False, None, 10.0, 8, 512, 0.5, 14),
    ((50, 50, 16), 0.05, True, 0.5, 5.0, 16, 1024, 0.4, 15),
]


def _setup(cb, orc):
    lo = cb.log_odds()
    grids, oracles, scenes = [], [], []
    for (half, res, ctr, thr, rng, *_rest, seed) in GRIDS:
        g = cb.OccupancyGrid(half, res, ctr, thr)
        o = orc.OracleGrid(half, res, ctr, thr)
        for x in (g, o):
            x.set_max_raycast_range(rng)
            x.set_log_odds_params(*lo)
        grids.append(g)
        oracles.append(o)
        scenes.append(cb.Scene(seed=seed, n_boxes=40, box_extent=10.0,
                               speed=_rest[2]))
    return grids, oracles, scenes


def _run(cb, orc, ticks, graph=False, tuning=None, max_tiles=None, max_points=None,
         setting=None, pipelined=False, check_every=1, max_slabs=None, prepared=False):
    """Ticks every grid with its own scan and recenter; checks every grid
    against its oracle after every check_every-th tick (and the last).
    Pipelined: ticks are submitted without waits between the checks, so
    consecutive ticks overlap. Prepared: the tick's arguments are packed by
    GridBatch.prepare and passed to tick_prepared / submit_prepared."""
    grids, oracles, scenes = _setup(cb, orc)
    total = max_points or sum(b * a for (_, _, _, _, _, b, a, _, _) in GRIDS)
    kw = {} if max_tiles is None else {"max_tiles": max_tiles}
    if max_slabs is not None:
        kw["max_slabs"] = max_slabs
    batch = cb.GridBatch.for_grids(grids, total, max_near_steps=12, min_chunk=256,
                                   **kw)
    if graph:
        batch.set_graph_mode(True)
    if tuning:
        batch.set_tuning(tuning, 1)
    if setting:
        batch.set_setting(*setting)
    if pipelined:
        batch.set_pipelined(True)
    live = []  # device columns of the ticks not yet waited for
    for k in range(ticks):
        scans, recenters, keep = [], [], []
        for i, (half, res, ctr, thr, rng, beams, az, speed, seed) in enumerate(GRIDS):
            s = scenes[i].scan(k, beams, az, -15.0, 15.0, rng)
            bufs = [cb.DeviceBuffer.from_array(np.ascontiguousarray(c))
                    for c in (s.xs, s.ys, s.zs)]
            keep.append(bufs)
            scans.append((grids[i], bufs[0], bufs[1], bufs[2], s.n, s.T))
            recenters.append((grids[i], s.base))
            oracles[i].update_sensor_frame((s.xs, s.ys, s.zs), s.T)
            oracles[i].recenter(s.base)
        print(f"[batch] tick {k} graph={graph} pipelined={pipelined} "
              f"setting={batch.setting}", flush=True)
        p = batch.prepare(scans, recenters) if prepared else None
        if not pipelined:
            if p is not None:
                batch.tick_prepared(p)
            else:
                batch.tick(scans, recenters)
        else:
            live.append(keep)
            if p is not None:
                batch.submit_prepared(p)
            else:
                batch.submit(scans, recenters)
            if (k + 1) % check_every and k + 1 < ticks:
                continue
            batch.wait()
            live.clear()
        for i in range(len(GRIDS)):
            assert grids[i].geometry.origin_i == oracles[i].origin()[0], (k, i)
            assert_occ_bitexact(grids[i].occupancy(), oracles[i].occupancy(),
                                f"tick {k} grid {i}")
    return batch


def test_batch_matches_oracle_per_grid(cb, orc, gpu_ready):
    _run(cb, orc, 6)


def test_batch_graph_replay_matches(cb, orc, gpu_ready):
    _run(cb, orc, 6, graph=True)


@pytest.mark.parametrize("pipelined", [False, True])
def test_batch_prepared_ticks_match_oracle(cb, orc, gpu_ready, pipelined):
    _run(cb, orc, 5, graph=True, pipelined=pipelined, check_every=2, prepared=True)


def test_batch_multiblock_tile_scan(cb, orc, gpu_ready):
    """Tile capacity above one workgroup's stage: the three-pass scan."""
    _run(cb, orc, 4, max_tiles=20000)


def test_batch_large_bin_blocks(cb, orc, gpu_ready):
    """Limits for 2M points per tick: the binning takes 1024 rays per
    workgroup (each grid's ray range starts at a multiple of 1024)."""
    _run(cb, orc, 3, max_points=1 << 21)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_autotuner_keeps_results(cb, orc, gpu_ready, graph):
    """The tuner switches near steps, chunk, lanes per ray and the hot-tile
    mode between ticks (with graph replay: the captured sequence follows the
    staged header); every tick stays bit-exact."""
    cands = [(4, 256, 1, 0), (8, 512, 4, 1), (12, 1024, 1, 1), (8, 2048, 4, 0),
             (8, 256, 0, 1)]
    batch = _run(cb, orc, 12, graph=graph, tuning=cands)
    assert not batch.tuning
    assert batch.setting in cands
    assert all(ms > 0 for ms in batch.tune_results(len(cands)))


@pytest.mark.parametrize("lanes", [1, 4])
def test_batch_shared_hot_tiles(cb, orc, gpu_ready, lanes):
    """Hot tiles split into 256-record chunks that add into one shared slab
    pair per tile with global atomics, under both walks."""
    batch = _run(cb, orc, 4, setting=(8, 256, lanes, 1))
    assert batch.counters()["split_tiles"] > 0



@pytest.mark.parametrize("graph", [False, True])
def test_batch_pipelined_matches_oracle(cb, orc, gpu_ready, graph):
    """Consecutive ticks overlap (a tick's walk and binning beside the
    previous tick's folds and recenters, on two streams and workspaces);
    every grid stays bit-exact at the checks between runs of four ticks."""
    batch = _run(cb, orc, 9, graph=graph, pipelined=True, check_every=4)
    assert batch.pipelined


def test_batch_pipelined_after_tuning(cb, orc, gpu_ready):
    """Autotuning ticks run one at a time, then the ticks overlap; the
    setting switches (walk and hot-tile mode) keep every grid exact."""
    cands = [(8, 512, 4, 1), (8, 256, 1, 0)]
    batch = _run(cb, orc, 10, graph=True, tuning=cands, pipelined=True, check_every=3)
    assert not batch.tuning


def test_batch_pipelined_rejects_changed_voxels(cb, gpu_ready):
    g = cb.OccupancyGrid((20, 20, 8), 0.1, False, 1.0)
    batch = cb.GridBatch.for_grids([g], 4096, changed_voxels=True)
    with pytest.raises(RuntimeError, match="changed-voxel"):
        batch.set_pipelined(True)
    batch.close()
    g.close()


def test_batch_unsplit_hot_tiles(cb, orc, gpu_ready):
    """Four scratch slabs for 256-record chunks: the hot tiles past them stay
    whole and accumulate in passes (the unsplit hot-tile kernel)."""
    batch = _run(cb, orc, 3, setting=(8, 256, 1, 0), max_slabs=4)
    c = batch.counters()
    assert c["unsplit_hot"] > 0 and c["split_tiles"] > 0, c
