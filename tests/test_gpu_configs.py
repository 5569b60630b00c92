"""GPU parity at the batched and served workload shapes.

Batched (64 grids of 401x401x101 @ 0.1 m, a scan per grid per tick of
8,192..131,072 points and 10..40 m range): six grids of that shape tick
through cublox::GridBatch, with and without HIP-graph replay, with each
tick's point count and range drawn per grid; the first tick takes the
largest scan (128 x 1024 points, 40 m) on every grid. After every tick each
grid must equal, bit for bit, the oracle grid updated alone with the same
scans and recentered after each (the per-grid semantics of reference
OccupancyGrid.cpp:181-238 and 141-179).

Served (sessions over 201x201x51, 401x401x101 and 701x701x151 @ 0.1 m,
16,384..131,072-point scans, sphere fetches of 0.5..3 m interleaved): a
MapServer mix holding two 701x701x151 sessions; every session ends bit-
identical to its requests issued alone on a plain grid, every fetch returns
the same samples as the alone run's, and one session per shape equals the
oracle.
"""
import numpy as np
import pytest

from helpers import assert_occ_bitexact, sample_multiset

pytestmark = pytest.mark.gpu

CB_GRIDS = 6
CB_TICKS = 7


def _cb_setup(cb, orc):
    from cublox_amd.workloads import batched_grid
    w = cb.WORKLOADS["batched"]
    lo = cb.log_odds()
    r_max = w.extra["range_range"][1]
    grids, oracles, scenes = [], [], []
    for i in range(CB_GRIDS):
        _, _, _, scene = batched_grid(w, i)
        g = cb.OccupancyGrid(w.half, w.resolution, False, w.threshold)
        o = orc.OracleGrid(w.half, w.resolution, False, w.threshold, threads=16)
        for x in (g, o):
            x.set_max_raycast_range(r_max)  # the segment bound of the batch limits
            x.set_log_odds_params(*lo)
        grids.append(g)
        oracles.append(o)
        scenes.append(scene)
    return w, grids, oracles, scenes


@pytest.mark.parametrize("graph,pipelined", [(False, False), (True, False), (True, True)])
def test_batched_shape_matches_oracle(cb, orc, gpu_ready, graph, pipelined):
    """Pipelined: ticks overlap in pairs (checked after every second tick)."""
    from cublox_amd.workloads import batched_scan_shape
    w, grids, oracles, scenes = _cb_setup(cb, orc)
    batch = cb.GridBatch.for_grids(grids, CB_GRIDS * 131072, max_near_steps=12)
    batch.set_graph_mode(graph)
    batch.set_pipelined(pipelined)
    moved = 0
    live = []
    before = [grids[i].geometry.origin_i for i in range(CB_GRIDS)]
    for k in range(CB_TICKS):
        scans, recenters, keep = [], [], []
        for i in range(CB_GRIDS):
            if k == 0:
                beams, az, rmax = 128, 1024, w.extra["range_range"][1]
            else:
                beams, az, rmax = batched_scan_shape(w, i, k)
            s = scenes[i].scan(k, beams, az, w.el_lo, w.el_hi, rmax)
            grids[i].set_max_raycast_range(rmax)
            oracles[i].set_max_raycast_range(rmax)
            bufs = [cb.DeviceBuffer.from_array(np.ascontiguousarray(c))
                    for c in (s.xs, s.ys, s.zs)]
            keep.append(bufs)
            scans.append((grids[i], bufs[0], bufs[1], bufs[2], s.n, s.T))
            recenters.append((grids[i], s.base))
            oracles[i].update_sensor_frame((s.xs, s.ys, s.zs), s.T)
            oracles[i].recenter(s.base)
        print(f"[cb] tick {k} graph={graph} pipelined={pipelined} "
              f"points={[x[4] for x in scans]}", flush=True)
        if pipelined:
            live.append(keep)
            batch.submit(scans, recenters)
            if k % 2 == 0 and k + 1 < CB_TICKS:
                continue
            batch.wait()
            live.clear()
        else:
            batch.tick(scans, recenters)
        for i in range(CB_GRIDS):
            gi = grids[i].geometry.origin_i
            moved += int(gi != before[i])
            assert gi == oracles[i].origin()[0], (k, i)
            assert_occ_bitexact(grids[i].occupancy(), oracles[i].occupancy(),
                                f"tick {k} grid {i}")
        before = [grids[i].geometry.origin_i for i in range(CB_GRIDS)]
    assert moved >= CB_GRIDS, "every grid must recenter at least once"
    c = batch.counters()
    assert c["overflow"] == 0, c
    batch.close()
    for g in grids:
        g.close()


# Served shapes: (half, scene seed, box extent, max range, speed m/scan)
SERVED = [
    ((100, 100, 25), 41, 10.0, 10.0, 0.5),
    ((200, 200, 50), 42, 40.0, 40.0, 1.0),
    ((350, 350, 75), 43, 60.0, 30.0, 1.0),
    ((350, 350, 75), 44, 60.0, 30.0, 1.0),
    ((200, 200, 50), 45, 40.0, 25.0, 0.7),
]
SERVED_SCANS = 4


def _served_requests(cb):
    """Per session: ("scan", Scan) / ("fetch", (centre, radius)) in order; a
    scan's point count is drawn in [16,384, 131,072], fetch radii in
    [0.5, 3] m."""
    out = []
    for i, (half, seed, extent, rng_m, speed) in enumerate(SERVED):
        sc = cb.Scene(seed=seed, n_boxes=200, box_extent=extent, speed=speed)
        draw = np.random.default_rng(seed)
        reqs = []
        for k in range(SERVED_SCANS):
            beams = int(draw.integers(16, 129))
            az = int(draw.integers(1024 if beams < 32 else 512, 1025))
            s = sc.scan(k, beams, az, -22.5, 22.5, rng_m)
            assert 16384 <= s.n <= 131072
            reqs.append(("scan", s))
            if k % 2 == 1 or i == 2:
                c = np.asarray(s.T, np.float32).reshape(3, 4)[:, 3]
                reqs.append(("fetch", (c, float(np.float32(draw.uniform(0.5, 3.0))))))
        out.append(reqs)
    return out


def _served_alone(cb, i, reqs):
    half, _, _, rng_m, _ = SERVED[i]
    g = cb.OccupancyGrid(half, 0.1, False, 2.0)
    g.set_max_raycast_range(rng_m)
    g.set_log_odds_params(*cb.log_odds())
    fetches = []
    for kind, x in reqs:
        if kind == "scan":
            g.update_sensor_frame((x.xs, x.ys, x.zs), x.T)
            g.recenter(x.base)
        else:
            fetches.append(g.fetch_occupancy_around(x[0], x[1]))
    return g, fetches


def test_served_shapes_mix_matches_alone_and_oracle(cb, orc, gpu_ready):
    """The served configuration: graph replay and pipelined ticks."""
    reqs = _served_requests(cb)
    srv = cb.MapServer(max_points_per_scan=131072, max_points_per_tick=1 << 20,
                       max_segments_per_ray=96, graph=True, pipelined=True)
    srv.keep_fetch_results(True)
    sids = [srv.open_session(half, 0.1, False, 2.0, rng_m, cb.log_odds())
            for (half, _, _, rng_m, _) in SERVED]
    print(f"[served] sessions={sids}", flush=True)
    srv.pause()
    order = np.random.default_rng(3)
    pos = [0] * len(reqs)
    fetch_ids = {i: [] for i in range(len(reqs))}
    while True:
        live = [i for i in range(len(reqs)) if pos[i] < len(reqs[i])]
        if not live:
            break
        i = int(order.choice(live))
        kind, x = reqs[i][pos[i]]
        pos[i] += 1
        if kind == "scan":
            srv.submit_scan(sids[i], x.xs, x.ys, x.zs, x.T, x.base)
        else:
            fetch_ids[i].append(srv.submit_fetch(sids[i], x[0], x[1]))
    srv.resume()
    srv.drain()
    st = srv.stats()
    assert st["scans"] == len(SERVED) * SERVED_SCANS, st
    for i, sid in enumerate(sids):
        print(f"[served] session {i} alone", flush=True)
        g, fetches = _served_alone(cb, i, reqs[i])
        assert srv.origin(sid) == g.geometry.origin_i, i
        assert_occ_bitexact(srv.occupancy(sid), g.occupancy(), f"session {i}")
        assert len(fetch_ids[i]) == len(fetches)
        for rid, (p_a, s_a) in zip(fetch_ids[i], fetches):
            p_s, s_s = srv.fetch_result(rid)
            assert np.array_equal(sample_multiset(p_s, s_s), sample_multiset(p_a, s_a)), \
                (i, rid)
        g.close()
    # One session per shape against the oracle.
    for i in (0, 1, 2):
        half, _, _, rng_m, _ = SERVED[i]
        print(f"[served] session {i} oracle", flush=True)
        o = orc.OracleGrid(half, 0.1, False, 2.0, threads=16)
        o.set_max_raycast_range(rng_m)
        o.set_log_odds_params(*cb.log_odds())
        for kind, x in reqs[i]:
            if kind == "scan":
                o.update_sensor_frame((x.xs, x.ys, x.zs), x.T)
                o.recenter(x.base)
        assert srv.origin(sids[i]) == o.origin()[0]
        assert_occ_bitexact(srv.occupancy(sids[i]), o.occupancy(), f"session {i} vs oracle")
        del o
    srv.close()
