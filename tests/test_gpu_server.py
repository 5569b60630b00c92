"""GPU checks of the map server (cublox::MapServer).

The property under test is "the same result alone and in the mix": every
session's grid, after the server has interleaved its scans and fetches with
other sessions' in batched ticks, is bit-identical to the same requests
issued one by one on a plain OccupancyGrid, and (for the smallest shapes)
to the oracle; each fetch returns the same sample multiset as the alone
run's fetch at the same point of the request sequence. Also covered:
HBM-budget admission with a waiting queue, completion bookkeeping, many
queued scans of one session (one per tick, in order), and every tick
formed from whatever arrived while the previous one ran.
"""
import numpy as np
import pytest

from helpers import assert_occ_bitexact, sample_multiset

pytestmark = pytest.mark.gpu

# half, resolution, centre binning, recenter threshold, max range, beams,
# azimuths, speed (m/scan), scene seed
SESSIONS = [
    ((40, 30, 12), 0.1, False, 1.0, 6.0, 16, 256, 0.5, 21),
    ((37, 21, 9), 0.1, True, 1.0, 12.0, 16, 128, 0.7, 22),
    ((60, 60, 20), 0.1, False, 2.0, 8.0, 32, 256, 1.0, 23),
    ((25, 45, 10), 0.2, False, None, 10.0, 8, 256, 0.5, 24),
    ((50, 50, 16), 0.05, True, 0.5, 5.0, 16, 512, 0.4, 25),
]
N_SCANS = 5
FETCH_R = 1.2


def _server(cb, **kw):
    kw.setdefault("max_points_per_scan", 1 << 14)
    kw.setdefault("max_points_per_tick", 1 << 16)
    kw.setdefault("max_tiles_per_tick", 1 << 16)
    kw.setdefault("max_segments_per_ray", 96)
    kw.setdefault("point_arena_bytes", 64 << 20)
    return cb.MapServer(**kw)


def _requests(cb):
    """Per session: list of ("scan", Scan) / ("fetch", centre) in order.
    Every second scan is followed by a fetch around the sensor."""
    out = []
    for (half, res, ctr, thr, rng, beams, az, speed, seed) in SESSIONS:
        sc = cb.Scene(seed=seed, n_boxes=40, box_extent=10.0, speed=speed)
        reqs = []
        for k in range(N_SCANS):
            s = sc.scan(k, beams, az, -15.0, 15.0, rng)
            reqs.append(("scan", s))
            if k % 2 == 1:
                reqs.append(("fetch", np.asarray(s.T, np.float32).reshape(3, 4)[:, 3]))
        out.append(reqs)
    return out


def _alone(cb, i, reqs):
    """Session i's requests issued one by one on a plain grid."""
    half, res, ctr, thr, rng, *_ = SESSIONS[i]
    g = cb.OccupancyGrid(half, res, ctr, thr)
    g.set_max_raycast_range(rng)
    g.set_log_odds_params(*cb.log_odds())
    fetches = []
    for kind, x in reqs:
        if kind == "scan":
            g.update_sensor_frame((x.xs, x.ys, x.zs), x.T)
            g.recenter(x.base)
        else:
            fetches.append(g.fetch_occupancy_around(x, FETCH_R))
    return g, fetches


def _open_all(cb, srv):
    sids = []
    for (half, res, ctr, thr, rng, *_rest) in SESSIONS:
        sids.append(srv.open_session(half, res, ctr, thr, rng, cb.log_odds()))
    return sids


def _submit_interleaved(srv, sids, reqs, seed):
    """Submits all sessions' requests in a seeded random interleaving (each
    session's own order kept); returns {session index: [fetch request ids]}."""
    rng = np.random.default_rng(seed)
    pos = [0] * len(reqs)
    fetch_ids = {i: [] for i in range(len(reqs))}
    scan_ids = {i: [] for i in range(len(reqs))}
    while True:
        live = [i for i in range(len(reqs)) if pos[i] < len(reqs[i])]
        if not live:
            break
        i = int(rng.choice(live))
        kind, x = reqs[i][pos[i]]
        pos[i] += 1
        if kind == "scan":
            scan_ids[i].append(srv.submit_scan(sids[i], x.xs, x.ys, x.zs, x.T, x.base))
        else:
            fetch_ids[i].append(srv.submit_fetch(sids[i], x, FETCH_R))
    return scan_ids, fetch_ids


@pytest.mark.parametrize("graph,pipelined", [(False, False), (True, False),
                                              (False, True), (True, True)])
def test_server_mix_matches_alone(cb, orc, gpu_ready, graph, pipelined):
    """Pipelined: two ticks in flight, each tick's walk beside the previous
    tick's folds, recenters and fetches."""
    reqs = _requests(cb)
    srv = _server(cb, graph=graph, pipelined=pipelined)
    srv.keep_fetch_results(True)
    sids = _open_all(cb, srv)
    print(f"[server] graph={graph} pipelined={pipelined} sessions={sids}", flush=True)
    srv.pause()  # everything queued first: ticks then batch deterministically
    scan_ids, fetch_ids = _submit_interleaved(srv, sids, reqs, seed=7)
    srv.resume()
    srv.drain()
    comp = srv.completions()
    n_req = sum(len(r) for r in reqs)
    assert len(comp) == n_req and len(set(comp["id"].tolist())) == n_req
    assert np.all(comp["done_ms"] >= comp["submit_ms"])
    st = srv.stats()
    assert st["scans"] == len(SESSIONS) * N_SCANS
    # Every tick takes the next scan of every session (all fit the caps).
    assert st["ticks"] == N_SCANS, st
    for i, sid in enumerate(sids):
        # A session's scans complete in submission order, one per tick.
        mine = comp[(comp["session"] == sid) & (comp["kind"] == 0)]
        mine = mine[np.argsort(mine["id"])]
        assert np.all(np.diff(mine["tick"]) > 0), mine["tick"]
        g, fetches = _alone(cb, i, reqs[i])
        assert srv.origin(sid) == g.geometry.origin_i, i
        assert_occ_bitexact(srv.occupancy(sid), g.occupancy(), f"session {i}")
        for rid, (p_a, s_a) in zip(fetch_ids[i], fetches):
            p_s, s_s = srv.fetch_result(rid)
            assert np.array_equal(sample_multiset(p_s, s_s), sample_multiset(p_a, s_a)), \
                (i, rid)
        g.close()
    srv.close()


@pytest.mark.parametrize("pipelined", [False, True])
def test_server_sessions_match_oracle(cb, orc, gpu_ready, pipelined):
    """Per-session oracle agreement for two of the shapes, served in a mix
    with requests arriving while ticks run."""
    reqs = _requests(cb)
    srv = _server(cb, pipelined=pipelined)
    sids = _open_all(cb, srv)
    _submit_interleaved(srv, sids, reqs, seed=11)
    srv.drain()
    for i in (0, 3):
        half, res, ctr, thr, rng, *_ = SESSIONS[i]
        o = orc.OracleGrid(half, res, ctr, thr)
        o.set_max_raycast_range(rng)
        o.set_log_odds_params(*cb.log_odds())
        for kind, x in reqs[i]:
            if kind == "scan":
                o.update_sensor_frame((x.xs, x.ys, x.zs), x.T)
                o.recenter(x.base)
        assert srv.origin(sids[i]) == o.origin()[0]
        assert_occ_bitexact(srv.occupancy(sids[i]), o.occupancy(), f"session {i} vs oracle")
    srv.close()


def test_server_admission_queue(cb, gpu_ready):
    """A budget for two sessions: the third waits (its requests queue) until
    a close frees room, then runs and matches its alone result."""
    reqs = _requests(cb)
    b = [cb.MapServer.session_bytes(s[0]) for s in SESSIONS[:3]]
    srv = _server(cb, hbm_budget_bytes=b[0] + b[1] + min(b[2], max(b[0], b[1])) // 2)
    sids = _open_all(cb, srv)[:3]  # sessions 3, 4 also wait, behind 2 (FIFO)
    for i in range(3):
        for kind, x in reqs[i]:
            if kind == "scan":
                srv.submit_scan(sids[i], x.xs, x.ys, x.zs, x.T, x.base)
    srv.drain()
    assert srv.admitted(sids[0]) and srv.admitted(sids[1])
    assert not srv.admitted(sids[2])
    assert srv.stats()["waiting"] == len(SESSIONS) - 2
    done = srv.completions()
    assert set(done["session"].tolist()) == {sids[0], sids[1]}
    srv.close_session(sids[0])
    srv.drain()
    assert srv.admitted(sids[2])
    done = srv.completions()
    assert set(done["session"].tolist()) == {sids[2]}
    scans_only = [r for r in reqs[2] if r[0] == "scan"]
    g, _ = _alone(cb, 2, scans_only)
    assert_occ_bitexact(srv.occupancy(sids[2]), g.occupancy(), "admitted late")
    g.close()
    srv.close()


def test_server_queued_scans_of_one_session(cb, gpu_ready):
    """Many scans of one session queued at once run one per tick, in order."""
    reqs = [r for r in _requests(cb)[2] if r[0] == "scan"]
    srv = _server(cb)
    half, res, ctr, thr, rng, *_ = SESSIONS[2]
    sid = srv.open_session(half, res, ctr, thr, rng, cb.log_odds())
    for _, x in reqs:
        srv.submit_scan(sid, x.xs, x.ys, x.zs, x.T, x.base)
    srv.drain()
    comp = srv.completions()
    assert len(comp) == len(reqs)
    assert len(set(comp["tick"].tolist())) == len(reqs)
    g, _ = _alone(cb, 2, reqs)
    assert_occ_bitexact(srv.occupancy(sid), g.occupancy(), "queued")
    g.close()
    srv.close()


def test_server_small_point_arena(cb, gpu_ready):
    """A point arena for about two scans: submissions wait for finished
    ticks to free their ranges (first-fit with coalescing), and the result
    still matches the session run alone."""
    reqs = [r for r in _requests(cb)[4] if r[0] == "scan"]
    n = max(x.n for _, x in reqs)
    srv = _server(cb, max_points_per_scan=n, max_points_per_tick=4 * n,
                  point_arena_bytes=2 * (12 * n + 256) + 256)
    half, res, ctr, thr, rng, *_ = SESSIONS[4]
    sid = srv.open_session(half, res, ctr, thr, rng, cb.log_odds())
    for _, x in reqs:
        srv.submit_scan(sid, x.xs, x.ys, x.zs, x.T, x.base)
    srv.drain()
    assert len(srv.completions()) == len(reqs)
    g, _ = _alone(cb, 4, reqs)
    assert_occ_bitexact(srv.occupancy(sid), g.occupancy(), "small arena")
    g.close()
    srv.close()


def _scans_of(reqs, i):
    return [r for r in reqs[i] if r[0] == "scan"]


def test_server_manual_eviction(cb, gpu_ready):
    """Sessions evicted to host snapshots between scans come back exactly:
    every session's map equals its alone run, evicted or resident."""
    reqs = _requests(cb)
    srv = _server(cb)
    sids = _open_all(cb, srv)
    for i, sid in enumerate(sids):
        for _, x in _scans_of(reqs, i)[:2]:
            srv.submit_scan(sid, x.xs, x.ys, x.zs, x.T, x.base)
    srv.drain()
    srv.evict_session(sids[0])
    srv.evict_session(sids[2])
    srv.drain()
    st = srv.stats()
    assert st["evictions"] == 2 and st["snapshot_bytes"] > 0, st
    assert not srv.admitted(sids[0]) and not srv.admitted(sids[2])
    # An evicted map reads back from its snapshot.
    g, _ = _alone(cb, 2, _scans_of(reqs, 2)[:2])
    assert_occ_bitexact(srv.occupancy(sids[2]), g.occupancy(), "evicted map")
    assert srv.origin(sids[2]) == g.geometry.origin_i
    g.close()
    for i, sid in enumerate(sids):
        for _, x in _scans_of(reqs, i)[2:]:
            srv.submit_scan(sid, x.xs, x.ys, x.zs, x.T, x.base)
    srv.drain()
    st = srv.stats()
    assert st["restores"] == 2 and st["snapshot_bytes"] == 0, st
    for i, sid in enumerate(sids):
        assert srv.admitted(sid)
        g, _ = _alone(cb, i, _scans_of(reqs, i))
        assert srv.origin(sid) == g.geometry.origin_i, i
        assert_occ_bitexact(srv.occupancy(sid), g.occupancy(), f"session {i}")
        g.close()
    srv.close()


def test_server_evicts_idle_sessions_for_waiting_ones(cb, gpu_ready):
    """A budget for two of three sessions with idle eviction on: each
    session's scans run in turn, the idle ones going to host snapshots and
    back, and every map ends as its alone run."""
    reqs = _requests(cb)
    b = [cb.MapServer.session_bytes(s[0]) for s in SESSIONS[:3]]
    srv = _server(cb, hbm_budget_bytes=sorted(b)[1] + sorted(b)[2] + 1, evict_idle=True)
    sids = _open_all(cb, srv)[:3]
    order = [0, 1, 2, 0, 2, 1, 0, 1, 2]  # rounds of one or two scans
    pos = [0, 0, 0]
    for i in order:
        scans = _scans_of(reqs, i)
        for _, x in scans[pos[i]:pos[i] + 2]:
            srv.submit_scan(sids[i], x.xs, x.ys, x.zs, x.T, x.base)
        pos[i] = min(pos[i] + 2, len(scans))
        srv.drain()
    st = srv.stats()
    assert st["evictions"] >= 2 and st["restores"] >= 1, st
    assert st["resident_bytes"] <= sorted(b)[1] + sorted(b)[2] + 1
    for i in range(3):
        g, _ = _alone(cb, i, _scans_of(reqs, i)[:pos[i]])
        assert srv.origin(sids[i]) == g.geometry.origin_i, i
        assert_occ_bitexact(srv.occupancy(sids[i]), g.occupancy(), f"session {i}")
        g.close()
    srv.close()
