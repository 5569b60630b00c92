#!/usr/bin/env python3
"""Occupancy-update benchmark on one MI355X.

One step = one LiDAR scan through the drop-in OccupancyGrid: fused scan
transform + ray cast + log-odds apply (update), then recenter on the robot
base position — the per-tick pair the reference driver times (reference
src/main.cpp:468-483). Inputs (sensor-frame SoA points) are resident in HBM
before the timed region starts; each call blocks, as in the reference.

Default workload: 701x701x151 voxels at 0.1 m, 65,536-point scans
(64 beams x 1024 azimuths, 30 m), 400-box scene, 1 m/scan, 2 m recenter
threshold. Prints ONE JSON line on stdout; details go to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
README_3090_MS = 6.167  # update 1.344 + recenter 4.822 ms (reference README.md:35)
METRIC = "ms/scan update+recenter @73.5M voxels; batched scans/s + p99 latency vs plain"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Sources of the fused update kernels: a recorded HBM traffic figure holds
# only for the kernels it was measured on.
UPDATE_SOURCES = ("cublox_amd/csrc/integrate.hip", "cublox_amd/csrc/raycast_common.hpp",
                  "cublox_amd/csrc/hip_common.hpp")


def code_digest(texts) -> str:
    """Digest of C++/HIP sources with comments and whitespace removed, so
    that only a change to the code itself makes a recorded figure stale."""
    import hashlib
    import re
    h = hashlib.sha256()
    for t in texts:
        t = re.sub(r"/\*.*?\*/", " ", t, flags=re.S)
        t = re.sub(r"//[^\n]*", " ", t)
        h.update(" ".join(t.split()).encode())
    return h.hexdigest()[:16]


def update_source_digest() -> str:
    texts = []
    for rel in UPDATE_SOURCES:
        with open(os.path.join(ROOT, rel), encoding="utf-8") as fh:
            texts.append(fh.read())
    return code_digest(texts)


def host_cpus() -> dict:
    """CPU counts of the host this runs on."""
    return dict(nproc=os.cpu_count(), affinity=len(os.sched_getaffinity(0)),
                omp_num_threads=os.environ.get("OMP_NUM_THREADS"))


def pct(xs, q):
    xs = sorted(xs)
    if not xs:
        return 0.0
    k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[k]


def make_scans(w, count, scene):
    scans = []
    for k in range(count):
        scans.append(scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range))
    return scans


def upload(cb, scans):
    """All scans' SoA columns into three device buffers (HBM-resident)."""
    n = scans[0].n
    total = n * len(scans)
    cols = [np.empty(total, np.float32) for _ in range(3)]
    for i, s in enumerate(scans):
        cols[0][i * n:(i + 1) * n] = s.xs
        cols[1][i * n:(i + 1) * n] = s.ys
        cols[2][i * n:(i + 1) * n] = s.zs
    bufs = [cb.DeviceBuffer.from_array(c) for c in cols]
    return bufs, n


def new_grid(cb, w, path=0):
    g = cb.OccupancyGrid(w.half, w.resolution, w.origin_at_center, w.threshold)
    g.set_max_raycast_range(w.max_range)
    g.set_log_odds_params(*cb.log_odds())
    g.set_update_path(path)
    return g


def run_gpu(cb, w, scans, bufs, n, warmup, steps, path=0, timing=True, native=True):
    """The timed per-scan loop: per scan the blocking update on the scan's
    HBM-resident columns, then the recenter. native: the loop runs in C++
    (cublox_seq_replay, as a C++ driver calls the drop-in API); else in
    Python, one binding call per call. With timing, HIP events around every
    kernel launch feed kernel_times(); the headline pass runs without them,
    so the wall clock holds only the drop-in calls."""
    from cublox_amd.grid import Replay
    g = new_grid(cb, w, path)
    g.enable_kernel_timing(timing)

    def spec(i):
        off = 4 * n * i
        return (g, bufs[0].ptr + off, bufs[1].ptr + off, bufs[2].ptr + off, n,
                scans[i].T, scans[i].base)

    if native:
        warm = Replay([spec(i) for i in range(warmup)])
        timed = Replay([spec(i) for i in range(warmup, warmup + steps)])
        warm.run()
        g.reset_kernel_times()
        cb.synchronize()
        t0 = time.perf_counter()
        per, moved = timed.run()
        cb.synchronize()
        t1 = time.perf_counter()
    else:
        moved = []

        def step(i):
            off = 4 * n * i
            s = scans[i]
            g.update_sensor_frame_device(bufs[0].ptr + off, bufs[1].ptr + off,
                                         bufs[2].ptr + off, n, s.T)
            moved.append(bool(g.recenter(s.base)))

        for i in range(warmup):
            step(i)
        g.reset_kernel_times()
        cb.synchronize()
        per = []
        moved.clear()
        t0 = time.perf_counter()
        for i in range(warmup, warmup + steps):
            a = time.perf_counter()
            step(i)
            per.append((time.perf_counter() - a) * 1e3)
        cb.synchronize()
        t1 = time.perf_counter()
    kt = g.kernel_times()
    run_gpu.moved = moved
    return g, per, (t1 - t0) * 1e3, kt


def recenter_cost(per, moved):
    """Critical-path time a recenter adds: the mean of a recentering step
    plus the step after it, less the same pair when no recenter happens
    (the clear runs beside the next update's ray walk)."""
    pairs_r = [per[k] + per[k + 1] for k in range(len(per) - 1) if moved[k]]
    pairs_n = [per[k] + per[k + 1] for k in range(len(per) - 1) if not moved[k]]
    if not pairs_r or not pairs_n:
        return None
    return dict(added_ms=round(statistics.mean(pairs_r) - statistics.mean(pairs_n), 5),
                events=len(pairs_r),
                step_ms_recenter=round(statistics.mean(per[k] for k in range(len(per))
                                                        if moved[k]), 5),
                step_ms_no_recenter=round(statistics.mean(per[k] for k in range(len(per))
                                                           if not moved[k]), 5))


def census(cb, w, scans, bufs, n, warmup, steps):
    """Replay of the same scans with a counter census after each ray cast
    (untimed): voxel visits and unique voxels per scan."""
    g = new_grid(cb, w)
    out = []
    for i in range(warmup + steps):
        off = 4 * n * i
        s = scans[i]
        g.cast_rays_sensor_frame_device(bufs[0].ptr + off, bufs[1].ptr + off,
                                        bufs[2].ptr + off, n, s.T)
        c = g.counter_census()
        g.apply_counters()
        g.recenter(s.base)
        if i >= warmup:
            out.append(c)
    g.close()
    return out


def run_pcie(cb, w, scans, warmup, steps):
    """Same steps with host-resident points (H2D inside update), in the
    native loop."""
    from cublox_amd.grid import Replay
    g = new_grid(cb, w)

    def spec(i):
        s = scans[i]
        return (g, s.xs, s.ys, s.zs, s.n, s.T, s.base)

    Replay([spec(i) for i in range(warmup)], on_host=True).run()
    per, _ = Replay([spec(i) for i in range(warmup, warmup + steps)], on_host=True).run()
    g.close()
    return per


# Autotuner candidates of the batched workload: (near steps, hot-tile chunk,
# lanes per ray = 64 / lanes rays per wavefront, hot-tile mode 0 chunk slabs
# / 1 shared slabs with global atomics).
TUNE_CANDIDATES = [(8, 2048, 1, 0), (8, 4096, 1, 0), (8, 8192, 1, 0), (8, 16384, 1, 0),
                   (8, 8192, 4, 0), (8, 2048, 1, 1), (6, 8192, 1, 0), (10, 8192, 1, 0)]


def run_batched(cb, tune_ticks_each, ticks):
    """The batched workload: 64 grids of 401x401x101, each with its own
    scene, trajectory and max range; every tick each grid gets a scan whose
    beams x azimuths (8,192..131,072 points) are drawn for that tick. The
    same scans run through the plain form (each grid's blocking update +
    recenter in turn) and through one GridBatch tick per step (graph
    replay, after online tuning)."""
    from cublox_amd.workloads import batched_grid, batched_scan_shape
    w = cb.WORKLOADS["batched"]
    G = w.extra["grids"]
    tune_ticks = len(TUNE_CANDIDATES) * (tune_ticks_each + 1)
    rc_ticks = 8  # after the timed ticks, without graph: recenter timed alone
    total = tune_ticks + 2 * ticks + rc_ticks
    timed = range(tune_ticks, tune_ticks + ticks)  # one tick at a time
    piped = range(tune_ticks + ticks, tune_ticks + 2 * ticks)  # overlapped ticks
    draws = [batched_grid(w, g) for g in range(G)]
    t = time.perf_counter()
    shapes = [[batched_scan_shape(w, g, k)[:2] for k in range(total)] for g in range(G)]
    scans = [[draws[g][3].scan(k, *shapes[g][k], w.el_lo, w.el_hi, draws[g][2])
              for k in range(total)] for g in range(G)]
    bufs = [[[cb.DeviceBuffer.from_array(np.ascontiguousarray(c))
              for c in (s.xs, s.ys, s.zs)] for s in gs] for gs in scans]
    pts = [sum(scans[g][k].n for g in range(G)) for k in range(total)]
    log(f"batched: {G} grids, {statistics.mean(pts):.0f} points/tick, generated + "
        f"uploaded {G * total} scans in {time.perf_counter() - t:.1f} s")

    def grids():
        out = []
        for g in range(G):
            x = cb.OccupancyGrid(w.half, w.resolution, w.origin_at_center, w.threshold)
            x.set_max_raycast_range(draws[g][2])
            x.set_log_odds_params(*cb.log_odds())
            out.append(x)
        return out

    # Plain form. Each grid's workspace is sized for its largest scan first
    # (allocation is set-up, not per-scan work).
    plain = grids()
    for g in range(G):
        kmax = max(range(total), key=lambda k: scans[g][k].n)
        b = bufs[g][kmax]
        plain[g].update_sensor_frame_device(b[0].ptr, b[1].ptr, b[2].ptr, scans[g][kmax].n,
                                            scans[g][kmax].T)
        plain[g].reset()
    # Each tick's 64 blocking update + recenter calls run in the native loop
    # (cublox_seq_replay), prepared before the clock starts.
    from cublox_amd.grid import Replay
    per_scan, tick_ms = [], []
    for k in range(total):
        rp = Replay([(plain[g], *(b.ptr for b in bufs[g][k]), scans[g][k].n, scans[g][k].T,
                      scans[g][k].base) for g in range(G)])
        a = time.perf_counter()
        step_ms, _ = rp.run()
        if k in timed or k in piped:
            tick_ms.append((time.perf_counter() - a) * 1e3)
            per_scan.extend(step_ms)
    for x in plain:
        x.close()
    plain_sps = G * len(tick_ms) / (sum(tick_ms) * 1e-3)

    # Batched form.
    bg = grids()
    batch = cb.GridBatch.for_grids(bg, max(pts), max_near_steps=12, min_chunk=512)
    batch.set_graph_mode(True)
    batch.set_tuning(TUNE_CANDIDATES, tune_ticks_each)
    btick, bdev, zeroed, rc_ms, rc_vox = [], [], [], [], []
    pipe_ms = 0.0
    # Every tick's arguments are packed before the clock starts, as the
    # plain leg's calls are.
    prepared = [batch.prepare([(bg[g], *bufs[g][k], scans[g][k].n, scans[g][k].T)
                               for g in range(G)],
                              [(bg[g], scans[g][k].base) for g in range(G)])
                for k in range(total)]
    for k in range(total):
        if k in piped:
            # Overlapped ticks: each submit waits only for the tick two back.
            if k == piped[0]:
                batch.set_pipelined(True)
                a0 = time.perf_counter()
            batch.submit_prepared(prepared[k])
            if k == piped[-1]:
                batch.wait()
                pipe_ms = (time.perf_counter() - a0) * 1e3
                batch.set_pipelined(False)
                batch.set_graph_mode(False)
            continue
        a = time.perf_counter()
        batch.tick_prepared(prepared[k])
        if k in timed:
            btick.append((time.perf_counter() - a) * 1e3)
            bdev.append(batch.last[0])
            zeroed.append(batch.last[2])
        elif k >= tune_ticks + ticks:
            rc_ms.append(batch.last[1])
            rc_vox.append(batch.last[2])
            records = batch.counters()["records"]
    batch_sps = G * len(btick) / (sum(btick) * 1e-3)
    pipe_sps = G * len(piped) / (pipe_ms * 1e-3)
    # Batched recenter: 4 B per zeroed voxel over the recenter sequence's
    # device time (descriptor copy + one zeroing launch for all grids).
    # Only ticks whose recenters moved some window count.
    rc_ev = [(ms, v) for ms, v in zip(rc_ms, rc_vox) if v > 0]
    rc_ms = [ms for ms, _ in rc_ev] or [0.0]
    rc_vox = [v for _, v in rc_ev] or [0]
    rc_gbs = 4.0 * sum(rc_vox) / (sum(rc_ms) * 1e-3) / 1e9 if sum(rc_ms) > 0 else 0.0
    tune = batch.tune_results(len(TUNE_CANDIDATES))
    out = dict(
        grids=G, grid=f"{w.map_size[0]}x{w.map_size[1]}x{w.map_size[2]} @ {w.resolution} m",
        points_per_tick=round(statistics.mean(pts)), ticks=len(btick),
        tile_records_per_tick=records,
        plain=dict(scans_per_s=round(plain_sps, 1), tick_ms=round(statistics.mean(tick_ms), 3),
                   scan_p50_ms=round(pct(per_scan, 50), 4),
                   scan_p95_ms=round(pct(per_scan, 95), 4),
                   scan_p99_ms=round(pct(per_scan, 99), 4)),
        batched=dict(scans_per_s=round(batch_sps, 1),
                     tick_p50_ms=round(pct(btick, 50), 4),
                     tick_p95_ms=round(pct(btick, 95), 4),
                     tick_p99_ms=round(pct(btick, 99), 4),
                     device_tick_ms=round(statistics.mean(bdev), 4),
                     graph=True,
                     setting=dict(zip(("near_steps", "chunk", "lanes_per_ray", "hot_mode"),
                                      batch.setting)),
                     tuning_ms={"/".join(map(str, c)): round(m, 4)
                                for c, m in zip(TUNE_CANDIDATES, tune)},
                     tuning_keys="near_steps/chunk/lanes_per_ray/hot_mode"),
        pipelined=dict(scans_per_s=round(pipe_sps, 1), ticks=len(piped),
                       ms_per_tick=round(pipe_ms / len(piped), 4), graph=True,
                       form="consecutive ticks overlapped: each tick's walk, binning, "
                            "tile scan and placing on one of two streams beside the "
                            "previous tick's folds and recenters"),
        speedup_scans_per_s=round(max(batch_sps, pipe_sps) / plain_sps, 2),
        speedup_one_tick_at_a_time=round(batch_sps / plain_sps, 2),
        recenter=dict(zeroed_voxels_per_tick=round(statistics.mean(zeroed)),
                      batched_recenter_ms=round(statistics.mean(rc_ms), 4),
                      batched_recenter_zeroed_voxels=round(statistics.mean(rc_vox)),
                      batched_recenter_gbs=round(rc_gbs, 1),
                      batched_recenter_frac_of_hbm_peak=round(rc_gbs / HBM_PEAK_GBS, 4)),
    )
    batch.close()
    for x in bg:
        x.close()
    return out


SERVED_SHAPES = ("small", "batched", "readme")  # 201x201x51, 401x401x101, 701x701x151


def _lat(c):
    d = c["done_ms"] - c["submit_ms"]
    if len(d) == 0:
        return None
    return dict(n=int(len(d)), p50_ms=round(pct(d, 50), 3), p95_ms=round(pct(d, 95), 3),
                p99_ms=round(pct(d, 99), 3))


def run_served(cb, n_sessions=256, seconds=3.0, loads=(0.5, 0.8, 1.0), depth=4,
               tick_points=1 << 20, chunk=0):
    """The map server workload: n_sessions sessions cycling over three grid
    shapes, HBM-budget admission (200 GB), continuous batching. Saturation
    is the drain rate of `depth` queued scans per session; then open-loop
    Poisson scan arrivals at each load fraction of it (fetches at 20 % of the
    scan rate, r uniform in 0.5-3 m) for `seconds`, with latency = completion - submit
    (host copy, H2D, batched update + recenter, fetch). The same sessions
    served one scan per tick give the unbatched saturation, and the same
    scans through one drop-in grid per session (blocking update + recenter
    per scan) the plain rate."""
    from cublox_amd.workloads import batched_grid
    W = cb.WORKLOADS
    t = time.perf_counter()
    specs, templates = [], []
    for k in range(n_sessions):
        w = W[SERVED_SHAPES[k % len(SERVED_SHAPES)]]
        if w.name == "batched":
            beams, az, rmax, scene = batched_grid(w, k)
        else:
            beams, az, rmax = w.beams, w.azimuths, w.max_range
            scene = cb.Scene(seed=w.seed(k), n_boxes=w.n_boxes, box_extent=w.box_extent,
                             speed=w.speed, yaw_rate_deg=w.yaw_rate_deg)
        s = scene.scan(0, beams, az, w.el_lo, w.el_hi, rmax)
        specs.append((w.half, w.resolution, w.origin_at_center, w.threshold, rmax))
        templates.append((s.xs, s.ys, s.zs, s.T, s.base))
    pts = [t_[0].size for t_ in templates]
    log(f"served: {n_sessions} sessions, {statistics.mean(pts):.0f} points/scan mean, "
        f"templates in {time.perf_counter() - t:.1f} s")

    def serve(max_grids, piped=False):
        srv = cb.MapServer(hbm_budget_bytes=200 << 30, max_sessions=n_sessions,
                           max_points_per_scan=131072, max_points_per_tick=tick_points,
                           max_grids_per_tick=max_grids, max_tiles_per_tick=1 << 21,
                           max_segments_per_ray=72, graph=True, pipelined=piped,
                           chunk=chunk)
        sids = [srv.open_session(h, r, c, th, mr, cb.log_odds()) for (h, r, c, th, mr) in specs]
        return srv, sids

    def queued(srv, sids, k0):
        """depth scans per session queued while paused, then drained."""
        srv.pause()
        for j in range(depth):
            for sid, (xs, ys, zs, T, base) in zip(sids, templates):
                Tj = np.array(T, np.float32)
                Tj[0, 3] += 0.5 * (k0 + j)
                b = np.array(base, np.float32)
                b[0] += 0.5 * (k0 + j)
                srv.submit_scan(sid, xs, ys, zs, Tj, b)
        srv.resume()
        srv.drain()
        c = srv.completions()
        return len(c) / ((c["done_ms"].max() - c["start_ms"].min()) * 1e-3)

    def measure(piped):
        """Saturation and the open-loop load levels of one server mode."""
        srv, sids = serve(256, piped=piped)
        t = time.perf_counter()
        queued(srv, sids, 0)  # warm-up: first origins, graph capture
        st0 = srv.stats()
        sat = queued(srv, sids, depth)
        st1 = srv.stats()
        log(f"served (pipelined={piped}): saturation {sat:.0f} scans/s over "
            f"{st1['ticks'] - st0['ticks']} ticks ({time.perf_counter() - t:.1f} s), "
            f"resident {st1['resident_bytes'] / 2**30:.1f} GiB")
        levels = []
        for f in loads:
            st_a = srv.stats()
            rate = f * sat
            n_sub, lag = srv.run_load(sids, templates, scan_rate_hz=rate, duration_s=seconds,
                                      fetch_ratio=0.2, fetch_radius=0.5, fetch_radius_max=3.0,
                                      speed=0.5, seed=17, clients=8)
            srv.drain()
            c = srv.completions()
            st_b = srv.stats()
            nt = max(st_b["ticks"] - st_a["ticks"], 1)
            scans = c[c["kind"] == 0]
            span = (c["done_ms"].max() - c["submit_ms"].min()) * 1e-3
            levels.append(dict(load=f, offered_scans_per_s=round(rate, 1),
                               achieved_scans_per_s=round(len(scans) / span, 1),
                               requests=int(n_sub), client_max_lag_ms=round(lag, 3),
                               ticks=nt, scans_per_tick=round(len(scans) / nt, 1),
                               device_ms_per_tick=round(
                                   (st_b["device_ms"] - st_a["device_ms"]) / nt, 3),
                               host_ms_between_ticks=round(
                                   (st_b["launch_ms"] - st_a["launch_ms"]) / nt, 3),
                               scan_latency=_lat(scans),
                               fetch_latency=_lat(c[c["kind"] == 1])))
            log(f"served (pipelined={piped}): load {f}: {levels[-1]}")
        st = srv.stats()
        srv.close()
        return sat, levels, st

    # One tick at a time, then pipelined ticks (two in flight while half a
    # tick of points is queued): throughput against tail latency.
    sat, levels, st = measure(False)
    pipe_sat, pipe_levels, _ = measure(True)
    # Unbatched: the same sessions, one scan per tick.
    srv1, sids1 = serve(1)
    queued(srv1, sids1, 0)
    plain_sat = queued(srv1, sids1, depth)
    srv1.close()
    log(f"served: unbatched saturation {plain_sat:.0f} scans/s")
    # Plain form: one drop-in grid per session, the same scans as blocking
    # update (host columns, so the H2D copy is included as in the server)
    # + recenter calls, one after another.
    plain = []
    for (h, r, c, th, mr) in specs:
        x = cb.OccupancyGrid(h, r, c, th)
        x.set_max_raycast_range(mr)
        x.set_log_odds_params(*cb.log_odds())
        plain.append(x)

    # The rounds run in the native loop (cublox_seq_replay), prepared
    # before the clock starts.
    from cublox_amd.grid import Replay

    def plain_round(j):
        steps = []
        for x, (xs, ys, zs, T, base) in zip(plain, templates):
            Tj = np.array(T, np.float32)
            Tj[0, 3] += 0.5 * j
            b = np.array(base, np.float32)
            b[0] += 0.5 * j
            steps.append((x, xs, ys, zs, xs.size, Tj, b))
        return Replay(steps, on_host=True)

    plain_round(0).run()  # first origins, workspaces
    rounds = [plain_round(j) for j in range(1, depth + 1)]
    t = time.perf_counter()
    for rp in rounds:
        rp.run()
    plain_sps = n_sessions * depth / (time.perf_counter() - t)
    for x in plain:
        x.close()
    log(f"served: plain per-scan calls {plain_sps:.0f} scans/s")
    return dict(sessions=n_sessions,
                shapes=[f"{W[s].map_size[0]}x{W[s].map_size[1]}x{W[s].map_size[2]}"
                        for s in SERVED_SHAPES],
                points_per_scan_mean=round(statistics.mean(pts)),
                hbm_budget_gib=200, resident_gib=round(st["resident_bytes"] / 2**30, 2),
                tick_points=tick_points, chunk=chunk or 4096,
                saturation_scans_per_s=round(sat, 1),
                plain_scans_per_s=round(plain_sps, 1),
                speedup_vs_plain=round(sat / plain_sps, 2),
                unbatched_saturation_scans_per_s=round(plain_sat, 1),
                batching_speedup=round(sat / plain_sat, 2), levels=levels,
                pipelined=dict(saturation_scans_per_s=round(pipe_sat, 1),
                               speedup_vs_plain=round(pipe_sat / plain_sps, 2),
                               levels=pipe_levels,
                               form="two ticks in flight while half a tick of points "
                                    "is queued, else one"),
                seconds_per_level=seconds)


def fill_bandwidth(cb, nbytes=256 << 20, reps=20):
    """Write bandwidth of a plain zero fill (cublox_fill_zero_async, 16 B
    stores) over nbytes, back to back on the default stream: the practical
    ceiling for the store-only recenter."""
    from cublox_amd import _ffi
    L = _ffi.hip()
    buf = cb.DeviceBuffer(nbytes)
    buf.zero()
    a, b = cb.Event(), cb.Event()
    a.record()
    for _ in range(reps):
        _ffi.check(L.cublox_fill_zero_async(buf.ptr, nbytes // 4, None))
    b.record()
    b.synchronize()
    ms = a.elapsed_ms(b) / reps
    # A device-to-device copy of the same size: a known read + write for
    # scaling hardware byte counters (FETCH_SIZE / WRITE_SIZE).
    dst = cb.DeviceBuffer(nbytes)
    a.record()
    _ffi.check(L.cublox_memcpy_d2d_async(dst.ptr, buf.ptr, nbytes, None))
    b.record()
    b.synchronize()
    copy_ms = a.elapsed_ms(b)
    dst.free()
    buf.free()
    return dict(bytes=nbytes, ms=round(ms, 5), gbs=round(nbytes / (ms * 1e-3) / 1e9, 1),
                copy_ms=round(copy_ms, 5),
                copy_gbs=round(2 * nbytes / (copy_ms * 1e-3) / 1e9, 1))


def run_fetch(cb, w, g, scene, k, reps=5):
    if not w.fetch_queries:
        return None
    qs = scene.queries(k, w.fetch_queries, w.fetch_spread)
    r_vox = int(np.ceil(w.fetch_radius / w.resolution))
    # Per-query output segments: the disk volume, at most 16,384 samples
    # (an OCCUPIED query returns surface voxels only; counts stay exact).
    cap = min((2 * r_vox + 1) ** 3, 16384)
    times = []
    counts = None
    k0 = g.kernel_times()
    for _ in range(reps):
        a = time.perf_counter()
        counts, _, _ = g.fetch_occupancy_batch(qs, w.fetch_radius, cb.OCCUPIED, cap)
        times.append((time.perf_counter() - a) * 1e3)
    k1 = g.kernel_times()
    launches = max(k1["fetch_launches"] - k0["fetch_launches"], 1)
    kernel_ms = (k1["fetch_ms"] - k0["fetch_ms"]) / launches
    tested = (k1["fetch_tested_voxels"] - k0["fetch_tested_voxels"]) / launches
    samples = (k1["fetch_samples"] - k0["fetch_samples"]) / launches
    # Algorithmic bytes of one batched launch: 4 B per voxel tested, 13 B
    # per sample written (position + state), 4 B per query count.
    fbytes = 4.0 * tested + 13.0 * samples + 4.0 * len(qs)
    gbs = fbytes / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else 0.0
    roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
            "bytes_per_launch": int(fbytes), "avg_launch_ms": round(kernel_ms, 5),
            "voxels_tested": int(tested), "samples": int(samples),
            "kernel": "fetch_batch_kernel"}
    return dict(queries=int(len(qs)), radius_m=w.fetch_radius, filter="occupied",
                ms_per_batch=statistics.median(times),
                samples=int(counts.sum()), per_query_cap=int(cap),
                queries_at_cap=int((counts >= cap).sum()), roofline=roof)


def _records(scans, n, step=16):
    """All scans as PointCloud2 payloads (x, y, z, intensity; 16 B/point)."""
    rec = np.zeros((len(scans) * n, step // 4), np.float32)
    for i, s in enumerate(scans):
        rec[i * n:(i + 1) * n, 0] = s.xs
        rec[i * n:(i + 1) * n, 1] = s.ys
        rec[i * n:(i + 1) * n, 2] = s.zs
        rec[i * n:(i + 1) * n, 3] = 1.0
    return rec.reshape(-1)


def _depth_frames(count, W=640, H=480, seed=5):
    rng = np.random.default_rng(seed)
    u = np.arange(W, dtype=np.float32)[None, :]
    v = np.arange(H, dtype=np.float32)[:, None]
    out = []
    for k in range(count):
        d = 1.5 + 5.5 * (v / H) + 0.8 * np.sin(u / 37.0 + k) + rng.normal(0, 0.02, (H, W))
        img = np.round(d * 1000).astype(np.uint16)
        img[rng.random((H, W)) < 0.1] = 0
        out.append(img)
    return out


def run_driver_steps(cb, w, scans, bufs, n, warmup, steps, g, base_update_ms):
    """The driver-side rows on the device: ingest of PointCloud2 payloads
    and depth images, the changed-voxel stream's cost, the viz cloud and
    map snapshots (all on the measured workload's grid shape)."""
    out = {}
    total = warmup + steps

    # Changed-voxel stream: the same update with the stream on.
    gd = new_grid(cb, w)
    gd.set_changed_capacity(1 << 24)
    gd.enable_kernel_timing(True)
    changed = []
    for i in range(total):
        off = 4 * n * i
        s = scans[i]
        if i == warmup:
            gd.reset_kernel_times()
        gd.update_sensor_frame_device(bufs[0].ptr + off, bufs[1].ptr + off,
                                      bufs[2].ptr + off, n, s.T)
        if i >= warmup:
            changed.append(gd.changed_voxels()[2])
        gd.recenter(s.base)
    kt = gd.kernel_times()
    upd = kt["update_ms"] / max(kt["update_launches"], 1)
    out["changed_voxels"] = dict(update_ms=upd, plain_update_ms=base_update_ms,
                                 overhead=upd / base_update_ms - 1.0,
                                 changed_per_scan=statistics.mean(changed))
    gd.close()

    # Ingest: PointCloud2 payloads read in place (range drop, transform).
    rec = cb.DeviceBuffer.from_array(_records(scans, n))
    gi = new_grid(cb, w)
    gi.enable_kernel_timing(True)
    per = []
    for i in range(total):
        s = scans[i]
        if i == warmup:
            gi.reset_kernel_times()
            cb.synchronize()
        a = time.perf_counter()
        gi.update_point_cloud2(rec.ptr + 16 * n * i, n, 1, 16, (0, 4, 8), s.T, device=True)
        gi.recenter(s.base)
        if i >= warmup:
            per.append((time.perf_counter() - a) * 1e3)
    kt = gi.kernel_times()
    r2 = np.float32(w.max_range) ** 2
    kept = [int(np.count_nonzero((s.xs * s.xs + s.ys * s.ys) + s.zs * s.zs <= r2))
            for s in scans[warmup:total]]
    out["ingest_point_cloud2"] = dict(
        ms_per_scan=statistics.mean(per), p99_ms=pct(per, 99),
        update_ms=kt["update_ms"] / max(kt["update_launches"], 1),
        point_step=16, kept_points_per_scan=statistics.mean(kept),
        note="misses beyond the range are dropped, as the driver does")
    gi.close()
    del rec

    # The same payloads from host memory (as a driver receives them): each
    # scan's records are copied into the grid's pinned stage, which the
    # ingest kernels read over the bus.
    hrec = _records(scans, n)
    gh = new_grid(cb, w)
    per = []
    for i in range(total):
        s = scans[i]
        payload = hrec[4 * n * i:4 * n * (i + 1)]  # n records of 4 floats
        a = time.perf_counter()
        gh.update_point_cloud2(payload, n, 1, 16, (0, 4, 8), s.T)
        gh.recenter(s.base)
        if i >= warmup:
            per.append((time.perf_counter() - a) * 1e3)
    out["ingest_point_cloud2_host"] = dict(ms_per_scan=statistics.mean(per),
                                           p99_ms=pct(per, 99), point_step=16)
    gh.close()
    del hrec

    # Depth images (640x480, 16-bit mm), device-resident.
    frames = _depth_frames(min(total, 30))
    dbuf = cb.DeviceBuffer.from_array(np.stack(frames).reshape(-1))
    gz = new_grid(cb, w)
    K = (525.0, 525.0, 319.5, 239.5)
    per = []
    for k in range(len(frames)):
        T = np.array([[0, 0, 1, 0.4 * k], [-1, 0, 0, 0.0], [0, -1, 0, 1.3]], np.float32)
        a = time.perf_counter()
        gz.update_depth_image(dbuf.ptr + k * 640 * 480 * 2, K, T, width=640, height=480,
                              step=1280, encoding=0, device=True)
        if k >= 3:
            per.append((time.perf_counter() - a) * 1e3)
    out["ingest_depth_640x480"] = dict(ms_per_frame=statistics.mean(per),
                                      p99_ms=pct(per, 99), frames=len(per))
    gz.close()
    del dbuf

    # Viz cloud on the measured grid (the driver's 25 m, every point).
    base = np.asarray(scans[total - 1].base, np.float32)
    times = []
    for _ in range(5):
        a = time.perf_counter()
        recs, fetched = g.occupancy_viz_cloud(base, 25.0, 1, 1_000_000)
        times.append((time.perf_counter() - a) * 1e3)
    out["viz_cloud"] = dict(radius_m=25.0, ms=statistics.median(times),
                            fetched=fetched, points=int(len(recs)))

    # Snapshot save / restore of the measured grid.
    times_s, times_l = [], []
    blob = None
    g2 = new_grid(cb, w)
    for _ in range(3):
        a = time.perf_counter()
        blob = g.save_snapshot()
        times_s.append((time.perf_counter() - a) * 1e3)
        a = time.perf_counter()
        g2.load_snapshot(blob)
        times_l.append((time.perf_counter() - a) * 1e3)
    same = bool(np.array_equal(g2.occupancy(), g.occupancy()))
    g2.close()
    enc = int(np.frombuffer(blob[12:16], "<u4")[0])
    out["snapshot"] = dict(bytes=len(blob), dense_bytes=4 * g.voxel_num,
                           encoding="sparse" if enc == 1 else "dense",
                           save_ms=statistics.median(times_s),
                           load_ms=statistics.median(times_l), round_trip_equal=same)
    return out


def cpu_baseline(cb, w, scans, budget_s, threads):
    """The oracle (CPU restatement, OpenMP) on the same scans."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    def timed(nthreads, max_scans, budget):
        o = oracle.OracleGrid(w.half, w.resolution, w.origin_at_center, w.threshold,
                              threads=nthreads)
        o.set_max_raycast_range(w.max_range)
        o.set_log_odds_params(*cb.log_odds())
        per = []
        start = time.perf_counter()
        for s in scans[:max_scans]:
            a = time.perf_counter()
            o.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
            o.recenter(s.base)
            per.append((time.perf_counter() - a) * 1e3)
            if time.perf_counter() - start > budget:
                break
        return per

    per = timed(threads, len(scans), budget_s)
    per1 = timed(1, 3, budget_s / 3)
    return dict(value=statistics.median(per), unit="ms/scan", cores=threads,
                kind="port",
                sample=f"first {len(per)} scans of the same sequence, update+recenter, "
                       f"OpenMP {threads} threads (median)",
                one_thread_ms=statistics.median(per1), one_thread_scans=len(per1),
                driver_steps=cpu_driver_steps(cb, w, scans, oracle))


def cpu_driver_steps(cb, w, scans, oracle, reps=5):
    """The driver's own CPU steps that the device rows replace, on one host
    thread as the driver runs them (the oracle's restatement of main.cpp):
    the PointCloud2 build (decode, filter, transform), the depth
    back-projection, and the viz cloud (OCCUPIED fetch + build loop) on
    the oracle grid after the first scans."""
    out = {}
    n = scans[0].n
    rec = _records(scans[:reps], n)
    r2 = np.float32(w.max_range) ** 2
    times = []
    for k in range(reps):
        buf = rec[k * n * 4:(k + 1) * n * 4].view(np.uint8)
        a = time.perf_counter()
        oracle.build_cloud(buf, n, 16, (0, 4, 8), scans[k].T, r2)
        times.append((time.perf_counter() - a) * 1e3)
    out["point_cloud2_build_ms"] = statistics.median(times)
    frames = _depth_frames(reps)
    K = (525.0, 525.0, 319.5, 239.5)
    times = []
    for img in frames:
        a = time.perf_counter()
        oracle.depth_to_cloud(img.view(np.uint8).reshape(-1), 640, 480, 1280, 0, K)
        times.append((time.perf_counter() - a) * 1e3)
    out["depth_back_projection_ms"] = statistics.median(times)
    o = oracle.OracleGrid(w.half, w.resolution, w.origin_at_center, w.threshold, threads=1)
    o.set_max_raycast_range(w.max_range)
    o.set_log_odds_params(*cb.log_odds())
    for s in scans[:reps]:
        o.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        o.recenter(s.base)
    base = np.asarray(scans[reps - 1].base, np.float32)
    a = time.perf_counter()
    pos, _st, fetched = o.fetch(base, 25.0, 2, 1_000_000, cap=1_000_000)
    b = time.perf_counter()
    recs = oracle.viz_points(pos, base, 25.0, 1.0 / w.resolution, w.origin_at_center, 1,
                             1_000_000)
    c = time.perf_counter()
    out["viz_cloud"] = dict(radius_m=25.0, fetch_ms=(b - a) * 1e3, build_ms=(c - b) * 1e3,
                            points=int(len(recs)), grid=f"after {reps} scans")
    out["threads"] = 1
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="readme")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip census, PCIe-inclusive, fetch and batched passes")
    ap.add_argument("--batched-ticks", type=int, default=20,
                    help="timed ticks of the 64-grid batched workload (0: skip)")
    ap.add_argument("--served-seconds", type=float, default=2.0,
                    help="seconds per load level of the 256-session map server (0: skip)")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run the zero-fill and copy calibration kernels (with --no-extras)")
    ap.add_argument("--only", choices=("batched", "served", "driver"),
                    help="run only that workload and print its record (for profiling)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(grids and sessions shard across GPUs with no exchange). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    # The drop-in C++ API prints to stdout (as the reference's does); keep
    # stdout for the one JSON line by pointing fd 1 at stderr meanwhile.
    sys.stdout.flush()
    json_fd = os.dup(1)
    os.dup2(2, 1)

    import cublox_amd as cb

    if cb.device_count() < 1:
        log("no HIP device visible")
        sys.exit(1)
    if args.only == "driver":
        w = cb.WORKLOADS[args.workload]
        scene = w.scene()
        scans = make_scans(w, args.warmup + args.steps, scene)
        bufs, n = upload(cb, scans)
        g, _per, _wall, kt = run_gpu(cb, w, scans, bufs, n, args.warmup, args.steps)
        rec = run_driver_steps(cb, w, scans, bufs, n, args.warmup, args.steps, g,
                               kt["update_ms"] / max(kt["update_launches"], 1))
        sys.stdout.flush()
        os.write(json_fd, (json.dumps({"workload": "driver", **rec}) + "\n").encode())
        os.close(json_fd)
        return
    if args.only:
        rec = (run_batched(cb, 1, args.batched_ticks) if args.only == "batched"
               else run_served(cb, seconds=args.served_seconds))
        sys.stdout.flush()
        os.write(json_fd, (json.dumps({"workload": args.only, **rec}) + "\n").encode())
        os.close(json_fd)
        return
    w = cb.WORKLOADS[args.workload]
    log(f"device: {cb.device_name()}")
    log(f"workload {w.name}: grid {w.map_size} ({w.voxels} voxels) @ {w.resolution} m, "
        f"{w.points} points/scan")
    total = args.warmup + args.steps
    scene = w.scene()
    t = time.perf_counter()
    scans = make_scans(w, total, scene)
    bufs, n = upload(cb, scans)
    log(f"generated + uploaded {total} scans in {time.perf_counter() - t:.2f} s")

    # Headline: the blocking calls alone (no timing events); then the same
    # scans again with HIP events for the per-kernel device times.
    g0, per, wall_ms, _ = run_gpu(cb, w, scans, bufs, n, args.warmup, args.steps,
                                  timing=False)
    rc_cost = recenter_cost(per, run_gpu.moved)
    log(f"recenter critical path: {rc_cost}")
    g0.close()
    # The same blocking calls made one by one from Python (one binding call
    # each): the interpreter's share of a step.
    gp, per_p, wall_p, _ = run_gpu(cb, w, scans, bufs, n, args.warmup, args.steps,
                                   timing=False, native=False)
    gp.close()
    log(f"python loop: {wall_p / args.steps:.4f} ms/scan")
    g, per_t, wall_t, kt = run_gpu(cb, w, scans, bufs, n, args.warmup, args.steps)
    log(f"with timing events: {wall_t / args.steps:.4f} ms/scan")
    ms_per_step = wall_ms / args.steps
    log(f"plain per-scan: {ms_per_step:.4f} ms/scan (p50 {pct(per, 50):.4f}, "
        f"p99 {pct(per, 99):.4f}) kernels {kt}")

    extras = {}
    roofline = None
    rooflines = {}
    if not args.no_extras:
        cen = census(cb, w, scans, bufs, n, args.warmup, args.steps)
        U = statistics.mean(c["touched"] for c in cen)
        visits = statistics.mean(c["visits"] for c in cen)
        Uh = statistics.mean(c["hit_voxels"] for c in cen)
        V = w.voxels
        up_ms = kt["update_ms"] / max(kt["update_launches"], 1)
        re_n = max(kt["recenter_launches"], 1)
        re_ms = kt["recenter_ms"] / re_n
        re_vox = kt["recenter_zeroed_voxels"] / re_n
        # Algorithmic bytes per launch: the fused update reads the points
        # once and reads + writes each touched voxel's log-odds once; the
        # recenter writes each zeroed voxel once.
        up_bytes = 12.0 * n + 8.0 * U
        re_bytes = 4.0 * re_vox

        def rl(bytes_, ms):
            gbs = bytes_ / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
            return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                    "traffic": None, "bytes_per_launch": int(bytes_),
                    "avg_launch_ms": round(ms, 5)}

        rooflines = {"update": rl(up_bytes, up_ms), "recenter": rl(re_bytes, re_ms)}
        # HBM bytes per update launch from the FETCH_SIZE / WRITE_SIZE runs
        # recorded in profiles/ (readme workload only).
        # The figure holds only for the kernel sources it was measured on:
        # with any other sources the traffic is reported as unmeasured.
        tpath = os.path.join(ROOT, "profiles", "update_traffic_readme_701x701x151.json")
        digest = update_source_digest()
        rooflines["update"]["source_digest"] = digest
        if w.name == "readme" and os.path.exists(tpath):
            with open(tpath) as fh:
                rec = json.load(fh)
            if rec.get("source_digest") == digest:
                rooflines["update"]["traffic"] = rec["update_traffic_bytes"]
            else:
                log(f"update traffic in {tpath} was measured on other kernel sources "
                    f"({rec.get('source_digest')} != {digest}): reported as null")
        rooflines["update"]["visits_per_s"] = visits / (up_ms * 1e-3) if up_ms else 0
        totals = {"update": kt["update_ms"], "recenter": kt["recenter_ms"]}
        dom = max(totals, key=totals.get)
        roofline = dict(rooflines[dom], kernel=dom)
        extras["census"] = dict(visits_per_scan=visits, unique_voxels_per_scan=U,
                                hit_voxels_per_scan=Uh,
                                recenter_events=kt["recenter_launches"],
                                zeroed_voxels_per_event=re_vox)
        gl, per_l, wall_l, kt_l = run_gpu(cb, w, scans, bufs, n, args.warmup,
                                          args.steps, path=1)
        gl.close()
        extras["per_voxel_atomics_path"] = dict(
            ms_per_scan=wall_l / args.steps,
            raycast_ms=kt_l["update_ms"] / max(kt_l["update_launches"], 1),
            apply_ms=kt_l["apply_ms"] / max(kt_l["apply_launches"], 1))
        log(f"per-voxel atomics path: {extras['per_voxel_atomics_path']}")
        pcie = run_pcie(cb, w, scans, args.warmup, args.steps)
        extras["pcie_inclusive_ms_per_scan"] = statistics.mean(pcie)
        # The reference's own figure (README.md:32-35) times update() on a
        # host PointCloud, i.e. with the points' host-to-device copies.
        extras["h2d_inclusive"] = dict(
            ms_per_scan=round(statistics.mean(pcie), 5),
            p50_ms=round(pct(pcie, 50), 5), p99_ms=round(pct(pcie, 99), 5),
            form="blocking updateSensorFrame on host columns (copied into a "
                 "grid-owned pinned buffer that the ray walk reads over the bus) "
                 "+ recenter, native loop",
            speedup_vs_3090_readme=round(README_3090_MS / statistics.mean(pcie), 2),
            target_ms=0.5)
        log(f"h2d-inclusive: {extras['h2d_inclusive']}")
        extras["fetch"] = run_fetch(cb, w, g, scene, total - 1)
        if extras["fetch"]:
            rooflines["fetch"] = extras["fetch"]["roofline"]
        extras["driver_steps"] = run_driver_steps(
            cb, w, scans, bufs, n, args.warmup, args.steps, g,
            kt["update_ms"] / max(kt["update_launches"], 1))
        log(f"driver steps: {extras['driver_steps']}")
        extras["zero_fill_256MiB"] = fill_bandwidth(cb)
        log(f"zero fill: {extras['zero_fill_256MiB']}")
        if args.batched_ticks > 0:
            extras["batched_workload"] = run_batched(cb, 1, args.batched_ticks)
            log(f"batched: {extras['batched_workload']}")
        if args.served_seconds > 0:
            extras["served_workload"] = run_served(cb, seconds=args.served_seconds)

    if args.no_extras and args.calibrate:
        extras["zero_fill_256MiB"] = fill_bandwidth(cb)

    cpu = None
    if not args.no_cpu_baseline:
        # The GPU host allots 16 CPU threads to a one-GPU job (its
        # OMP_NUM_THREADS); the OpenMP restatement runs on that share and on
        # one thread, and the host's own counts are reported beside them.
        cpus = host_cpus()
        threads = max(1, min(16, cpus["affinity"]))
        cpu = cpu_baseline(cb, w, scans, args.cpu_budget, threads)
        cpu["host_cpus"] = cpus
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": METRIC,
        "value": round(ms_per_step, 5),
        "unit": "ms/scan",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": round(ms_per_step / README_3090_MS, 5),
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{w.name}: {w.map_size[0]}x{w.map_size[1]}x{w.map_size[2]} "
                        f"@ {w.resolution} m, {w.beams}x{w.azimuths} points/scan, "
                        f"max range {w.max_range} m, {w.n_boxes} boxes, "
                        f"{w.speed} m/scan, recenter threshold {w.threshold} m",
            "voxels": w.voxels,
            "points_per_scan": w.points,
            "form": "plain per-scan, blocking drop-in calls from a native loop "
                    "(cublox_seq_replay), inputs HBM-resident",
        },
        "p50_ms": round(pct(per, 50), 5),
        "python_loop_ms_per_scan": round(wall_p / args.steps, 5),
        "recenter_critical_path": rc_cost,
        "p99_ms": round(pct(per, 99), 5),
        "scans_per_s": round(1e3 / ms_per_step, 2),
        "kernel_ms": {k: round(v / max(kt[k.replace('_ms', '_launches')], 1), 5)
                      for k, v in kt.items() if k.endswith("_ms")},
        "speedup_vs_3090_readme": round(README_3090_MS / ms_per_step, 2),
        "speedup_note": "speedup_vs_3090_readme divides the 3090's H2D-inclusive "
                        "figure by the HBM-resident value; h2d_inclusive holds the "
                        "like-for-like ratio",
        "roofline": roofline,
        "rooflines": rooflines,
        "cpu_baseline": cpu,
    }
    out.update(extras)
    sys.stdout.flush()
    os.write(json_fd, (json.dumps(out) + "\n").encode())
    os.close(json_fd)


if __name__ == "__main__":
    main()
