"""The native scan-loop replay (include/cublox/seq_replay.h) makes exactly
the drop-in calls a driver makes (blocking update with the scan's pose,
then recenter at the base; reference main.cpp:437-456): grids driven by it,
from device and from host columns, equal the oracle replaying the same
calls bit for bit, and report the same recenter moves."""
import numpy as np
import pytest

from helpers import assert_occ_bitexact, make_pair

pytestmark = pytest.mark.gpu


def test_replay_matches_oracle(cb, orc, gpu_ready):
    from cublox_amd.grid import Replay
    w = cb.WORKLOADS["small"]
    scene = w.scene()
    scans = [scene.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
             for k in range(8)]
    g_dev, o = make_pair(cb, orc, w, threads=16)
    g_host, _ = make_pair(cb, orc, w)
    bufs = [[cb.DeviceBuffer.from_array(np.ascontiguousarray(c)) for c in (s.xs, s.ys, s.zs)]
            for s in scans]
    print(f"[replay] small workload, {len(scans)} scans of {scans[0].n} points", flush=True)
    dev = Replay([(g_dev, b[0].ptr, b[1].ptr, b[2].ptr, s.n, s.T, s.base)
                  for s, b in zip(scans, bufs)])
    host = Replay([(g_host, s.xs, s.ys, s.zs, s.n, s.T, s.base) for s in scans], on_host=True)
    ms_dev, moved_dev = dev.run()
    ms_host, moved_host = host.run()
    moved_o = []
    for s in scans:
        o.update_sensor_frame((s.xs, s.ys, s.zs), s.T)
        moved_o.append(bool(o.recenter(s.base)))
    assert moved_dev == moved_o and moved_host == moved_o
    assert any(moved_o), "the sequence should recenter at least once"
    assert all(t > 0 for t in ms_dev + ms_host)
    ref = o.occupancy()
    assert_occ_bitexact(g_dev.occupancy(), ref, "native replay, device columns")
    assert_occ_bitexact(g_host.occupancy(), ref, "native replay, host columns")
    for x in (g_dev, g_host):
        x.close()
