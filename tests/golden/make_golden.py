#!/usr/bin/env python3
"""Writes the committed golden fixtures of this directory.

All values come from this repository's own generator and CPU oracle; none
are taken from the reference. Fixtures:
  scangen.json       first points (float32 bit patterns), pose and miss count
                     of selected scans of the small and readme workloads;
  small_oracle.json  per scan of the 20-scan small workload: voxel visits,
                     touched voxels, hits, SHA-256 of the op/hit counters,
                     of the occupancy after apply and after recenter, the
                     window origin, plus fetch multiset digests at the end.
Run from the repository root: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import cublox_amd as cb  # noqa: E402
import oracle  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def multiset_digest(pos, st) -> str:
    bits = np.ascontiguousarray(pos, np.float32).view(np.uint32).reshape(-1, 3)
    rows = np.concatenate([bits.astype(np.uint64), st.astype(np.uint64)[:, None]], 1)
    if len(rows):
        rows = rows[np.lexsort(rows.T[::-1])]
    return sha(rows)


def scangen_fixture():
    out = {}
    for name, scans in (("small", (0, 7)), ("readme", (0, 3))):
        w = cb.WORKLOADS[name]
        sc = w.scene()
        for k in scans:
            s = sc.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
            out[f"{name}/{k}"] = dict(
                xs=s.xs[:16].view(np.uint32).tolist(),
                ys=s.ys[:16].view(np.uint32).tolist(),
                zs=s.zs[:16].view(np.uint32).tolist(),
                T=s.T.reshape(-1).view(np.uint32).tolist(),
                base=s.base.view(np.uint32).tolist(),
                n_miss=int(s.n_miss),
                digest=sha(np.stack([s.xs, s.ys, s.zs])))
    return out


def small_oracle_fixture():
    w = cb.WORKLOADS["small"]
    sc = w.scene()
    g = oracle.OracleGrid(w.half, w.resolution, w.origin_at_center, w.threshold)
    g.set_max_raycast_range(w.max_range)
    g.set_log_odds_params(*cb.log_odds())
    scans = []
    for k in range(w.scans):
        s = sc.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
        g.raycast(s.world(), s.sensor_origin)
        op, hit = g.counters()
        g.apply()
        occ_a = g.occupancy()
        action = g.recenter(s.base)
        occ_r = g.occupancy()
        scans.append(dict(visits=int(op.sum()), touched=int((op != 0).sum()),
                          hits=int(hit.sum()), op=sha(op), hit=sha(hit),
                          occ_apply=sha(occ_a), occ_recenter=sha(occ_r),
                          action=int(action), origin=list(g.origin()[0])))
    qs = sc.queries(w.scans - 1, 6, 6.0)
    fetches = []
    for c in qs:
        for filt in (None, 2):
            pos, st, n = g.fetch(c, 1.5, filt)
            fetches.append(dict(center=np.asarray(c, np.float32).view(np.uint32).tolist(),
                                filter=-1 if filt is None else filt, count=int(n),
                                digest=multiset_digest(pos, st)))
    return dict(scans=scans, fetches=fetches)


def main():
    with open(os.path.join(HERE, "scangen.json"), "w") as f:
        json.dump(scangen_fixture(), f, indent=1)
    with open(os.path.join(HERE, "small_oracle.json"), "w") as f:
        json.dump(small_oracle_fixture(), f, indent=1)
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
