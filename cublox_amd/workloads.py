"""The benchmark workloads (grid shape, sensor, scene, trajectory).

Shapes are realised as map = 2*half + 1 per axis. Log-odds come from the
probabilities of the reference driver config (config/cublox_driver.yaml:
35-41) through logit(); corner binning (yaml:19).

  small     201x201x51 @ 0.1 m, 16x1024 beams (+-15 deg), 10 m, 40 boxes in
            +-10 m, 0.5 m/scan, 2 deg/scan, recenter threshold 2 m, 20 scans
  readme    701x701x151 @ 0.1 m (74.2 M voxels), 64x1024 (+-22.5 deg), 30 m,
            400 boxes in +-60 m, 1 m/scan, threshold 2 m, 200 scans, 1000
            fetch queries of r = 1.5 m within +-20 m of the sensor
  batched   64 grids of 401x401x101 @ 0.1 m, per tick 16..128 beams x
            512..1024 azimuths, max range 10..40 m, 500 ticks
  stress    1401x1401x151 @ 0.05 m, 128x2048, 60 m, 2000 boxes in +-70 m,
            1 m/scan, threshold 0.5 m, 100 scans, 100000 fetches of r = 1 m
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Tuple

# Probabilities of the reference driver config.
PROBABILITIES = dict(hit=0.9, miss=0.49, min=0.12, max=0.98, free=0.499,
                     occupied=0.85)


@dataclass
class Workload:
    name: str
    config_id: int
    half: Tuple[int, int, int]
    resolution: float
    beams: int
    azimuths: int
    el_lo: float
    el_hi: float
    max_range: float
    n_boxes: int
    box_extent: float
    speed: float
    yaw_rate_deg: float
    threshold: Optional[float]
    scans: int
    fetch_queries: int = 0
    fetch_radius: float = 0.0
    fetch_spread: float = 20.0
    origin_at_center: bool = False
    extra: dict = field(default_factory=dict)

    @property
    def map_size(self):
        return tuple(2 * h + 1 for h in self.half)

    @property
    def voxels(self) -> int:
        m = self.map_size
        return m[0] * m[1] * m[2]

    @property
    def points(self) -> int:
        return self.beams * self.azimuths

    def seed(self, grid_id: int = 0) -> int:
        return 1000 * self.config_id + grid_id

    def scene(self, grid_id: int = 0):
        from .scene import Scene
        return Scene(seed=self.seed(grid_id), n_boxes=self.n_boxes,
                     box_extent=self.box_extent, speed=self.speed,
                     yaw_rate_deg=self.yaw_rate_deg)


SMALL = Workload("small", 1, (100, 100, 25), 0.1, 16, 1024, -15.0, 15.0, 10.0,
                 40, 10.0, 0.5, 2.0, 2.0, 20)
README = Workload("readme", 2, (350, 350, 75), 0.1, 64, 1024, -22.5, 22.5,
                  30.0, 400, 60.0, 1.0, 2.0, 2.0, 200, fetch_queries=1000,
                  fetch_radius=1.5, fetch_spread=20.0)
BATCHED = Workload("batched", 3, (200, 200, 50), 0.1, 128, 1024, -22.5, 22.5,
                   40.0, 200, 40.0, 0.5, 2.0, 2.0, 500,
                   extra=dict(grids=64, beams_range=(16, 128),
                              az_range=(512, 1024), range_range=(10.0, 40.0)))
STRESS = Workload("stress", 5, (700, 700, 75), 0.05, 128, 2048, -22.5, 22.5,
                  60.0, 2000, 70.0, 1.0, 2.0, 0.5, 100, fetch_queries=100000,
                  fetch_radius=1.0, fetch_spread=20.0)

ALL = {w.name: w for w in (SMALL, README, BATCHED, STRESS)}


def log_odds():
    """(l_hit, l_miss, l_min, l_max, l_free, l_occupied) via the host logit."""
    from .grid import logit
    p = PROBABILITIES
    return (logit(p["hit"]), logit(p["miss"]), logit(p["min"]), logit(p["max"]),
            logit(p["free"]), logit(p["occupied"]))


def batched_grid(w: Workload, grid_id: int):
    """Per-grid draw of the batched workload: (beams, azimuths, max range,
    scene) from a generator seeded with the grid's seed."""
    import numpy as np
    from .scene import Scene
    rng = np.random.default_rng(w.seed(grid_id))
    lo_b, hi_b = w.extra["beams_range"]
    lo_a, hi_a = w.extra["az_range"]
    lo_r, hi_r = w.extra["range_range"]
    beams = int(rng.integers(lo_b, hi_b + 1))
    az = int(rng.integers(lo_a, hi_a + 1))
    rmax = float(np.float32(rng.uniform(lo_r, hi_r)))
    scene = Scene(seed=w.seed(grid_id), n_boxes=w.n_boxes, box_extent=w.box_extent,
                  speed=w.speed, yaw_rate_deg=w.yaw_rate_deg)
    return beams, az, rmax, scene


def batched_scan_shape(w: Workload, grid_id: int, tick: int):
    """Per (grid, tick) draw of the batched workload's scan: (beams,
    azimuths, max range), i.e. a point count in [8,192, 131,072] and a range
    in [10, 40] m, from a generator seeded with the grid's seed and the
    tick."""
    import numpy as np
    rng = np.random.default_rng([w.seed(grid_id), int(tick)])
    lo_b, hi_b = w.extra["beams_range"]
    lo_a, hi_a = w.extra["az_range"]
    lo_r, hi_r = w.extra["range_range"]
    beams = int(rng.integers(lo_b, hi_b + 1))
    az = int(rng.integers(lo_a, hi_a + 1))
    rmax = float(np.float32(rng.uniform(lo_r, hi_r)))
    return beams, az, rmax

