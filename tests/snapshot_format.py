"""Independent reader of the map snapshot format (include/cublox/snapshot.h)
for the tests: the header as a numpy structured dtype and the payload
decoded into a dense occupancy array."""
import numpy as np

HEADER = np.dtype([
    ("magic", "S8"), ("version", "<u4"), ("encoding", "<u4"),
    ("half_map_size", "<i4", 3), ("resolution", "<f4"),
    ("origin_at_center", "<i4"), ("has_threshold", "<i4"),
    ("recenter_threshold", "<f4"), ("origin_i", "<i4", 3),
    ("origin_f", "<f4", 3), ("first_run", "<i4"),
    ("l_hit", "<f4"), ("l_miss", "<f4"), ("l_min", "<f4"), ("l_max", "<f4"),
    ("l_free", "<f4"), ("l_occupied", "<f4"), ("l_unknown", "<f4"),
    ("max_raycast_range", "<f4"), ("voxel_num", "<i4"), ("reserved", "<u4"),
    ("payload_bytes", "<u8"),
])
assert HEADER.itemsize == 120


def decode(blob: bytes):
    """(header record, dense float32 occupancy, sparse ids or None)."""
    h = np.frombuffer(blob[:HEADER.itemsize], HEADER)[0]
    assert h["magic"] == b"CBLXSNAP" and h["version"] == 1
    payload = np.frombuffer(blob[HEADER.itemsize:], np.uint8)
    assert payload.size == h["payload_bytes"]
    V = int(h["voxel_num"])
    if h["encoding"] == 0:
        return h, payload.view("<f4").copy(), None
    count = int(payload[:4].view("<u4")[0])
    ids = payload[8:8 + 4 * count].view("<i4")
    vals = payload[8 + 4 * count:8 + 8 * count].view("<f4")
    occ = np.zeros(V, np.float32)
    occ[ids] = vals
    return h, occ, ids.copy()
