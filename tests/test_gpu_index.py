"""Device-side index math: the reference's device round trip
(test/test_utils_index_roundtrip.cu:160-273: its three map fixtures, seed
48, 128 cases per fixture), plus 4,096 more cases per fixture and cases at
the window faces, run through cublox_launch_index_chain. Every link of the
chain the kernel writes must equal the oracle's host index math on the same
inputs, and the round-trip invariants must hold:
local -> global -> local, local -> hash -> local, hash -> global = local ->
global, local -> pos -> global = local -> global.
"""
import numpy as np
import pytest

from test_index_math import FIXTURES

pytestmark = pytest.mark.gpu


def _cases(half, origin, res, center, seed, n):
    from cublox_amd.device import INDEX_CASE
    rng = np.random.default_rng(seed)
    h = np.array(half)
    c = np.zeros(n + 6, INDEX_CASE)
    c["local"][:n] = rng.integers(-h, h + 1, size=(n, 3))
    for k, (a, s) in enumerate([(0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1)]):
        c["local"][n + k] = 0
        c["local"][n + k, a] = s * h[a]
    c["map"] = 2 * h + 1
    c["half"] = h
    c["origin"] = origin
    res32 = np.float32(res)
    c["res"] = res32
    c["inv"] = np.float32(1.0) / res32
    c["center"] = int(center)
    return c


@pytest.mark.parametrize("fx", range(len(FIXTURES)))
def test_device_index_chain_matches_oracle(cb, orc, gpu_ready, fx):
    from cublox_amd.device import index_chain
    half, origin, res, center = FIXTURES[fx]
    cases = _cases(half, origin, res, center, 48 + fx, 128 + 4096)
    out = index_chain(cases)
    m = cases["map"][0]
    h = cases["half"][0]
    inv = float(cases["inv"][0])
    bad = 0
    for c, r in zip(cases, out):
        l = c["local"]
        g = [orc.local_to_global(int(l[a]), int(m[a]), int(h[a]), int(origin[a]))
             for a in range(3)]
        assert list(r["g"]) == g, (c, r)
        assert list(r["l_back"]) == list(l)
        assert list(r["l_back"]) == [orc.global_to_local(g[a], int(m[a]), int(h[a]))
                                     for a in range(3)]
        hid = orc.local_to_hash(l, m, h)
        assert int(r["hash"]) == hid == orc.global_to_hash(g, m, h)
        assert list(r["l_hash"]) == list(orc.hash_to_local(hid, m, h)) == list(l)
        assert list(r["g_hash"]) == g
        pos = np.array([orc.global_to_pos(g[a], float(res), center) for a in range(3)],
                       np.float32)
        assert np.array_equal(r["pos"].view(np.uint32), pos.view(np.uint32)), (c, r)
        gp = [orc.pos_to_global(float(pos[a]), inv, center) for a in range(3)]
        assert list(r["g_pos"]) == gp
        bad += int(gp != g)
    # The reference's invariant: the position of a voxel maps back to it.
    assert bad == 0
