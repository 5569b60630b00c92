"""Index math of the toroidal window: the oracle and the drop-in host API
(include/cublox/utils.hpp through libcublox.so).

Pins: the reference's five round-trip invariants with its three fixtures
and seeds 42-46, 500 samples each (reference
test/test_utils_index_roundtrip.cpp:44-180), exhaustive checks on tiny maps
against Python integer modular arithmetic, and oracle == product on random
samples.
"""
import ctypes as C

import numpy as np
import pytest

FIXTURES = [  # (half, origin, resolution, centre binning)
    ((200, 200, 100), (17, -42, 8), 0.05, True),
    ((350, 350, 75), (-120, 55, 0), 0.1, False),
    ((3, 5, 7), (2, -1, 4), 0.25, True),
]


def _i3(v):
    return (C.c_int * 3)(*[int(x) for x in v])


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


class Host:
    """The product's host index math (utils.hpp) via libcublox.so."""

    def __init__(self, L):
        self.L = L

    def pos_to_global(self, p, inv, c):
        return self.L.cublox_utils_pos_to_global(float(p), float(inv), int(c))

    def global_to_pos(self, g, res, c):
        return self.L.cublox_utils_global_to_pos(int(g), float(res), int(c))

    def global_to_local(self, g, m, h):
        return self.L.cublox_utils_global_to_local(int(g), int(m), int(h))

    def local_to_global(self, l, m, h, o):
        return self.L.cublox_utils_local_to_global(int(l), int(m), int(h), int(o))

    def local_to_hash(self, l, m, h):
        return self.L.cublox_utils_local_to_hash(_i3(l), _i3(m), _i3(h))

    def hash_to_local(self, hid, m, h):
        out = (C.c_int * 3)()
        self.L.cublox_utils_hash_to_local(int(hid), _i3(m), _i3(h), out)
        return tuple(out)


@pytest.fixture(scope="module")
def host(cb):
    from cublox_amd import _ffi
    return Host(_ffi.host())


def _impls(orc, host):
    return {"oracle": orc, "product": host}


def _fixture_arrays(fx):
    half, origin, res, ctr = fx
    m = tuple(2 * h + 1 for h in half)
    return half, origin, res, ctr, m


@pytest.mark.parametrize("fx", FIXTURES)
@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_reference_round_trips(orc, host, fx, impl):
    I = _impls(orc, host)[impl]
    half, origin, res, ctr, m = _fixture_arrays(fx)
    inv = np.float32(1.0) / np.float32(res)

    def g2l(g):
        return tuple(I.global_to_local(g[a], m[a], half[a]) for a in range(3))

    def l2g(l):
        return tuple(I.local_to_global(l[a], m[a], half[a], origin[a]) for a in range(3))

    def p2g(p):
        return tuple(I.pos_to_global(p[a], inv, ctr) for a in range(3))

    def g2p(g):
        return tuple(I.global_to_pos(g[a], np.float32(res), ctr) for a in range(3))

    # local -> global -> local (seed 42)
    rng = np.random.default_rng(42)
    for _ in range(500):
        l = tuple(int(rng.integers(-half[a], half[a] + 1)) for a in range(3))
        assert g2l(l2g(l)) == l
    # local -> hash -> local (seed 43)
    rng = np.random.default_rng(43)
    for _ in range(500):
        l = tuple(int(rng.integers(-half[a], half[a] + 1)) for a in range(3))
        assert I.hash_to_local(I.local_to_hash(l, m, half), m, half) == l
    # local -> global -> hash -> global (seed 44)
    rng = np.random.default_rng(44)
    for _ in range(500):
        l = tuple(int(rng.integers(-half[a], half[a] + 1)) for a in range(3))
        g = l2g(l)
        hid = I.local_to_hash(g2l(g), m, half)
        assert l2g(I.hash_to_local(hid, m, half)) == g
    # pos -> global -> pos -> global (seed 45)
    rng = np.random.default_rng(45)
    for _ in range(500):
        p = rng.uniform(-20, 20, 3).astype(np.float32)
        g = p2g(p)
        assert p2g(g2p(g)) == g
    # pos -> hash -> pos -> hash (seed 46)
    rng = np.random.default_rng(46)
    for _ in range(500):
        p = rng.uniform(-20, 20, 3).astype(np.float32)
        hid = I.local_to_hash(g2l(p2g(p)), m, half)
        pb = g2p(l2g(I.hash_to_local(hid, m, half)))
        assert I.local_to_hash(g2l(p2g(pb)), m, half) == hid


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_tiny_map_exhaustive_modular(orc, host, impl):
    """global->local is the unique l in [-h, h] congruent to g (mod m);
    local->global is the unique congruent g in the window around origin;
    hashes enumerate [0, V) x-major, z fastest."""
    I = _impls(orc, host)[impl]
    half = (3, 5, 7)
    m = tuple(2 * h + 1 for h in half)
    for a in range(3):
        for g in range(-60, 61):
            assert I.global_to_local(g, m[a], half[a]) == ((g + half[a]) % m[a]) - half[a]
        for o in range(-20, 21, 3):
            window = range(o - half[a], o + half[a] + 1)
            for l in range(-half[a], half[a] + 1):
                want = [g for g in window if (g - l) % m[a] == 0]
                assert I.local_to_global(l, m[a], half[a], o) == want[0]
    seen = []
    for lx in range(-half[0], half[0] + 1):
        for ly in range(-half[1], half[1] + 1):
            for lz in range(-half[2], half[2] + 1):
                seen.append(I.local_to_hash((lx, ly, lz), m, half))
    assert seen == list(range(m[0] * m[1] * m[2]))


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_binning_closed_form(orc, host, impl):
    """Corner binning = floor(p/res); centre binning rounds half away from
    zero; both in float32 (p * inv_res rounded once)."""
    I = _impls(orc, host)[impl]
    rng = np.random.default_rng(7)
    inv = np.float32(1.0) / np.float32(0.1)
    for p in rng.uniform(-50, 50, 2000).astype(np.float32):
        q = np.float32(p) * inv
        assert I.pos_to_global(p, inv, False) == int(np.floor(q))
        r = np.float32(q + (np.float32(-0.5) if p < 0 else np.float32(0.5)))
        assert I.pos_to_global(p, inv, True) == int(np.trunc(r))
    # exact bin edges
    assert I.pos_to_global(0.0, inv, False) == 0
    assert I.pos_to_global(-0.0, inv, False) == 0
    assert I.pos_to_global(np.float32(-1e-7), inv, False) == -1
    assert I.pos_to_global(np.float32(0.05), inv, True) == 1
    assert I.pos_to_global(np.float32(-0.05), inv, True) == -1


def test_oracle_equals_product_random(orc, host):
    rng = np.random.default_rng(11)
    for _ in range(3000):
        m = int(rng.integers(1, 40)) * 2 + 1
        h = (m - 1) // 2
        g = int(rng.integers(-10**6, 10**6))
        o = int(rng.integers(-10**5, 10**5))
        l = int(rng.integers(-h, h + 1))
        assert orc.global_to_local(g, m, h) == host.global_to_local(g, m, h)
        assert orc.local_to_global(l, m, h, o) == host.local_to_global(l, m, h, o)
        p = np.float32(rng.uniform(-500, 500))
        for ctr in (False, True):
            assert orc.pos_to_global(p, 20.0, ctr) == host.pos_to_global(p, 20.0, ctr)
            assert np.float32(orc.global_to_pos(g, 0.05, ctr)).tobytes() == \
                np.float32(host.global_to_pos(g, 0.05, ctr)).tobytes()
