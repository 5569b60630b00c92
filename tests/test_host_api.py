"""Host-side checks that need no GPU: the C ABI libraries load and export
every declared symbol, log-odds constants, the workload generator, and the
drop-in C++ headers compile and run the reference's index round trips."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# logit() of the driver config probabilities (config/cublox_driver.yaml:
# 35-41) with glibc logf on x86-64, as float32 bit patterns.
LOGIT_BITS = {0.9: 0x400C9F53, 0.49: 0xBD23DC91, 0.12: 0xBFFF07F3,
              0.98: 0x4079139A, 0.499: 0xBB831205, 0.85: 0x3FDE076A}


def _declared(header):
    text = open(os.path.join(ROOT, "include", "cublox", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cublox_[a-z0-9_]+)\s*\(", text)))


def test_capi_exports_every_declared_symbol(cb):
    from cublox_amd import _ffi
    L = _ffi.hip()
    names = _declared("capi.h")
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_host_library_exports(cb):
    """Every entry point og_capi.h declares is exported, and the header
    declares every one the binding defines."""
    from cublox_amd import _ffi
    L = _ffi.host()
    names = _declared("og_capi.h")
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    src = open(os.path.join(ROOT, "cublox_amd", "csrc", "og_capi.cpp")).read()
    defined = sorted(set(re.findall(r"^\S.*?\b(cublox_(?:og|gb|utils)_[a-z0-9_]+)\(", src,
                                    flags=re.M)))
    assert sorted(set(defined) - set(names)) == []


def test_replay_library_exports(cb):
    """seq_replay.h's entry point is exported by libcublox_seq.so, and a
    negative count or missing steps is refused before any call is made."""
    import ctypes as C
    from cublox_amd import _ffi
    L = _ffi.seq()
    names = _declared("seq_replay.h")
    assert names == ["cublox_seq_replay"]
    assert hasattr(L, "cublox_seq_replay")
    assert L.cublox_seq_replay(None, -1, 0, None, None) == 1
    assert L.cublox_seq_replay(None, 3, 0, None, None) == 1
    assert L.cublox_seq_replay(None, 0, 0, None, None) == 0
    assert C.sizeof(_ffi.SeqStep) == 96  # pointer fields, n, T[12], base[3]


def test_server_library_exports(cb):
    """ms_capi.h (the map server binding) and ms_capi.cpp agree, and every
    declared entry point is exported."""
    from cublox_amd import _ffi
    L = _ffi.host()
    names = _declared("ms_capi.h")
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    src = "".join(open(os.path.join(ROOT, "cublox_amd", "csrc", f)).read()
                  for f in ("ms_capi.cpp", "ServerLoad.cpp"))
    defined = sorted(set(re.findall(r"^\S.*?\b(cublox_ms_[a-z0-9_]+)\(", src, flags=re.M)))
    assert defined == names


def test_server_session_bytes(cb):
    """Admission charge: 4 B per voxel of the realised map (2 * half + 1 per
    axis) plus the fetch output reserve (500000 samples of 13 B)."""
    reserve = 500000 * 13 + 256
    for half, vox in [((100, 100, 25), 201 * 201 * 51), ((350, 350, 75), 74201551),
                      ((200, 200, 50), 16240901)]:
        assert cb.MapServer.session_bytes(half) == 4 * vox + reserve


@pytest.mark.parametrize("p", sorted(LOGIT_BITS))
def test_logit_constants(cb, orc, p):
    want = LOGIT_BITS[p]
    assert np.float32(cb.logit(p)).view(np.uint32) == want
    assert np.float32(orc.logit(p)).view(np.uint32) == want


def test_scene_is_deterministic(cb):
    w = cb.WORKLOADS["small"]
    a = w.scene().scan(3, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
    b = w.scene().scan(3, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
    for x, y in ((a.xs, b.xs), (a.ys, b.ys), (a.zs, b.zs)):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    assert np.array_equal(a.T, b.T)
    c = cb.Scene(seed=w.seed(1), n_boxes=w.n_boxes, box_extent=w.box_extent,
                 speed=w.speed).scan(3, w.beams, w.azimuths, w.el_lo, w.el_hi,
                                     w.max_range)
    assert not np.array_equal(a.xs, c.xs)


def test_scene_geometry(cb):
    """Returns lie on the ground or a box face (within 5 sigma of noise);
    misses sit at max_range + 1 m; the sensor is 1.8 m above the ground."""
    w = cb.WORKLOADS["small"]
    sc = w.scene()
    boxes = sc.boxes()
    for k in (0, 5):
        s = sc.scan(k, w.beams, w.azimuths, w.el_lo, w.el_hi, w.max_range)
        assert abs(s.T[2, 3] - 1.8) < 1e-5
        r = np.sqrt(s.xs.astype(np.float64) ** 2 + s.ys ** 2 + s.zs ** 2)
        miss = np.abs(r - (w.max_range + 1.0)) < 1e-3
        assert miss.sum() == s.n_miss
        R = s.T[:, :3].astype(np.float64)
        t = s.T[:, 3].astype(np.float64)
        pw = np.stack([s.xs, s.ys, s.zs], 1).astype(np.float64) @ R.T + t
        hits = pw[~miss]
        d_ground = np.abs(hits[:, 2])
        d_box = np.full(len(hits), np.inf)
        for b in boxes:
            lo, hi = b[:3], b[3:]
            inside = np.clip(hits, lo, hi)
            out = np.linalg.norm(hits - inside, axis=1)
            # distance to the surface for points inside the box
            inner = np.min(np.minimum(hits - lo, hi - hits), axis=1)
            dist = np.where(out > 0, out, np.maximum(inner, 0))
            d_box = np.minimum(d_box, dist)
        assert np.all(np.minimum(d_ground, d_box) < 0.1)


def test_workload_shapes(cb):
    w = cb.WORKLOADS
    assert w["small"].map_size == (201, 201, 51) and w["small"].points == 16384
    assert w["readme"].voxels == 74_201_551 and w["readme"].points == 65536
    assert w["batched"].map_size == (401, 401, 101)
    assert w["stress"].voxels == 296_382_951 and w["stress"].points == 262144


def test_cpp_dropin_index_roundtrips(cb, tmp_path):
    """The Eigen-typed drop-in headers compile (against the bundled Eigen
    subset) and pass the reference's index invariants."""
    src = os.path.join(ROOT, "tests", "cpp", "test_utils_roundtrip.cpp")
    exe = tmp_path / "roundtrip"
    inc = [f"-I{ROOT}/include", f"-I{ROOT}/include/cublox/compat"]
    subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", *inc, src,
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "0 failures" in out.stdout


def test_snapshot_header_layout(tmp_path):
    """The C header's struct has the field offsets the format documents
    (the tests' numpy reader uses the same table)."""
    from snapshot_format import HEADER
    fields = [n for n in HEADER.names]
    src = tmp_path / "off.c"
    src.write_text(
        "#include <stddef.h>\n#include <stdio.h>\n#include <cublox/snapshot.h>\n"
        "int main(void){printf(\"%zu\\n\", sizeof(cublox_snapshot_header));\n"
        + "".join(f"printf(\"%zu\\n\", offsetof(cublox_snapshot_header, {f}));\n"
                  for f in fields) + "return 0;}\n")
    exe = tmp_path / "off"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src),
                    "-o", str(exe)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    assert out[0] == HEADER.itemsize == 120
    assert out[1:] == [HEADER.fields[f][1] for f in fields]


def test_index_chain_struct_layout(cb, tmp_path):
    """The numpy records of the device index chain match the C structs."""
    from cublox_amd.device import INDEX_CASE, INDEX_CHAIN
    src = tmp_path / "idx.c"
    src.write_text(
        "#include <stdio.h>\n#include <stddef.h>\n#include <cublox/capi.h>\n"
        "int main(void){printf(\"%zu %zu %zu %zu %zu\\n\", sizeof(cublox_index_case),"
        " sizeof(cublox_index_chain), offsetof(cublox_index_case, resolution),"
        " offsetof(cublox_index_chain, pos), offsetof(cublox_index_chain, hash_id));"
        "return 0;}\n")
    exe = tmp_path / "idx"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src),
                    "-o", str(exe)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    assert out == [INDEX_CASE.itemsize, INDEX_CHAIN.itemsize,
                   INDEX_CASE.fields["res"][1], INDEX_CHAIN.fields["pos"][1],
                   INDEX_CHAIN.fields["hash"][1]]


def test_cpp_dropin_caller_links(cb, tmp_path):
    """A reference-shaped C++ caller (OccupancyGrid.hpp, Eigen PointCloud,
    config_ members, update / recenter / fetchOccupancyAround / isOccupied)
    compiles against the drop-in headers and links against libcublox.so
    (it runs in the GPU tests)."""
    src = os.path.join(ROOT, "tests", "cpp", "dropin_caller.cpp")
    exe = tmp_path / "caller"
    pkg = os.path.join(ROOT, "cublox_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-Wall", "-Werror",
                    f"-I{ROOT}/include", f"-I{ROOT}/include/cublox/compat", src,
                    f"-L{pkg}", "-lcublox", f"-Wl,-rpath,{pkg}", "-o", str(exe)],
                   check=True)
    assert os.path.exists(exe)
