"""The .bgeo particle export (include/mpm_engine.cuh:446-452 through partio, absent here).

No partio-written file exists in the reference, so the byte layout is pinned by its
structure (header fields, sizes, byte order, the trailer) and by two independent writers,
Python (ckmpm_amd/bgeo.py) and C++ (include/ckmpm.hpp), agreeing byte for byte."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from ckmpm_amd import bgeo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _points(n, seed=0):
    return np.random.default_rng(seed).uniform(-2, 2, (n, 3)).astype(np.float32)


def test_header_fields_and_layout():
    xyz = _points(5)
    data = bgeo.encode_positions(xyz)
    assert len(data) == 41 + 16 * 5 + 2
    assert data[:5] == b"BgeoV"
    fields = struct.unpack(">9i", data[5:41])
    assert fields == (5, 5, 0, 0, 0, 0, 0, 0, 0)
    for i in range(5):  # big-endian x y z, then w = 1
        x, y, z, w = struct.unpack(">4f", data[41 + 16 * i:57 + 16 * i])
        assert (x, y, z) == tuple(float(v) for v in xyz[i]) and w == 1.0
    assert data[-2:] == b"\x00\xff"


@pytest.mark.parametrize("n", [0, 1, 1000])
def test_round_trip(tmp_path, n):
    xyz = _points(n, n)
    p = tmp_path / "model_0_particle_frame_0.bgeo"
    bgeo.write_positions(p, xyz)
    np.testing.assert_array_equal(bgeo.read_positions(p), xyz)


def test_reader_rejects_malformed_files():
    good = bgeo.encode_positions(_points(3))
    with pytest.raises(ValueError):
        bgeo.decode_positions(b"PGEO" + good[4:])
    with pytest.raises(ValueError):
        bgeo.decode_positions(good[:-3])
    bad_version = good[:5] + struct.pack(">i", 4) + good[9:]
    with pytest.raises(ValueError):
        bgeo.decode_positions(bad_version)
    with_prims = good[:13] + struct.pack(">i", 1) + good[17:]
    with pytest.raises(ValueError):
        bgeo.decode_positions(with_prims)


CPP = r'''
#include "ckmpm.hpp"
#include <cstdio>
int main(int argc, char** argv) {
    std::vector<float> xyz = ckmpm::ReadRawPositions(argv[1]);
    ckmpm::WriteBgeoPositions(argv[2], xyz);
    try {
        ckmpm::WriteBgeoPositions(argv[2] + std::string(".bad"), std::vector<float>(4, 0.0f));
    } catch (const ckmpm::Error&) {
        return 0;
    }
    return 3;
}
'''


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs the host C++ compiler")
@pytest.mark.parametrize("n", [0, 7, 4096])
def test_cpp_writer_matches_python_writer(tmp_path, n):
    src = tmp_path / "w.cpp"
    src.write_text(CPP)
    exe = tmp_path / "w"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    xyz = _points(n, 3 + n)
    raw = tmp_path / "in.bin"
    xyz.tofile(raw)
    out = tmp_path / "out.bgeo"
    subprocess.run([str(exe), str(raw), str(out)], check=True)
    assert out.read_bytes() == bgeo.encode_positions(xyz)
