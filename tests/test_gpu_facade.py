"""The C++ front end (include/eikonalxx_amd.hpp) on the GPU: the
reference's own usage pattern compiled against the header, its fields
against the oracle."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "eikonalxx_amd")


def test_cpp_front_end_solves(oracle, tmp_path):
    exe = str(tmp_path / "facade_solve")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I",
                    os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "facade_solve.cpp"),
                    "-L", LIBDIR, "-leikonalxx_amd", "-Wl,-rpath," + LIBDIR,
                    "-o", exe], check=True)
    out = str(tmp_path / "fields.bin")
    r = subprocess.run([exe, out], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    data = np.fromfile(out, dtype=np.float64)
    nx, ny, nz, dx = 9, 10, 11, 100.0
    n3 = nx * ny * nz
    n2 = 21 * 20
    assert data.size == n3 + n2 + 2
    # 3D: FP64 on the GPU equals the oracle bit for bit
    g = oracle.geometry3d(nx, ny, nz, dx, dx, dx, 5.0, 6.0, 7.0)
    slow = oracle.nodal_to_cell3d(nx, ny, nz, np.full(n3, 3000.0))
    src = (5 + dx * (nx // 2) + dx / 4, 6 + dx * (ny // 2) - dx / 4,
           7 + dx * (nz // 4))
    ref = oracle.solve3d(g, slow, src, 1, 2)
    assert np.array_equal(data[:n3], ref)
    # 2D and its stations
    g2 = oracle.geometry2d(21, 20, 100.0, 100.05)
    s2 = oracle.nodal_to_cell2d(21, 20, np.full(n2, 5000.0))
    xs, zs = 100.0 * 10 + 25.0, 100.05 * 10
    ref2 = oracle.solve2d(g2, s2, (xs, zs), 0, 3)
    t2 = data[n3:n3 + n2]
    assert np.max(np.abs(t2 - ref2) / np.maximum(ref2, 1e-3)) <= 1e-12
    sta = data[n3 + n2:]
    for k, rcv in enumerate(((xs, zs), (xs + 200.0, zs + 3.1 * 100.05))):
        want = oracle.station_time2d(g2, s2, ref2, (xs, zs), rcv)
        assert abs(sta[k] - want) <= 1e-12 * max(abs(want), 1e-3)
