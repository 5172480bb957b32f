"""The C-ABI library and the host-side front ends, without a GPU: the
library loads and exports every entry point its header declares, the
header is valid C, the C++ front end compiles and its host-only logic
holds, and the Python mirror keeps the reference defaults."""
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "eikonalxx_amd.h")
LIBDIR = os.path.join(ROOT, "eikonalxx_amd")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"EKX_API[^;(]*?\b(ekx_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    assert len(names) >= 40
    assert "ekx_solver3d_solve" in names and "ekx_batch3d_solve" in names


def test_library_exports_every_declared_symbol():
    import eikonalxx_amd as ekx
    lib = ekx.library()
    for name in _declared():
        assert hasattr(lib, name), name
    assert sorted(ekx.EXPORTED_SYMBOLS) == _declared()
    out = subprocess.run(["nm", "-D", "--defined-only",
                          os.path.join(LIBDIR, "libeikonalxx_amd.so")],
                         check=True, capture_output=True, text=True).stdout
    exported = set(re.findall(r"\sT\s(ekx_\w+)", out))
    assert exported == set(_declared())     # nothing else leaks out


def test_header_is_c():
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only",
                    "-x", "c", HEADER], check=True)


def test_host_only_entry_points():
    import ctypes
    import eikonalxx_amd as ekx
    lib = ekx.library()
    assert lib.ekx_version()
    o = ekx.SolverOptions()
    # src/solverOptions.cpp:5-21
    assert o.number_of_sweeps == 5
    assert o.factored_radius == 3
    assert o.algorithm == ekx.SolverAlgorithm.LevelSetMethod
    assert abs(o.convergence_tolerance - 1e-5) < 1e-15
    assert o.precision == ekx.Precision.FP64
    h = ctypes.c_void_p()
    assert lib.ekx_solver3d_create(ctypes.byref(h)) == 0
    assert lib.ekx_solver3d_solve(h) == 2               # not initialized
    assert b"initialized" in lib.ekx_last_error().lower()
    lib.ekx_solver3d_destroy(h)


def _compile(src, out):
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra",
                    "-Werror", "-I", os.path.join(ROOT, "include"), src,
                    "-L", LIBDIR, "-leikonalxx_amd",
                    "-Wl,-rpath," + LIBDIR, "-o", out], check=True)


def test_cpp_front_end_host_logic(tmp_path):
    exe = str(tmp_path / "facade_host")
    _compile(os.path.join(ROOT, "tests", "cpp", "facade_host.cpp"), exe)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok"


def test_cpp_front_end_gpu_program_compiles(tmp_path):
    _compile(os.path.join(ROOT, "tests", "cpp", "facade_solve.cpp"),
             str(tmp_path / "facade_solve"))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_workloads_are_seeded():
    import numpy as np
    from eikonalxx_amd import workloads as W
    a = W.config3_mixed(nx=33, ny=29, nz=21, h=25.0)
    b = W.config3_mixed(nx=33, ny=29, nz=21, h=25.0)
    assert np.array_equal(a["velocity"], b["velocity"])
    assert a["source"] == b["source"]
    s = W.cell_slowness(a["velocity"], 33, 29, 21)
    assert s.size == 32 * 28 * 20 and np.all(s > 0)
    t = W.config5_tables(nx=41, ny=37, nz=21, h=50.0, n_sources=16)
    assert len(t["sources"]) == 16
    assert all(src[2] == 0.0 for src in t["sources"])   # surface stations
