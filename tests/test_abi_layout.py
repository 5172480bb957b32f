"""The Python ctypes mirror of the C ABI's structs has the C compiler's
sizes and field offsets (a mismatch would pass wrong options or read wrong
statistics silently)."""
import ctypes
import os
import subprocess

import pytest

import eikonalxx_amd as ekx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def layout(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("abi") / "abi_layout")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror",
                    "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "abi_layout.c"),
                    "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True,
                         text=True).stdout
    sizes, offsets = {}, {}
    for line in out.splitlines():
        t, f, v = line.split()
        if f == "size":
            sizes[t] = int(v)
        else:
            offsets[(t, f)] = int(v)
    return sizes, offsets


@pytest.mark.parametrize("cname,pytype", [
    ("ekx_solver_options", ekx._Options),
    ("ekx_solve_stats", ekx.SolveStats),
    ("ekx_geometry3d", ekx._Geometry3D),
    ("ekx_layered_gaussian", ekx._LayeredGaussian),
    ("ekx_problem2d", ekx._Problem2D),
])
def test_struct_layout_matches(layout, cname, pytype):
    sizes, offsets = layout
    assert ctypes.sizeof(pytype) == sizes[cname]
    for (t, f), off in offsets.items():
        if t == cname:
            assert getattr(pytype, f).offset == off, f
