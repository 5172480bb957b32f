"""The VTK rectilinear-grid writers of the C++ front end (the layout of
src/io/vtkRectilinearGrid3d.cpp and vtkRectilinearGrid2d.cpp): a small
program writes models and datasets, this test parses the files back."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "eikonalxx_amd")


class Reader:
    """Walks a legacy VTK file: text lines and big-endian float blocks."""

    def __init__(self, path):
        self.data = open(path, "rb").read()
        self.pos = 0

    def line(self):
        end = self.data.index(b"\n", self.pos)
        text = self.data[self.pos:end].decode()
        self.pos = end + 1
        return text

    def floats(self, n):
        a = np.frombuffer(self.data, dtype=">f4", count=n, offset=self.pos)
        self.pos += 4 * n
        return a.astype(np.float32)

    def ascii(self, n):
        return np.array([float(self.line()) for _ in range(n)])

    def done(self):
        return self.pos == len(self.data)


def _coords(n, d, o, negate=False):
    i = np.arange(n, dtype=np.float32)
    v = np.float32(o) + np.float32(d) * i
    return -v if negate else v


def _header(r, title, kind, dims):
    assert r.line() == "# vtk DataFile Version 3.0"
    assert r.line() == title
    assert r.line() == kind
    assert r.line() == "DATASET RECTILINEAR_GRID"
    assert r.line() == "DIMENSIONS %d %d %d" % dims


def _binary_axes(r, axes):
    for name, vals in zip("XYZ", axes):
        assert r.line() == "%s_COORDINATES %d float" % (name, len(vals))
        np.testing.assert_array_equal(r.floats(len(vals)), vals)
        assert r.line() == ""


def _run(tmp_path):
    exe = str(tmp_path / "vtk_host")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror",
                    "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "vtk_host.cpp"),
                    "-L", LIBDIR, "-leikonalxx_amd",
                    "-Wl,-rpath," + LIBDIR, "-o", exe], check=True)
    out = subprocess.run([exe, str(tmp_path)], capture_output=True,
                         text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    return tmp_path


def test_vtk_files(tmp_path):
    d = _run(tmp_path)
    nx, ny, nz = 4, 3, 5
    v = 1000.0 + 7.0 * np.arange(nx * ny * nz)
    x = _coords(nx, 10.0, 1.0)
    y = _coords(ny, 20.0, 2.0)
    zb = _coords(nz, 5.0, 3.0, negate=True)

    # Model3D::writeVTK: cell velocities 1 / (8 / sum of corners)
    r = Reader(os.path.join(d, "sub", "model3d.vtk"))
    _header(r, "velocity_m/s", "BINARY", (nx, ny, nz))
    _binary_axes(r, (x, y, zb))
    V = v.reshape(nz, ny, nx)
    corners = sum(V[k:k + nz - 1, j:j + ny - 1, i:i + nx - 1]
                  for k in (0, 1) for j in (0, 1) for i in (0, 1))
    vel = (1.0 / (8.0 / corners)).ravel().astype(np.float32)
    assert r.line() == "CELL_DATA %d" % vel.size
    assert r.line() == "SCALARS velocity_m/s float 1"
    assert r.line() == "LOOKUP_TABLE default"
    np.testing.assert_array_equal(r.floats(vel.size), vel)
    assert r.done()

    # ASCII grid: z written as z0 - dz i; XYZ cell data transposed
    r = Reader(os.path.join(d, "grid3d_ascii.vtk"))
    _header(r, "my_title", "ASCII", (nx, ny, nz))
    for name, n, want in (("X", nx, x), ("Y", ny, y),
                          ("Z", nz, np.float32(3.0) - np.float32(5.0)
                           * np.arange(nz, dtype=np.float32))):
        assert r.line() == "%s_COORDINATES %d float" % (name, n)
        np.testing.assert_allclose(r.ascii(n), want, rtol=1e-6)
    nodes = 0.25 * np.arange(nx * ny * nz)
    assert r.line() == "POINT_DATA %d" % nodes.size
    assert r.line() == "SCALARS t_field float 1"
    assert r.line() == "LOOKUP_TABLE default"
    np.testing.assert_allclose(r.ascii(nodes.size), nodes, rtol=1e-6)
    ncx, ncy, ncz = nx - 1, ny - 1, nz - 1
    xyz = 1.5 * np.arange(ncx * ncy * ncz)        # z fastest
    natural = xyz.reshape(ncx, ncy, ncz).transpose(2, 1, 0).ravel()
    assert r.line() == "CELL_DATA %d" % natural.size
    assert r.line() == "SCALARS cells float 1"
    np.testing.assert_allclose(r.ascii(natural.size), natural, rtol=1e-6)
    assert r.done()

    # binary grid: default names, one POINT_DATA, 3D vectors
    r = Reader(os.path.join(d, "grid3d_binary.vtk"))
    _header(r, "vtk_dataset", "BINARY", (nx, ny, nz))
    _binary_axes(r, (x, y, zb))
    assert r.line() == "POINT_DATA %d" % nodes.size
    assert r.line() == "SCALARS dataset float 1"
    assert r.line() == "LOOKUP_TABLE default"
    np.testing.assert_array_equal(r.floats(nodes.size),
                                  nodes.astype(np.float32))
    assert r.line() == "VECTORS grad float "
    vec = (-0.5 * np.arange(3 * nodes.size)).astype(np.float32)
    np.testing.assert_array_equal(r.floats(vec.size), vec)
    assert r.done()


def test_vtk_2d_files(tmp_path):
    d = _run(tmp_path)
    nx, nz = 5, 4
    x = _coords(nx, 2.5, -3.0)
    zb = _coords(nz, 4.0, 1.5, negate=True)
    v = 1500.0 + 11.0 * np.arange(nx * nz)
    r = Reader(os.path.join(d, "model2d.vtk"))
    _header(r, "vp_model", "BINARY", (nx, 1, nz))
    _binary_axes(r, (x, np.zeros(1, np.float32), zb))
    V = v.reshape(nz, nx)
    s = 4.0 / ((V[:-1, :-1] + V[:-1, 1:]) + (V[1:, :-1] + V[1:, 1:]))
    vel = (1.0 / s).ravel().astype(np.float32)
    assert r.line() == "CELL_DATA %d" % vel.size
    assert r.line() == "SCALARS vp_model float 1"
    assert r.line() == "LOOKUP_TABLE default"
    np.testing.assert_array_equal(r.floats(vel.size), vel)
    assert r.done()

    r = Reader(os.path.join(d, "grid2d.vtk"))
    _header(r, "field_2d", "BINARY", (nx, 1, nz))
    _binary_axes(r, (x, np.zeros(1, np.float32), zb))
    xz = 3.0 * np.arange(nx * nz)                  # z fastest
    assert r.line() == "POINT_DATA %d" % (nx * nz)
    assert r.line() == "SCALARS times float 1"
    assert r.line() == "LOOKUP_TABLE default"
    np.testing.assert_array_equal(
        r.floats(nx * nz),
        xz.reshape(nx, nz).T.ravel().astype(np.float32))
    assert r.line() == "VECTORS times_gradient float "
    g = 0.1 * np.arange(2 * nx * nz)
    want = np.zeros((nx * nz, 3), np.float32)
    want[:, 0] = g[0::2]
    want[:, 2] = g[1::2]
    np.testing.assert_array_equal(r.floats(3 * nx * nz), want.ravel())
    assert r.done()
