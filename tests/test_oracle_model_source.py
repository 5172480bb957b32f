"""Source location, model preparation and the analytic fields of the
oracle, against the reference's unit tests (testing/source.cpp,
testing/model.cpp, testing/analytic.cpp, testing/solver3d.cpp:19-62,
testing/solver2d.cpp:61-136)."""
import math

import numpy as np
import pytest


def test_source3d(oracle):
    """testing/source.cpp:62-126."""
    nx, ny, nz = 21, 18, 20
    dx, dy, dz = 100.0, 200.0, 100.05
    x0, y0, z0 = 100.0, -10.0, 200.0
    xs = x0 + dx * (nx // 2) + dx / 4
    ys = y0 + dy * (ny // 2) - dy / 4
    zs = z0 + dz * (nz // 2)
    g = oracle.geometry3d(nx, ny, nz, dx, dy, dz, x0, y0, z0)
    off, cell, idx = oracle.source3d(g, xs, ys, zs)
    assert np.allclose(off, (xs - x0, ys - y0, zs - z0), atol=1e-14, rtol=0)
    assert tuple(cell) == (10, 8, 10)
    assert idx == 10 * (nx - 1) * (ny - 1) + 8 * (nx - 1) + 10
    # free surface: z at the origin -> cell 0
    _, cell, _ = oracle.source3d(g, xs, ys, z0)
    assert cell[2] == 0
    # the last node belongs to the last cell (min(nCell - 1, .))
    _, cell, _ = oracle.source3d(g, x0 + (nx - 1) * dx, ys, zs)
    assert cell[0] == nx - 2
    for bad in ((x0 - 1, ys, zs), (xs, y0 + (ny - 1) * dy + 1, zs),
                (xs, ys, z0 - 1e-6)):
        with pytest.raises(ValueError):
            oracle.source3d(g, *bad)


def test_source2d(oracle):
    """testing/source.cpp:12-60."""
    nx, nz, dx, dz, x0, z0 = 21, 20, 100.0, 100.05, 100.0, 200.0
    xs = x0 + dx * (nx // 2) + dx / 4
    zs = z0 + dz * (nz // 2)
    g = oracle.geometry2d(nx, nz, dx, dz, x0, z0)
    off, cell, idx = oracle.source2d(g, xs, zs)
    assert np.allclose(off, (xs - x0, zs - z0), atol=1e-14, rtol=0)
    assert tuple(cell) == (10, 10)
    assert idx == (nx - 1) * 10 + 10


def _nodal_to_cell_ref3d(nx, ny, nz, v):
    """The test's own interpolation (testing/model.cpp:47-75): the cell's
    slowness is one over the mean of its 8 nodal velocities."""
    v = v.reshape(nz, ny, nx)
    avg = (v[:-1, :-1, :-1] + v[:-1, :-1, 1:] + v[:-1, 1:, :-1]
           + v[:-1, 1:, 1:] + v[1:, :-1, :-1] + v[1:, :-1, 1:]
           + v[1:, 1:, :-1] + v[1:, 1:, 1:]) / 8.0
    return (1.0 / avg).ravel()


def test_model3d(oracle):
    """testing/model.cpp:174-266 (50 x 51 x 60, velocities 1..N): nodal ->
    cell in natural and transposed (XYZ) order, and cell velocities."""
    nx, ny, nz = 50, 51, 60
    v = np.arange(1, nx * ny * nz + 1, dtype=np.float64)
    ref = _nodal_to_cell_ref3d(nx, ny, nz, v)
    got = oracle.nodal_to_cell3d(nx, ny, nz, v, 0)
    assert np.max(np.abs(got - ref)) < 1e-10
    vt = v.reshape(nz, ny, nx).transpose(2, 1, 0).ravel()   # z fastest
    got = oracle.nodal_to_cell3d(nx, ny, nz, vt, 1)
    assert np.max(np.abs(got - ref)) < 1e-10
    ncell = (nx - 1) * (ny - 1) * (nz - 1)
    vc = np.arange(1, ncell + 1, dtype=np.float64)
    got = oracle.cell_velocity_to_slowness3d(nx, ny, nz, vc, 0)
    assert np.max(np.abs(got - 1.0 / vc)) < 1e-10
    vct = vc.reshape(nz - 1, ny - 1, nx - 1).transpose(2, 1, 0).ravel()
    got = oracle.cell_velocity_to_slowness3d(nx, ny, nz, vct, 1)
    assert np.max(np.abs(got - 1.0 / vc)) < 1e-10
    with pytest.raises(ValueError):
        oracle.nodal_to_cell3d(3, 3, 3, np.r_[np.ones(26), 0.0])


def test_model2d(oracle):
    """testing/model.cpp:94-172 (420 x 220)."""
    nx, nz = 420, 220
    v = np.arange(1, nx * nz + 1, dtype=np.float64)
    w = v.reshape(nz, nx)
    ref = (1.0 / ((w[:-1, :-1] + w[:-1, 1:] + w[1:, :-1] + w[1:, 1:])
                  / 4.0)).ravel()
    got = oracle.nodal_to_cell2d(nx, nz, v, 0)
    assert np.max(np.abs(got - ref)) < 1e-10
    got = oracle.nodal_to_cell2d(nx, nz, w.T.ravel(), 1)
    assert np.max(np.abs(got - ref)) < 1e-10


def test_analytic_time3d(oracle):
    """testing/solver3d.cpp:19-62 (1e-14)."""
    d = (100.0, 150.0, 200.0)
    src = (500.0, 750.0, 1000.0)
    s = 1.0 / 5000.0
    x, y, z = d
    dist = math.sqrt((x - src[0]) ** 2 + (y - src[1]) ** 2
                     + (z - src[2]) ** 2)
    t = oracle.analytic_time3d((1, 1, 1), d, src, s)
    assert abs(t - dist * s) < 1e-14
    t, gx, gy, gz = oracle.analytic_time_deriv3d((1, 1, 1), d, src, s)
    assert abs(t - dist * s) < 1e-14
    for gval, dv, sv in ((gx, x, src[0]), (gy, y, src[1]), (gz, z, src[2])):
        assert abs(gval - (dv - sv) / dist * s) < 1e-14
    # at the source the gradient is defined as 0
    t, gx, gy, gz = oracle.analytic_time_deriv3d((0, 0, 0), (1, 1, 1),
                                                 (0, 0, 0), s)
    assert (t, gx, gy, gz) == (0.0, 0.0, 0.0, 0.0)


def test_analytic_time2d(oracle):
    """testing/solver2d.cpp:61-136: the four corners of a cell around the
    source are equidistant."""
    dx, dz, xs, zs, s = 100.0, 200.0, 50.0, 100.0, 1.0 / 5000.0
    tref = math.sqrt(xs * xs + zs * zs) * s
    for ik in ((0, 0), (1, 0), (1, 1), (0, 1)):
        assert abs(oracle.analytic_time2d(ik, (dx, dz), (xs, zs), s)
                   - tref) < 1e-10
    assert oracle.analytic_time2d((0, 0), (1.0, 1.0), (0.0, 0.0), s) == 0.0


def test_homogeneous3d(oracle):
    """testing/analytic.cpp:64-116 (1e-15)."""
    nx, ny, nz, d = 43, 43, 23, (100.0, 101.0, 102.0)
    o = (1.0, 1.5, 2.0)
    v = 4000.0
    src = ((nx - 1) / 4.0 * d[0] + o[0] + 4,
           3 * (ny - 1) / 4.0 * d[1] + o[1] - 80,
           (nz - 1) / 3.0 * d[2] + o[2] - 40)
    g = oracle.geometry3d(nx, ny, nz, *d, *o)
    T = oracle.homogeneous3d(g, src, v)
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx),
                             indexing="ij")
    ref = np.sqrt((o[0] + ix * d[0] - src[0]) ** 2
                  + (o[1] + iy * d[1] - src[1]) ** 2
                  + (o[2] + iz * d[2] - src[2]) ** 2) / v
    assert np.max(np.abs(T - ref.ravel())) < 1e-15


def test_homogeneous2d(oracle):
    """testing/analytic.cpp:15-62 (1e-15)."""
    nx, nz, dx, dz, x0, z0, v = 55, 23, 100.0, 101.0, 1.0, 2.0, 3000.0
    xs = (nx - 1) / 4.0 * dx + x0 + 4
    zs = (nz - 1) / 3.0 * dz + z0 - 40
    T = oracle.homogeneous2d(oracle.geometry2d(nx, nz, dx, dz, x0, z0),
                             (xs, zs), v)
    iz, ix = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    ref = np.hypot(x0 + ix * dx - xs, z0 + iz * dz - zs) / v
    assert np.max(np.abs(T - ref.ravel())) < 1e-15


def test_linear_gradient2d(oracle):
    """testing/analytic.cpp:117-190 (1e-12)."""
    v1, v2 = 1000.0, 4000.0
    nx, nz, dx, dz, x0, z0 = 56, 24, 200.0, 201.0, 0.0, 1.0
    xs = (nx - 1) / 4.0 * dx + x0 + 4
    zs = (nz - 1) / 4.0 * dz + z0 - 40
    T = oracle.linear_gradient2d(oracle.geometry2d(nx, nz, dx, dz, x0, z0),
                                 (xs, zs), v1, v2)
    gz = abs(v2 - v1) / ((nz - 1) * dz)
    vs = v1 + (zs - z0) * gz
    worst = 0.0
    for iz in range(nz):
        for ix in range(nx):
            ddx, ddz = x0 + ix * dx - xs, z0 + iz * dz - zs
            vel = vs + gz * ddz
            arg = 1 + (0.5 / vel) * (gz * gz / vs) * (ddx * ddx + ddz * ddz)
            worst = max(worst, abs(math.acosh(arg) / gz - T[iz * nx + ix]))
    assert worst < 1e-12


def test_linear_gradient3d_satisfies_eikonal(oracle):
    """No reference test covers the 3D linear gradient field: it is pinned
    by the eikonal equation itself.  Central differences of T on a fine
    grid give |grad T| = 1 / v(z) to O(h^2), and at the source depth the
    field reduces to the 2D field along any horizontal direction."""
    n, h = 61, 5.0
    v1, v2 = 2000.0, 5000.0
    src = (151.0, 150.0, 100.0)
    g = oracle.geometry3d(n, n, n, h)
    T = oracle.linear_gradient3d(g, src, v1, v2).reshape(n, n, n)
    gz = (v2 - v1) / ((n - 1) * h)
    z = np.arange(n) * h
    gx_, gy_, gz_ = np.gradient(T, h, h, h)   # arrays are [z, y, x]
    mag = np.sqrt(gx_ ** 2 + gy_ ** 2 + gz_ ** 2)
    slow = 1.0 / (v1 + gz * z)
    inner = (slice(3, -3), slice(3, -3), slice(3, -3))
    # stay away from the source singularity
    zz, yy, xx = np.meshgrid(z, z, z, indexing="ij")
    far = np.sqrt((xx - src[0]) ** 2 + (yy - src[1]) ** 2
                  + (zz - src[2]) ** 2) > 120.0
    rel = np.abs(mag / slow[:, None, None] - 1.0)
    assert np.max(rel[inner][far[inner]]) < 3e-3
    # the 2D field in the x-z plane through the source
    g2 = oracle.geometry2d(n, n, h)
    T2 = oracle.linear_gradient2d(g2, (src[0], src[2]), v1, v2).reshape(n, n)
    iy = int(round(src[1] / h))
    if abs(iy * h - src[1]) < 1e-12:
        assert np.max(np.abs(T[:, iy, :] - T2)) < 1e-12
    # constant velocity: the homogeneous field
    Th = oracle.linear_gradient3d(g, src, 3000.0, 3000.0)
    assert np.max(np.abs(Th - oracle.homogeneous3d(g, src, 3000.0))) < 1e-15
