"""The oracle's outputs derived from a travel-time field: receiver times and
travel-time gradients.

What pins them:
  * the reference's own gradient tests (testing/gradient.cpp:14-88 and
    :90-172): central differences of the analytic homogeneous and
    linear-gradient fields against the analytic gradients, to 5e-5 and 1e-4;
  * exactness: the central and second-order one-sided stencils reproduce the
    gradient of any quadratic to rounding, and trilinear interpolation
    reproduces any multilinear field;
  * the analytic source-cell values (source corners, receivers in the
    source cell) against the closed form of a homogeneous medium;
  * the 3D rules reduce to the 2D ones on a field that does not vary in y.
"""
import math

import numpy as np
import pytest


def _xz(g):
    ix = np.arange(g.nx) * g.dx
    iz = np.arange(g.nz) * g.dz
    return np.meshgrid(ix, iz, indexing="xy")     # [iz, ix]


def test_gradient2d_reference_homogeneous(oracle):
    """testing/gradient.cpp:14-88."""
    vel, nx, nz, dx, dz = 2500.0, 201, 51, 10.0, 10.5
    xs = (nx - 1) / 4.0 * dx + 0.3 * dx
    zs = (nz - 1) / 3.0 * dz - 0.4 * dx
    g = oracle.geometry2d(nx, nz, dx, dz)
    slow = np.full((nx - 1) * (nz - 1), 1.0 / vel)
    T = oracle.homogeneous2d(g, (xs, zs), vel)
    grad = oracle.gradient2d(g, slow, T, (xs, zs)).reshape(nz, nx, 2)
    x, z = _xz(g)
    r = np.hypot(x - xs, z - zs)
    assert np.max(np.abs(grad[..., 0] - (x - xs) / r / vel)) < 5e-5
    assert np.max(np.abs(grad[..., 1] - (z - zs) / r / vel)) < 5e-5


def test_gradient2d_reference_linear_gradient(oracle):
    """testing/gradient.cpp:90-172: v(z) linear from 1000 to 4000 m/s, source
    at the surface; the 3 surface nodes around the source are excluded from
    the z check as the reference excludes them."""
    v_top, v_bot, nx, nz, dx, dz = 1000.0, 4000.0, 401, 201, 100.0, 101.0
    xs, zs = (nx - 1) / 2.0 * dx, 0.0
    g = oracle.geometry2d(nx, nz, dx, dz)
    T = oracle.linear_gradient2d(g, (xs, zs), v_top, v_bot)
    # cell slowness as Model2D::setNodalVelocities averages it
    zn = np.arange(nz) * dz
    vn = v_top + zn * (v_bot - v_top) / ((nz - 1) * dz)
    vel = np.repeat(vn[:, None], nx, axis=1)
    slow = oracle.nodal_to_cell2d(nx, nz, vel.ravel())
    grad = oracle.gradient2d(g, slow, T, (xs, zs)).reshape(nz, nx, 2)
    # analytic gradient of T = acosh(1 + G^2 r^2 / (2 v0 v)) / G
    gz = (v_bot - v_top) / ((nz - 1) * dz)
    x, z = _xz(g)
    ddx, ddz = x - xs, z - zs
    v0, v = v_top + gz * zs, v_top + gz * z
    arg = 1.0 + gz * gz * (ddx * ddx + ddz * ddz) / (2.0 * v0 * v)
    with np.errstate(divide="ignore", invalid="ignore"):
        scale = 1.0 / (gz * np.sqrt(arg * arg - 1.0))
        dargx = gz * gz * ddx / (v0 * v)
        dargz = (gz * gz * ddz / (v0 * v)
                 - gz * gz * (ddx * ddx + ddz * ddz) * gz / (2 * v0 * v * v))
        ax, az = scale * dargx, scale * dargz
    src = (slice(0, 1), slice(int(xs / dx), int(xs / dx) + 1))
    ax[src] = 0.0
    az[src] = 0.0
    assert np.max(np.abs(grad[..., 0] - ax)) < 1e-4
    skip = np.zeros((nz, nx), dtype=bool)
    skip[0, 199:202] = True
    assert np.max(np.abs(grad[..., 1] - az)[~skip]) < 1e-4


def test_gradient2d_quadratic_exact(oracle):
    """Every stencil (interior, both edges, both axes) is exact on a
    quadratic; the source sits on a grid node so no corner is overridden."""
    g = oracle.geometry2d(9, 7, 2.0, 3.0)
    x, z = _xz(g)
    T = 1.0 + 0.5 * x - 0.25 * z + 0.03 * x * x + 0.02 * x * z - 0.01 * z * z
    slow = np.full(8 * 6, 1.0)
    grad = oracle.gradient2d(g, slow, T.ravel(), (4.0, 6.0)).reshape(7, 9, 2)
    np.testing.assert_allclose(grad[..., 0], 0.5 + 0.06 * x + 0.02 * z,
                               rtol=0, atol=1e-13)
    np.testing.assert_allclose(grad[..., 1], -0.25 + 0.02 * x - 0.02 * z,
                               rtol=0, atol=1e-13)


def test_gradient2d_source_corners(oracle):
    """An off-node source: the 4 corners of its cell carry the analytic
    gradient of the source-cell slowness; only the axes along which the
    source is off the lower corner are replaced."""
    g = oracle.geometry2d(8, 8, 10.0)
    slow = np.linspace(1e-3, 2e-3, 49)
    T = np.random.default_rng(3).random(64)
    xs, zs = 33.0, 40.0                      # off in x only (cell 3, 4)
    grad = oracle.gradient2d(g, slow, T, (xs, zs)).reshape(8, 8, 2)
    plain = oracle.gradient2d(g, slow, T, (30.0, 40.0)).reshape(8, 8, 2)
    s0 = slow[4 * 7 + 3]
    for k in range(2):
        for j in range(2):
            ddx, ddz = (3 + j) * 10.0 - xs, (4 + k) * 10.0 - zs
            r = math.hypot(ddx, ddz)
            assert grad[4 + k, 3 + j, 0] == pytest.approx(ddx / r * s0,
                                                           rel=1e-15)
            assert grad[4 + k, 3 + j, 1] == plain[4 + k, 3 + j, 1]
    mask = np.ones((8, 8), dtype=bool)
    mask[4:6, 3:5] = False
    np.testing.assert_array_equal(grad[mask], plain[mask])


def test_gradient3d_quadratic_exact(oracle):
    g = oracle.geometry3d(6, 5, 7, 2.0, 1.5, 2.5)
    z, y, x = np.meshgrid(np.arange(7) * 2.5, np.arange(5) * 1.5,
                          np.arange(6) * 2.0, indexing="ij")
    T = (0.1 * x + 0.2 * y - 0.3 * z + 0.01 * x * y + 0.02 * y * z
         - 0.03 * x * z + 0.04 * x * x - 0.05 * y * y + 0.06 * z * z)
    slow = np.ones(5 * 4 * 6)
    grad = oracle.gradient3d(g, slow, T.ravel(), (4.0, 3.0, 5.0))
    grad = grad.reshape(7, 5, 6, 3)
    np.testing.assert_allclose(grad[..., 0], 0.1 + 0.01 * y - 0.03 * z
                               + 0.08 * x, rtol=0, atol=1e-13)
    np.testing.assert_allclose(grad[..., 1], 0.2 + 0.01 * x + 0.02 * z
                               - 0.1 * y, rtol=0, atol=1e-13)
    np.testing.assert_allclose(grad[..., 2], -0.3 + 0.02 * y - 0.03 * x
                               + 0.12 * z, rtol=0, atol=1e-13)


def test_gradient3d_homogeneous(oracle):
    """The 3D analytic field of a constant velocity: the source corners are
    exact and the finite differences converge to s (x - xs) / r."""
    vel, n, h = 3000.0, 41, 10.0
    g = oracle.geometry3d(n, n, n, h)
    src = (203.0, 187.0, 96.0)
    T = oracle.homogeneous3d(g, src, vel)
    slow = np.full((n - 1) ** 3, 1.0 / vel)
    grad = oracle.gradient3d(g, slow, T, src).reshape(n, n, n, 3)
    z, y, x = np.meshgrid(*(np.arange(n) * h,) * 3, indexing="ij")
    r = np.sqrt((x - src[0]) ** 2 + (y - src[1]) ** 2 + (z - src[2]) ** 2)
    want = np.stack([(x - src[0]), (y - src[1]), (z - src[2])], -1)
    want = want / r[..., None] / vel
    err = np.abs(grad - want)
    # corners of the source cell: the analytic values
    assert np.max(err[9:11, 18:20, 20:22]) < 1e-18
    assert np.max(err) < 5e-5
    # far from the source the error falls with h^2 / r^2
    assert np.max(err[r > 200.0]) < 2e-6


def test_gradient3d_reduces_to_2d(oracle):
    """A field constant in y, source on a y grid line: the x and z
    components on the source's y plane equal the 2D gradient."""
    g2 = oracle.geometry2d(11, 9, 10.0, 12.0)
    g3 = oracle.geometry3d(11, 4, 9, 10.0, 7.0, 12.0)
    rng = np.random.default_rng(5)
    T2 = rng.random(11 * 9) + 1.0
    s2 = rng.random(10 * 8) * 1e-3 + 1e-4
    T3 = np.repeat(T2.reshape(9, 1, 11), 4, axis=1).ravel()
    s3 = np.repeat(s2.reshape(8, 1, 10), 3, axis=1).ravel()
    src2, src3 = (43.0, 50.0), (43.0, 14.0, 50.0)
    G2 = oracle.gradient2d(g2, s2, T2, src2).reshape(9, 11, 2)
    G3 = oracle.gradient3d(g3, s3, T3, src3).reshape(9, 4, 11, 3)
    np.testing.assert_array_equal(G3[:, 2, :, 0], G2[..., 0])
    np.testing.assert_array_equal(G3[:, 2, :, 2], G2[..., 1])
    # the one-sided stencils leave rounding noise on a constant
    np.testing.assert_allclose(G3[..., 1], 0.0, rtol=0, atol=1e-15)


def test_station3d_multilinear_exact(oracle):
    g = oracle.geometry3d(7, 6, 5, 10.0, 8.0, 12.0, 100.0, -50.0, 3.0)
    z, y, x = np.meshgrid(np.arange(5) * 12.0, np.arange(6) * 8.0,
                          np.arange(7) * 10.0, indexing="ij")

    def f(x, y, z):
        return 1.0 + 0.1 * x - 0.2 * y + 0.3 * z + 0.01 * x * y \
            - 0.02 * y * z + 0.003 * x * z + 0.0004 * x * y * z

    T = f(x, y, z).ravel()
    slow = np.ones(6 * 5 * 4)
    src = (100.0 + 5.0, -50.0 + 3.0, 3.0 + 4.0)
    rng = np.random.default_rng(11)
    for _ in range(50):
        r = rng.random(3) * (60.0, 40.0, 48.0)
        if (r < (10.0, 8.0, 12.0)).all():
            continue                           # the source cell
        got = oracle.station_time3d(g, slow, T, src,
                                    (100.0 + r[0], -50.0 + r[1], 3.0 + r[2]))
        assert got == pytest.approx(f(*r), rel=1e-13, abs=1e-13)
    # nodes read back exactly, the last node included
    for node in ((0, 5, 4), (6, 5, 4), (3, 2, 1)):
        r = np.array(node) * (10.0, 8.0, 12.0)
        got = oracle.station_time3d(g, slow, T, src,
                                    (100.0 + r[0], -50.0 + r[1], 3.0 + r[2]))
        assert got == pytest.approx(f(*r), rel=1e-14)


def test_station3d_source_cell_analytic(oracle):
    g = oracle.geometry3d(5, 5, 5, 10.0)
    slow = np.arange(64, dtype=float) * 1e-5 + 1e-4
    T = np.full(125, 99.0)
    src = (21.0, 25.0, 33.0)                   # cell (2, 2, 3)
    s0 = slow[3 * 16 + 2 * 4 + 2]
    rcv = (28.0, 22.0, 39.0)
    d = np.subtract(rcv, src)
    got = oracle.station_time3d(g, slow, T, src, rcv)
    assert got == pytest.approx(np.sqrt(d @ d) * s0, rel=1e-15)
    assert oracle.station_time3d(g, slow, T, src, src) == 0.0
    assert oracle.station_time3d(g, slow, T, src, (60.0, 0.0, 0.0)) == -1.0


def test_station3d_reduces_to_2d(oracle):
    g2 = oracle.geometry2d(9, 8, 10.0, 11.0)
    g3 = oracle.geometry3d(9, 3, 8, 10.0, 5.0, 11.0)
    rng = np.random.default_rng(2)
    T2 = rng.random(72)
    s2 = rng.random(56) + 0.5
    T3 = np.repeat(T2.reshape(8, 1, 9), 3, axis=1).ravel()
    s3 = np.repeat(s2.reshape(7, 1, 8), 2, axis=1).ravel()
    src2, src3 = (31.0, 42.0), (31.0, 5.0, 42.0)
    for _ in range(40):
        x, z = rng.random() * 80.0, rng.random() * 77.0
        want = oracle.station_time2d(g2, s2, T2, src2, (x, z))
        got = oracle.station_time3d(g3, s3, T3, src3, (x, 5.0, z))
        assert got == pytest.approx(want, rel=1e-14, abs=1e-15)


def test_gradient_needs_three_points(oracle):
    """src/gradient.hpp:31-41: central differences need 3 points an axis."""
    g = oracle.geometry2d(2, 5, 1.0)
    with pytest.raises(ValueError, match="3 grid points"):
        oracle.gradient2d(g, np.ones(4), np.zeros(10), (0.5, 0.5))
    g = oracle.geometry3d(4, 2, 4, 1.0)
    with pytest.raises(ValueError, match="3 grid points"):
        oracle.gradient3d(g, np.ones(9), np.zeros(32), (0.5, 0.5, 0.5))
