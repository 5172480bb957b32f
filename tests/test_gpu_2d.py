"""GPU parity of the 2D fast-sweeping solver (src/solver2d.cpp:462-565)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solve_gpu(ekx, g, slow, src, n_sweeps, radius, stations=None):
    s = ekx.Solver2D()
    s.initialize(ekx.SolverOptions(number_of_sweeps=n_sweeps,
                                   factored_radius=radius), g)
    s.set_slowness(slow)
    s.set_source(src)
    if stations is not None:
        s.set_stations(stations)
    s.solve()
    return s


def test_config1_homogeneous(ekx, oracle):
    """201 x 201, h = 10 m, v = 5000 m/s, source (1005, 995) m, radius 3,
    2 iterations: GPU against the oracle (the contract bar is 1e-12
    relative) and against the analytic |x - xs| / v (first-order scheme)."""
    n, h, v = 201, 10.0, 5000.0
    g = ekx.Geometry2D(n, n, h)
    vel = np.full(n * n, v)
    og = oracle.geometry2d(n, n, h)
    slow = oracle.nodal_to_cell2d(n, n, vel)
    src = (1005.0, 995.0)
    s = ekx.Solver2D()
    s.initialize(ekx.SolverOptions(number_of_sweeps=2, factored_radius=3), g)
    s.set_nodal_velocities(vel)
    s.set_source(src)
    s.solve()
    got = s.get_travel_time_field()
    ref = oracle.solve2d(og, slow, src, 2, 3)
    rel = np.abs(got - ref) / np.maximum(ref, slow[0] * h)
    print("2D config 1: max rel vs oracle", rel.max(), "bitwise",
          np.array_equal(got, ref), s.get_stats().as_dict())
    assert rel.max() <= 1e-12
    exact = oracle.homogeneous2d(og, src, v)
    err = np.abs(got - exact)
    print("vs analytic: max abs %.3e s, mean %.3e s" % (err.max(),
                                                       err.mean()))
    assert err.max() < 1e-3


@pytest.mark.parametrize("n_sweeps", [0, 1, 3])
@pytest.mark.parametrize("radius", [-1, 2, 4])
def test_random_model(ekx, oracle, n_sweeps, radius):
    nx, nz, h = 37, 29, 12.5
    rng = np.random.default_rng(17 + radius)
    slow = 1.0 / rng.uniform(1500.0, 4500.0, (nx - 1) * (nz - 1))
    src = (201.0, 160.0)
    s = _solve_gpu(ekx, ekx.Geometry2D(nx, nz, h), slow, src, n_sweeps,
                   radius)
    ref = oracle.solve2d(oracle.geometry2d(nx, nz, h), slow, src, n_sweeps,
                         radius)
    got = s.get_travel_time_field()
    rel = np.abs(got - ref) / np.maximum(ref, slow.min() * h)
    assert rel.max() <= 1e-12


def test_nonuniform_grid(ekx, oracle):
    nx, nz, dx, dz = 41, 33, 10.0, 13.0
    rng = np.random.default_rng(3)
    slow = 1.0 / rng.uniform(2000.0, 3000.0, (nx - 1) * (nz - 1))
    src = (187.0, 211.0)
    s = _solve_gpu(ekx, ekx.Geometry2D(nx, nz, dx, dz), slow, src, 2, 3)
    ref = oracle.solve2d(oracle.geometry2d(nx, nz, dx, dz), slow, src, 2, 3)
    got = s.get_travel_time_field()
    rel = np.abs(got - ref) / np.maximum(ref, slow.min() * dx)
    assert rel.max() <= 1e-12


def test_stations_reference_case(ekx, oracle):
    """testing/solver2d.cpp:1195-1302: a station at the source reads ~0 and
    a station elsewhere equals the hand-computed bilinear value."""
    nx, nz, dx, dz, vel = 21, 20, 100.0, 100.05, 5000.0
    xs = dx * (nx // 2) + dx / 4
    zs = dz * (nz // 2)
    sta = [(xs, zs), (xs + 2 * dx, zs + 3.1 * dz)]
    g = ekx.Geometry2D(nx, nz, dx, dz)
    s = ekx.Solver2D()
    s.initialize(ekx.SolverOptions(number_of_sweeps=0, factored_radius=3), g)
    s.set_nodal_velocities(np.full(nx * nz, vel))
    s.set_source((xs, zs))
    s.set_stations(sta)
    s.solve()
    t = s.get_travel_times_at_stations()
    field = s.get_travel_time_field()
    assert abs(t[0]) < 1e-7
    cx, cz = int((sta[1][0]) / dx), int((sta[1][1]) / dz)
    i00 = cz * nx + cx
    t00, t10, t01, t11 = (field[i00], field[i00 + 1], field[i00 + nx],
                          field[i00 + nx + 1])
    lx, lz = sta[1]
    fxz0 = ((cx + 1) * dx - lx) / dx * t00 + (lx - cx * dx) / dx * t10
    fxz1 = ((cx + 1) * dx - lx) / dx * t01 + (lx - cx * dx) / dx * t11
    fi = ((cz + 1) * dz - lz) / dz * fxz0 + (lz - cz * dz) / dz * fxz1
    assert abs(fi - t[1]) < 1e-8
    slow = oracle.nodal_to_cell2d(nx, nz, np.full(nx * nz, vel))
    og = oracle.geometry2d(nx, nz, dx, dz)
    for k, r in enumerate(sta):
        ref = oracle.station_time2d(og, slow, field, (xs, zs), r)
        assert abs(ref - t[k]) <= 1e-12 * max(abs(ref), 1e-3)


def test_errors(ekx):
    s = ekx.Solver2D()
    with pytest.raises(RuntimeError):
        s.solve()
    s.initialize(ekx.SolverOptions(), ekx.Geometry2D(11, 11, 10.0))
    with pytest.raises(RuntimeError):
        s.solve()                       # no model
    with pytest.raises(ValueError):
        s.set_source((200.0, 5.0))      # outside [0, 100]
    with pytest.raises(ValueError):
        s.set_slowness(np.full(7, 1e-3))
    with pytest.raises(ValueError):
        s.set_nodal_velocities(np.zeros(121))
    s.set_slowness(np.full(100, 1e-3))
    with pytest.raises(RuntimeError):
        s.solve()                       # no source
    with pytest.raises(ValueError):
        s.set_stations([(5.0, 500.0)])
