"""GPU parity of the 3D solver: the FP64 HIP path against the CPU oracle.

Both kernel families are exact topological reorderings of the serial
fast-sweeping loop, and the library is built with -ffp-contract=off, so the
FP64 fields must equal the oracle's bit for bit on every case here (the
bar is equality; the 129^3 case also reports the relative error the
north-star tolerance of 1e-12 is stated for).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def smooth_slowness(nx, ny, nz, seed, vmin=2000.0, vmax=6000.0):
    rng = np.random.default_rng(seed)
    v = rng.uniform(vmin, vmax, size=(nz - 1, ny - 1, nx - 1))
    # mild smoothing so that refracted and diagonal operators all occur
    for ax in range(3):
        v = 0.5 * v + 0.25 * (np.roll(v, 1, ax) + np.roll(v, -1, ax))
    return (1.0 / v).reshape(-1)


CASES = [
    # (nx, ny, nz), (dx, dy, dz), source (fractions of the extent), radius
    ((3, 3, 3), (10.0, 10.0, 10.0), (0.5, 0.5, 0.5), 3),
    ((5, 4, 6), (10.0, 10.0, 10.0), (0.0, 0.0, 0.0), 2),
    ((9, 10, 11), (10.0, 10.0, 10.0), (0.55, 0.45, 0.25), 2),
    ((17, 12, 9), (25.0, 25.0, 25.0), (1.0, 1.0, 1.0), 3),
    ((20, 23, 19), (10.0, 10.0, 10.0), (0.31, 0.77, 0.52), -1),
    ((33, 17, 25), (10.0, 10.0, 10.0), (0.5, 0.25, 0.9), 4),
    ((12, 14, 10), (10.0, 12.0, 15.0), (0.4, 0.6, 0.3), 2),
    ((16, 16, 16), (10.0, 10.0, 10.0), (0.5, 0.5, 0.5), 40),
]


def _source(shape, h, frac, x0=(0.0, 0.0, 0.0)):
    return tuple(x0[k] + frac[k] * (shape[k] - 1) * h[k] for k in range(3))


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("kernel", ["HYPERPLANE", "COLUMN"])
@pytest.mark.parametrize("n_sweeps", [0, 1, 2])
def test_fp64_bitwise_vs_oracle(ekx, oracle, case, kernel, n_sweeps):
    (nx, ny, nz), h, frac, radius = CASES[case]
    print("case", case, (nx, ny, nz), h, frac, radius, kernel, n_sweeps,
          flush=True)
    slow = smooth_slowness(nx, ny, nz, seed=100 + case)
    src = _source((nx, ny, nz), h, frac)
    g = oracle.geometry3d(nx, ny, nz, h[0], h[1], h[2])
    ref = oracle.solve3d(g, slow, src, n_sweeps, radius)
    opts = ekx.SolverOptions(number_of_sweeps=n_sweeps,
                             factored_radius=radius,
                             kernel=getattr(ekx.Kernel, kernel))
    solver = ekx.Solver3D()
    solver.initialize(opts, ekx.Geometry3D(nx, ny, nz, *h))
    solver.set_slowness(slow)
    solver.set_source(src)
    solver.solve()
    assert solver.have_travel_time_field()
    got = solver.get_travel_time_field()
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, (
        "%d nodes differ, first %s: gpu %r oracle %r"
        % (bad.size, bad[:5], got[bad[:5]], ref[bad[:5]]))


@pytest.mark.parametrize("kernel", ["HYPERPLANE", "COLUMN"])
@pytest.mark.parametrize("algorithm", ["FastSweepingMethod",
                                       "LevelSetMethod"])
def test_shipped_schedule(ekx, oracle, kernel, algorithm):
    nx, ny, nz = 13, 11, 9
    slow = smooth_slowness(nx, ny, nz, seed=7)
    src = (61.0, 42.0, 33.0)
    g = oracle.geometry3d(nx, ny, nz, 10.0)
    alg = getattr(ekx.SolverAlgorithm, algorithm)
    ref = oracle.solve3d(g, slow, src, 3, 3, algorithm=int(alg),
                         schedule=oracle.SCHEDULE_SHIPPED)
    opts = ekx.SolverOptions(number_of_sweeps=3, factored_radius=3,
                             algorithm=alg, schedule=ekx.Schedule.SHIPPED,
                             kernel=getattr(ekx.Kernel, kernel))
    solver = ekx.Solver3D()
    solver.initialize(opts, ekx.Geometry3D(nx, ny, nz, 10.0))
    solver.set_slowness(slow)
    solver.set_source(src)
    solver.solve()
    assert np.array_equal(solver.get_travel_time_field(), ref)


def test_linear_gradient_129(ekx, oracle):
    """The 129^3 linear-gradient workload, 3 iterations, radius 3: GPU FP64
    against the oracle (bitwise; the tolerance of the contract is 1e-12)."""
    n, h = 129, 10.0
    z = np.arange(n) * h
    vel = np.repeat((3000.0 + 0.5 * z)[:, None], n * n, axis=1).reshape(-1)
    slow = oracle.nodal_to_cell3d(n, n, n, vel)
    src = (640.0, 640.0, 200.0)
    g = oracle.geometry3d(n, n, n, h)
    ref = oracle.solve3d(g, slow, src, 3, 3)
    for kernel in (ekx.Kernel.COLUMN, ekx.Kernel.HYPERPLANE):
        solver = ekx.Solver3D()
        solver.initialize(ekx.SolverOptions(number_of_sweeps=3,
                                            factored_radius=3, kernel=kernel),
                          ekx.Geometry3D(n, n, n, h))
        solver.set_nodal_velocities(vel)
        solver.set_source(src)
        solver.solve()
        got = solver.get_travel_time_field()
        rel = np.abs(got - ref) / np.maximum(ref, 2e-4 * h)
        print(kernel.name, "max rel", rel.max(), "bitwise",
              np.array_equal(got, ref), solver.get_stats().as_dict(),
              flush=True)
        assert rel.max() <= 1e-12
        assert np.array_equal(got, ref)


def test_batch_uneven_fp64(ekx, oracle):
    rng = np.random.default_rng(11)
    problems, refs = [], []
    shared = None
    for i in range(6):
        dims = tuple(int(v) for v in rng.integers(5, 27, size=3))
        if i == 4:
            dims = problems[3][0].nx, problems[3][0].ny, problems[3][0].nz
        h = 20.0
        slow = smooth_slowness(*dims, seed=200 + i)
        if i == 4:
            slow = shared           # shares problem 3's model
        if i == 3:
            shared = slow
        frac = rng.uniform(0, 1, size=3)
        src = _source(dims, (h, h, h), frac)
        problems.append((ekx.Geometry3D(*dims, h), slow, src))
        g = oracle.geometry3d(*dims, h)
        refs.append(oracle.solve3d(g, slow, src, 2, 3))
    batch = ekx.Batch3D()
    batch.initialize(ekx.SolverOptions(number_of_sweeps=2,
                                       factored_radius=3), problems)
    batch.solve()
    for i, ref in enumerate(refs):
        got = batch.get_travel_time_field(i)
        assert np.array_equal(got, ref), "problem %d differs" % i
    # a second solve with moved sources reuses the captured schedule
    new_src = _source((problems[0][0].nx, problems[0][0].ny,
                       problems[0][0].nz), (20.0,) * 3, (0.1, 0.9, 0.4))
    batch.set_source(0, new_src)
    batch.solve()
    g0 = problems[0][0]
    ref0 = oracle.solve3d(oracle.geometry3d(g0.nx, g0.ny, g0.nz, 20.0),
                          problems[0][1], new_src, 2, 3)
    assert np.array_equal(batch.get_travel_time_field(0), ref0)
    assert np.array_equal(batch.get_travel_time_field(1), refs[1])
