"""Work-queue launch order of the column kernel (one launch per sweep whose
waves take column chunks in dependency order) against one launch per
diagonal: bit-identical FP64 fields over seeded random batches, chunk
depths and chain counts, the FP64 polish segment, early exit, eager timed
launches and repeated graph replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _random_batch(rng, ekx):
    n = int(rng.integers(1, 6))
    problems = []
    uniform = bool(rng.integers(0, 2))
    for _ in range(n):
        dims = tuple(int(v) for v in rng.integers(3, 41, size=3))
        d = (10.0, 10.0, 10.0) if uniform else tuple(
            float(v) for v in rng.uniform(5.0, 20.0, size=3))
        g = ekx.Geometry3D(*dims, *d)
        ncell = (dims[0] - 1) * (dims[1] - 1) * (dims[2] - 1)
        slow = 1.0 / rng.uniform(1500.0, 6000.0, ncell)
        src = tuple(float(rng.uniform(0, (m - 1) * dd))
                    for m, dd in zip(dims, d))
        problems.append((g, slow, src))
    return problems


def _batch(ekx, problems, order, **kw):
    b = ekx.Batch3D()
    b.initialize(ekx.SolverOptions(launch_order=order, **kw), problems)
    b.solve()
    return [b.get_travel_time_field(i) for i in range(len(problems))], b


@pytest.mark.parametrize("seed", range(8))
def test_queue_equals_diagonals_random(ekx, seed):
    rng = np.random.default_rng(7000 + seed)
    problems = _random_batch(rng, ekx)
    kw = dict(number_of_sweeps=int(rng.integers(1, 3)),
              factored_radius=int(rng.choice([-1, 2, 3])),
              chunk_planes=int(rng.choice([0, 3, 8, 13, 40])),
              chains=int(rng.choice([0, 1, 2, 3])))
    print("case", seed, len(problems), kw, flush=True)
    a, ba = _batch(ekx, problems, ekx.LaunchOrder.DIAGONALS, **kw)
    q, bq = _batch(ekx, problems, ekx.LaunchOrder.QUEUE, **kw)
    for x, y in zip(a, q):
        assert np.array_equal(x, y)
    sa, sq = ba.get_stats(), bq.get_stats()
    assert sa.updates == sq.updates
    # one launch per sweep and chain
    n_sweeps = 8 * (kw["number_of_sweeps"] + 1)
    assert sq.kernel_launches == n_sweeps * sq.chains


def test_queue_batch_of_surface_sources(ekx):
    """16 sources on one layered model, 4 chains: the tables workload's
    shape at a small size."""
    from eikonalxx_amd import workloads
    w = workloads.config5_tables(nx=73, ny=65, nz=41, h=100.0,
                                 n_sources=16)
    slow = workloads.cell_slowness(w["velocity"], w["nx"], w["ny"], w["nz"])
    g = ekx.Geometry3D(w["nx"], w["ny"], w["nz"], w["h"])
    problems = [(g, slow, s) for s in w["sources"]]
    kw = dict(number_of_sweeps=2, factored_radius=3)
    a, _ = _batch(ekx, problems, ekx.LaunchOrder.DIAGONALS, **kw)
    q, bq = _batch(ekx, problems, ekx.LaunchOrder.QUEUE, **kw)
    assert bq.get_stats().chains == 4
    for x, y in zip(a, q):
        assert np.array_equal(x, y)
    # a second solve replays the captured graph: flags start over
    bq.solve()
    for i, y in enumerate(q):
        assert np.array_equal(bq.get_travel_time_field(i), y)


def _single(ekx, order, **kw):
    from eikonalxx_amd import workloads
    w = workloads.config3_mixed(nx=97, ny=89, nz=65, h=25.0)
    slow = workloads.cell_slowness(w["velocity"], w["nx"], w["ny"], w["nz"])
    s = ekx.Solver3D()
    s.initialize(ekx.SolverOptions(factored_radius=w["radius"],
                                   launch_order=order, **kw),
                 ekx.Geometry3D(w["nx"], w["ny"], w["nz"], w["h"]))
    s.set_slowness(slow)
    s.set_source(w["source"])
    s.solve()
    return s.get_travel_time_field(), s.get_stats()


@pytest.mark.parametrize("chunk", [0, 16, 65])
def test_queue_single_grid(ekx, chunk):
    a, _ = _single(ekx, ekx.LaunchOrder.DIAGONALS, number_of_sweeps=2,
                   chunk_planes=chunk)
    q, st = _single(ekx, ekx.LaunchOrder.QUEUE, number_of_sweeps=2,
                    chunk_planes=chunk)
    assert np.array_equal(a, q)
    assert st.kernel_launches == 24


def test_queue_fp64_polish(ekx):
    """FP32 sweeps, then the FP64 polish iteration in work-queue form."""
    p = ekx.Precision.FP32_POLISH
    a, _ = _single(ekx, ekx.LaunchOrder.DIAGONALS, number_of_sweeps=2,
                   precision=p)
    q, _ = _single(ekx, ekx.LaunchOrder.QUEUE, number_of_sweeps=2,
                   precision=p)
    assert np.array_equal(a, q)


def test_queue_early_exit_and_timing(ekx):
    a, _ = _single(ekx, ekx.LaunchOrder.DIAGONALS, number_of_sweeps=3)
    q, st = _single(ekx, ekx.LaunchOrder.QUEUE, number_of_sweeps=3,
                    early_exit=True, convergence_tolerance=0.0)
    assert np.array_equal(a, q)
    assert st.iterations_run == 3
    t, st = _single(ekx, ekx.LaunchOrder.QUEUE, number_of_sweeps=3,
                    kernel_timing=True)
    assert np.array_equal(a, t)
    assert st.timed_launches == 32
    assert 0 < st.timed_kernel_ms <= st.sweep_ms_fp64 * 1.01


def test_queue_long_dependency_chain(ekx):
    """One 321 x 321 grid in whole columns: its 1681 chunks fit the resident
    waves at once, so the wave of the last diagonal waits for about 80
    column times in a row (the dependency waits are bounded in time, not in
    polls, for this)."""
    nx, ny, nz, h = 321, 321, 161, 20.0
    g = ekx.Geometry3D(nx, ny, nz, h)
    z = np.arange(nz - 1) + 0.5
    slow = np.repeat(1.0 / (3000.0 + 2.0 * z * h), (nx - 1) * (ny - 1))
    src = (100.0 * h, 200.0 * h, 20.0 * h)

    def run(order):
        s = ekx.Solver3D()
        s.initialize(ekx.SolverOptions(number_of_sweeps=1, factored_radius=3,
                                       chunk_planes=nz, launch_order=order),
                     g)
        s.set_slowness(slow)
        s.set_source(src)
        s.solve()
        return s.get_travel_time_field(), s.get_stats()

    a, _ = run(ekx.LaunchOrder.DIAGONALS)
    q, st = run(ekx.LaunchOrder.QUEUE)
    assert st.chunk_planes == nz
    assert np.array_equal(a, q)
