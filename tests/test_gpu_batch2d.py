"""Batched 2D solves (ekx_batch2d_*, no reference counterpart): every
problem's field equals a Solver2D solve of the same geometry, model and
source bit for bit -- grids up to 256 rows (LDS level buffers) and a batch
with a taller grid (memory-only kernel) -- with models shared by pointer."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _single(ekx, g, slow, src, n_sweeps, radius):
    s = ekx.Solver2D()
    s.initialize(ekx.SolverOptions(number_of_sweeps=n_sweeps,
                                   factored_radius=radius), g)
    s.set_slowness(slow)
    s.set_source(src)
    s.solve()
    return s.get_travel_time_field(), s.get_stats()


def _problems(ekx, rng, dims_list, share):
    problems = []
    shared = {}
    for nx, nz, dx, dz in dims_list:
        g = ekx.Geometry2D(nx, nz, dx, dz)
        key = (nx, nz)
        if share and key in shared:
            slow = shared[key]
        else:
            slow = 1.0 / rng.uniform(1500.0, 6000.0, (nx - 1) * (nz - 1))
            shared[key] = slow
        src = (float(rng.uniform(0, (nx - 1) * dx)),
               float(rng.uniform(0, (nz - 1) * dz)))
        problems.append((g, slow, src))
    return problems


@pytest.mark.parametrize("tall", [False, True])
def test_batch2d_equals_single_solves(ekx, tall):
    rng = np.random.default_rng(2200 + tall)
    dims = [(201, 201, 10.0, 10.0), (57, 33, 10.0, 10.0),
            (201, 201, 10.0, 10.0), (120, 250, 7.5, 12.0), (3, 3, 5.0, 5.0),
            (201, 201, 10.0, 10.0)]
    if tall:
        dims.append((40, 300, 10.0, 10.0))
    problems = _problems(ekx, rng, dims, share=True)
    print("case tall=%s, %d problems" % (tall, len(problems)), flush=True)
    b = ekx.Batch2D()
    b.initialize(ekx.SolverOptions(number_of_sweeps=2, factored_radius=3),
                 problems)
    b.solve()
    total = 0
    for i, (g, slow, src) in enumerate(problems):
        ref, st = _single(ekx, g, slow, src, 2, 3)
        assert np.array_equal(b.get_travel_time_field(i), ref), i
        total += st.updates
    st = b.get_stats()
    assert st.updates == total
    assert st.kernel_launches == 1 and st.last_solve_ms > 0
    # a second solve gives the same fields
    b.solve()
    ref0, _ = _single(ekx, *problems[0], 2, 3)
    assert np.array_equal(b.get_travel_time_field(0), ref0)


def test_batch2d_rejects_bad_input(ekx):
    g = ekx.Geometry2D(11, 11, 10.0)
    slow = np.full(100, 2e-4)
    b = ekx.Batch2D()
    with pytest.raises(ValueError):
        b.initialize(ekx.SolverOptions(), [(g, slow, (500.0, 5.0))])
    with pytest.raises(ValueError):
        b.initialize(ekx.SolverOptions(), [(g, -slow, (5.0, 5.0))])
    g2 = ekx.Geometry2D(12, 11, 10.0)       # one model, two grids
    with pytest.raises(ValueError):
        b.initialize(ekx.SolverOptions(), [(g, slow, (5.0, 5.0)),
                                           (g2, slow, (5.0, 5.0))])
    with pytest.raises(RuntimeError):
        b.solve()                               # not initialized
