"""The launch-shape choices (no reference counterpart): the size-class
default of the column z-chunk depth and the measured autotuner.  Neither
changes a field: every launch shape is a topological order of the same
updates."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problems(ekx, n, dims, h=12.0, seed=5):
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    g = ekx.Geometry3D(nx, ny, nz, h)
    slow = 1.0 / rng.uniform(1500.0, 5000.0, (nx - 1) * (ny - 1) * (nz - 1))
    return [(g, slow, tuple(float(rng.uniform(0, (m - 1) * h))
                            for m in dims)) for _ in range(n)]


def test_single_grid_default_chunks_and_matches_whole_columns(ekx, oracle):
    problems = _problems(ekx, 1, (41, 37, 33))
    g, slow, src = problems[0]
    fields = {}
    for chunk in (0, 10 ** 6):
        s = ekx.Solver3D()
        s.initialize(ekx.SolverOptions(number_of_sweeps=2, factored_radius=3,
                                       chunk_planes=chunk), g)
        s.set_slowness(slow)
        s.set_source(src)
        s.solve()
        fields[chunk] = s.get_travel_time_field()
        st = s.get_stats()
        # one small grid: the shortest chunks; an explicit depth larger
        # than the grid: whole columns
        assert st.chunk_planes == (8 if chunk == 0 else 33)
        assert st.chains == 1
    assert np.array_equal(fields[0], fields[10 ** 6])
    og = oracle.geometry3d(41, 37, 33, 12.0)
    assert np.array_equal(fields[0], oracle.solve3d(og, slow, src, 2, 3))


def test_batch_autotune_keeps_fields(ekx):
    problems = _problems(ekx, 20, (33, 29, 27), seed=9)
    b = ekx.Batch3D()
    b.initialize(ekx.SolverOptions(number_of_sweeps=2, factored_radius=3),
                 problems)
    b.solve()
    before = [b.get_travel_time_field(i) for i in range(len(problems))]
    default = b.get_stats()
    chunk, chains = b.autotune()
    assert chunk in (27, 24, 16, 12, 8)
    assert chains in (1, 2, 4)
    with pytest.raises(RuntimeError):
        b.get_travel_time_field(0)          # the trials discard the field
    b.solve()
    st = b.get_stats()
    assert (st.chunk_planes, st.chains) == (chunk, chains)
    assert st.updates == default.updates
    for i in range(len(problems)):
        assert np.array_equal(b.get_travel_time_field(i), before[i])


@pytest.mark.parametrize("precision", ["FP32_POLISH", "FIXED32"])
def test_autotune_reduced_precision(ekx, precision):
    problems = _problems(ekx, 6, (25, 21, 30), seed=13)
    opts = ekx.SolverOptions(number_of_sweeps=2, factored_radius=3,
                             precision=getattr(ekx.Precision, precision))
    b = ekx.Batch3D()
    b.initialize(opts, problems)
    b.solve()
    before = [b.get_travel_time_field(i) for i in range(len(problems))]
    b.autotune()
    b.solve()
    for i in range(len(problems)):
        assert np.array_equal(b.get_travel_time_field(i), before[i])


def test_solver_autotune_needs_model_and_source(ekx):
    g = ekx.Geometry3D(17, 17, 17, 10.0)
    s = ekx.Solver3D()
    s.initialize(ekx.SolverOptions(), g)
    with pytest.raises(RuntimeError, match="model"):
        s.autotune()
    s.set_slowness(np.full(16 ** 3, 1 / 3000.0))
    with pytest.raises(RuntimeError, match="Source"):
        s.autotune()
    s.set_source((80.0, 80.0, 80.0))
    chunk, chains = s.autotune()
    assert chains == 1 and 8 <= chunk <= 17
    s.solve()
    assert s.get_stats().chunk_planes == chunk
