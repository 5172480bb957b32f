"""Grouped refinement sweeps (option column_group = 1): batches whose
problems share one model and one geometry run G problems of the same column
per block with one shared slowness ring.  The fields must not change:
bit-identical to the oracle in FP64 and to the ungrouped launches in every
precision."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NX, NY, NZ, H = 29, 27, 23, 15.0


def _problems(ekx, n, seed, geom=None):
    rng = np.random.default_rng(seed)
    g = geom or ekx.Geometry3D(NX, NY, NZ, H)
    slow = 1.0 / rng.uniform(1500.0, 5000.0,
                             (g.nx - 1) * (g.ny - 1) * (g.nz - 1))
    ext = ((g.nx - 1) * g.dx, (g.ny - 1) * g.dy, (g.nz - 1) * g.dz)
    srcs = [tuple(float(v) for v in rng.random(3) * ext) for _ in range(n)]
    # corners and faces: factored boxes that span column boundaries
    srcs[0] = (0.0, 0.0, 0.0)
    srcs[1] = ext
    srcs[2] = (7.5 * g.dx, 8.2 * g.dy, 0.0)
    return [(g, slow, s) for s in srcs]


def _run(ekx, problems, **kw):
    kw.setdefault("column_group", 1)
    b = ekx.Batch3D()
    b.initialize(ekx.SolverOptions(factored_radius=3, **kw), problems)
    b.solve()
    fields = [b.get_travel_time_field(i) for i in range(len(problems))]
    return fields, b.get_stats()


@pytest.mark.parametrize("n,chains", [(13, 1), (14, 2), (6, 1)])
def test_grouped_fp64_matches_oracle(ekx, oracle, n, chains):
    """13 problems: two full groups of 6 and one padded group; 14 problems
    in 2 chains of 7; exactly one group."""
    problems = _problems(ekx, n, 100 + n)
    got, st = _run(ekx, problems, number_of_sweeps=2, chains=chains)
    assert st.group_waves == 6
    off, st0 = _run(ekx, problems, number_of_sweeps=2, chains=chains,
                    column_group=0)
    assert st0.group_waves == 0
    assert st.updates == st0.updates
    og = oracle.geometry3d(NX, NY, NZ, H)
    for i, (g, slow, src) in enumerate(problems):
        assert np.array_equal(got[i], off[i]), i
        if i < 5:
            ref = oracle.solve3d(og, slow, src, 2, 3)
            assert np.array_equal(got[i], ref), i


@pytest.mark.parametrize("precision", ["FP32", "FIXED32", "FP32_POLISH",
                                       "FIXED32_POLISH"])
def test_grouped_reduced_precision_is_unchanged(ekx, precision):
    """17 problems in 2 chains: every chain holds a full group of 8."""
    problems = _problems(ekx, 17, 7)
    prec = getattr(ekx.Precision, precision)
    got, st = _run(ekx, problems, number_of_sweeps=3, precision=prec,
                   chains=2)
    assert st.group_waves in (6, 8)
    off, _ = _run(ekx, problems, number_of_sweeps=3, precision=prec,
                  chains=2, column_group=0)
    for a, b in zip(got, off):
        assert np.array_equal(a, b)


def test_grouped_nonuniform_and_chunked(ekx, oracle):
    g = ekx.Geometry3D(NX, NY, NZ, 12.0, 15.0, 9.0)
    problems = _problems(ekx, 8, 31, geom=g)
    got, st = _run(ekx, problems, number_of_sweeps=2, chunk_planes=7)
    assert st.group_waves == 6
    og = oracle.geometry3d(NX, NY, NZ, 12.0, 15.0, 9.0)
    for i in (0, 3, 7):
        _, slow, src = problems[i]
        assert np.array_equal(got[i], oracle.solve3d(og, slow, src, 2, 3))


def test_grouped_early_exit(ekx):
    """The convergence monitor sees the grouped launches: early exit stops
    at the same iteration and gives the same fields as ungrouped."""
    problems = _problems(ekx, 12, 5)
    kw = dict(number_of_sweeps=10, early_exit=True,
              convergence_tolerance=1e-9)
    got, st = _run(ekx, problems, **kw)
    off, st0 = _run(ekx, problems, column_group=0, **kw)
    assert st.group_waves == 6
    assert st.iterations_run == st0.iterations_run
    for a, b in zip(got, off):
        assert np.array_equal(a, b)


def test_not_grouped_when_models_differ(ekx):
    p = _problems(ekx, 7, 1)
    q = _problems(ekx, 7, 2)
    _, st = _run(ekx, p + q, number_of_sweeps=1)
    assert st.group_waves == 0
    g2 = ekx.Geometry3D(NX, NY, NZ + 1, H)
    r = [(g2, np.resize(p[0][1], (NX - 1) * (NY - 1) * NZ), s)
         for _, _, s in p]
    _, st = _run(ekx, p + r, number_of_sweeps=1)
    assert st.group_waves == 0
