"""Seeded random sweep over grid shapes (including sizes that are not whole
tiles and grids of 3 points per axis), source positions, spacings, chunk
depths, chain counts and precision modes: the column kernel (tiled fields)
against the hyperplane kernel (natural fields) bit for bit, and in FP64
against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = 12


def _case(rng, ekx):
    n = int(rng.integers(1, 6))
    precision = str(rng.choice(["FP64", "FP64", "FIXED32", "FP32_POLISH"]))
    chunk = int(rng.choice([0, 0, 3, 5, 8, 13, 40]))
    chains = int(rng.choice([0, 1, 2, 3]))
    problems = []
    uniform = bool(rng.integers(0, 2))
    for _ in range(n):
        dims = tuple(int(v) for v in rng.integers(3, 36, size=3))
        d = (10.0, 10.0, 10.0) if uniform else tuple(
            float(v) for v in rng.uniform(5.0, 20.0, size=3))
        g = ekx.Geometry3D(*dims, *d)
        ncell = (dims[0] - 1) * (dims[1] - 1) * (dims[2] - 1)
        slow = 1.0 / rng.uniform(1500.0, 6000.0, ncell)
        corner = rng.integers(0, 3)
        if corner == 0:
            src = (0.0, 0.0, 0.0)
        elif corner == 1:
            src = tuple((m - 1) * dd for m, dd in zip(dims, d))
        else:
            src = tuple(float(rng.uniform(0, (m - 1) * dd))
                        for m, dd in zip(dims, d))
        problems.append((g, slow, src, dims, d))
    return problems, precision, chunk, chains


@pytest.mark.parametrize("seed", range(CASES))
def test_random_shapes_tiled_equals_natural(ekx, oracle, seed):
    rng = np.random.default_rng(1000 + seed)
    problems, precision, chunk, chains = _case(rng, ekx)
    n_sweeps = int(rng.integers(1, 3))
    radius = int(rng.choice([-1, 2, 3]))

    def run(kernel):
        b = ekx.Batch3D()
        b.initialize(ekx.SolverOptions(
            number_of_sweeps=n_sweeps, factored_radius=radius,
            chunk_planes=chunk, chains=chains, kernel=kernel,
            precision=getattr(ekx.Precision, precision)),
            [(g, s, src) for g, s, src, _, _ in problems])
        b.solve()
        return [b.get_travel_time_field(i) for i in range(len(problems))]
    tiled = run(ekx.Kernel.COLUMN)
    natural = run(ekx.Kernel.HYPERPLANE)
    for i, (a, b) in enumerate(zip(tiled, natural)):
        assert np.array_equal(a, b), (seed, i, precision, chunk, chains)
    if precision == "FP64":
        g, slow, src, dims, d = problems[0]
        og = oracle.geometry3d(*dims, *d)
        assert np.array_equal(tiled[0], oracle.solve3d(og, slow, src,
                                                       n_sweeps, radius))
