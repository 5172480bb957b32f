"""GPU parity of the quantities derived from a solved field: travel times at
3D stations (trilinear, analytic in the source cell) and the 2D/3D
travel-time gradients (src/gradient.hpp), against the oracle on the same
field.  FP64 fields give bit-identical results."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _model3d(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    return 1.0 / rng.uniform(2000.0, 5000.0, (nx - 1) * (ny - 1) * (nz - 1))


def _solver3d(ekx, geom, slow, src, precision=None, n_sweeps=2, radius=3):
    kw = {} if precision is None else {"precision": precision}
    s = ekx.Solver3D()
    s.initialize(ekx.SolverOptions(number_of_sweeps=n_sweeps,
                                   factored_radius=radius, **kw), geom)
    s.set_slowness(slow)
    s.set_source(src)
    return s


def _stations(rng, n, extent, origin=(0.0, 0.0, 0.0)):
    pts = rng.random((n, 3)) * np.asarray(extent) + np.asarray(origin)
    return pts


def test_stations3d_match_oracle(ekx, oracle):
    nx, ny, nz, h = 23, 19, 17, 10.0
    origin = (100.0, -40.0, 5.0)
    geom = ekx.Geometry3D(nx, ny, nz, h, x0=origin[0], y0=origin[1],
                          z0=origin[2])
    og = oracle.geometry3d(nx, ny, nz, h, h, h, *origin)
    slow = _model3d(nx, ny, nz, 1)
    src = (origin[0] + 93.0, origin[1] + 71.0, origin[2] + 44.0)
    rng = np.random.default_rng(9)
    extent = ((nx - 1) * h, (ny - 1) * h, (nz - 1) * h)
    sta = _stations(rng, 300, extent, origin)
    # a station in the source cell, one at the source, nodes, the far corner
    extra = [(src[0] + 2.0, src[1] - 1.0, src[2] + 3.0), src,
             (origin[0], origin[1], origin[2]),
             (origin[0] + extent[0], origin[1] + extent[1],
              origin[2] + extent[2]),
             (origin[0] + 50.0, origin[1] + 80.0, origin[2] + 160.0)]
    sta = np.vstack([sta, extra])
    s = _solver3d(ekx, geom, slow, src)
    s.set_stations(sta)
    assert s.get_number_of_stations() == len(sta)
    with pytest.raises(RuntimeError):
        s.get_travel_times_at_stations()
    s.solve()
    got = s.get_travel_times_at_stations()
    field = s.get_travel_time_field()
    ref = oracle.solve3d(og, slow, src, 2, 3)
    assert np.array_equal(field, ref)
    want = np.array([oracle.station_time3d(og, slow, ref, src, r)
                     for r in sta])
    np.testing.assert_array_equal(got, want)
    assert got[-4] == 0.0                           # at the source
    # stations set after a solve are interpolated at once
    s.set_stations(sta[:7])
    np.testing.assert_array_equal(s.get_travel_times_at_stations(),
                                  want[:7])
    with pytest.raises(ValueError, match="station position"):
        s.set_stations([(origin[0] - 1.0, origin[1], origin[2])])


@pytest.mark.parametrize("precision", ["FP32", "FIXED32", "FP32_POLISH"])
def test_stations3d_reduced_precision(ekx, oracle, precision):
    """The interpolation reads the stored field of every precision."""
    n, h = 33, 10.0
    geom = ekx.Geometry3D(n, n, n, h)
    slow = _model3d(n, n, n, 4)
    src = (161.0, 143.0, 97.0)
    sta = _stations(np.random.default_rng(2), 64, ((n - 1) * h,) * 3)
    ref = _solver3d(ekx, geom, slow, src, n_sweeps=3)
    ref.set_stations(sta)
    ref.solve()
    s = _solver3d(ekx, geom, slow, src, getattr(ekx.Precision, precision),
                  n_sweeps=3)
    s.set_stations(sta)
    s.solve()
    got = s.get_travel_times_at_stations()
    field = s.get_travel_time_field()
    og = oracle.geometry3d(n, n, n, h)
    want = np.array([oracle.station_time3d(og, slow, field, src, r)
                     for r in sta])
    # the same interpolation of the field the solver returns; float fields
    # are read as float and widened, as the oracle sees them
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(got, ref.get_travel_times_at_stations(),
                               rtol=0, atol=2e-3)


def test_batch_station_table(ekx, oracle):
    """Station-reciprocal tables: problems x stations after one solve."""
    n, h = 25, 10.0
    geom = ekx.Geometry3D(n, n, n, h)
    og = oracle.geometry3d(n, n, n, h)
    slow = _model3d(n, n, n, 7)
    rng = np.random.default_rng(12)
    sources = [tuple(p) for p in _stations(rng, 5, ((n - 1) * h,) * 3)]
    b = ekx.Batch3D()
    b.initialize(ekx.SolverOptions(number_of_sweeps=2, factored_radius=3),
                 [(geom, slow, src) for src in sources])
    sta = _stations(rng, 40, ((n - 1) * h,) * 3)
    b.set_stations(sta)
    b.solve()
    table = b.get_station_table()
    assert table.shape == (5, 40)
    for k, src in enumerate(sources):
        ref = oracle.solve3d(og, slow, src, 2, 3)
        want = [oracle.station_time3d(og, slow, ref, src, r) for r in sta]
        np.testing.assert_array_equal(table[k], want)
        np.testing.assert_array_equal(b.get_travel_times_at_stations(k),
                                      want)


@pytest.mark.parametrize("src", [(93.0, 71.0, 44.0), (90.0, 70.0, 40.0),
                                 (93.0, 70.0, 40.0), (220.0, 180.0, 160.0)])
def test_gradient3d_matches_oracle(ekx, oracle, src):
    """Off-node, on-node, off in x only, and on the far corner of the
    model (the source cell is the last one)."""
    nx, ny, nz, h = 23, 19, 17, 10.0
    geom = ekx.Geometry3D(nx, ny, nz, h)
    og = oracle.geometry3d(nx, ny, nz, h)
    slow = _model3d(nx, ny, nz, 3)
    s = _solver3d(ekx, geom, slow, src)
    with pytest.raises(RuntimeError):
        s.compute_travel_time_gradient_field()
    s.solve()
    assert not s.have_travel_time_gradient_field()
    with pytest.raises(RuntimeError):
        s.get_travel_time_gradient_field()
    s.compute_travel_time_gradient_field()
    assert s.have_travel_time_gradient_field()
    got = s.get_travel_time_gradient_field()
    ref = oracle.solve3d(og, slow, src, 2, 3)
    want = oracle.gradient3d(og, slow, ref, src)
    np.testing.assert_array_equal(got, want)
    assert s.get_travel_time_gradient_field_device()
    # a new solve invalidates the gradient
    s.set_source((55.0, 66.0, 77.0))
    s.solve()
    assert not s.have_travel_time_gradient_field()
    s.compute_travel_time_gradient_field()
    ref = oracle.solve3d(og, slow, (55.0, 66.0, 77.0), 2, 3)
    np.testing.assert_array_equal(
        s.get_travel_time_gradient_field(),
        oracle.gradient3d(og, slow, ref, (55.0, 66.0, 77.0)))


def test_gradient3d_nonuniform_and_batch(ekx, oracle):
    nx, ny, nz = 21, 18, 15
    dx, dy, dz = 10.0, 12.0, 8.0
    geom = ekx.Geometry3D(nx, ny, nz, dx, dy, dz)
    og = oracle.geometry3d(nx, ny, nz, dx, dy, dz)
    slow = _model3d(nx, ny, nz, 5)
    sources = [(33.0, 45.0, 17.0), (150.0, 100.0, 60.0), (0.0, 0.0, 0.0)]
    b = ekx.Batch3D()
    b.initialize(ekx.SolverOptions(number_of_sweeps=2, factored_radius=2),
                 [(geom, slow, src) for src in sources])
    b.solve()
    for k in (2, 0, 1):
        ref = oracle.solve3d(og, slow, sources[k], 2, 2)
        np.testing.assert_array_equal(b.get_travel_time_gradient_field(k),
                                      oracle.gradient3d(og, slow, ref,
                                                        sources[k]))


def test_gradient3d_fixed32(ekx, oracle):
    """A scaled-integer field: the gradient of the decoded field."""
    n, h = 29, 10.0
    geom = ekx.Geometry3D(n, n, n, h)
    og = oracle.geometry3d(n, n, n, h)
    slow = _model3d(n, n, n, 8)
    src = (131.0, 97.0, 155.0)
    s = _solver3d(ekx, geom, slow, src, ekx.Precision.FIXED32, n_sweeps=3)
    s.solve()
    s.compute_travel_time_gradient_field()
    field = s.get_travel_time_field()
    np.testing.assert_array_equal(s.get_travel_time_gradient_field(),
                                  oracle.gradient3d(og, slow, field, src))


def test_gradient2d_matches_oracle(ekx, oracle):
    nx, nz, dx, dz = 41, 33, 10.0, 12.0
    rng = np.random.default_rng(21)
    slow = 1.0 / rng.uniform(1500.0, 4500.0, (nx - 1) * (nz - 1))
    og = oracle.geometry2d(nx, nz, dx, dz)
    for src in [(187.0, 211.0), (190.0, 204.0), (400.0, 384.0)]:
        s = ekx.Solver2D()
        s.initialize(ekx.SolverOptions(number_of_sweeps=2, factored_radius=3),
                     ekx.Geometry2D(nx, nz, dx, dz))
        s.set_slowness(slow)
        s.set_source(src)
        with pytest.raises(RuntimeError):
            s.compute_travel_time_gradient_field()
        s.solve()
        assert not s.have_travel_time_gradient_field()
        s.compute_travel_time_gradient_field()
        assert s.have_travel_time_gradient_field()
        ref = oracle.solve2d(og, slow, src, 2, 3)
        np.testing.assert_array_equal(s.get_travel_time_gradient_field(),
                                      oracle.gradient2d(og, slow, ref, src))
