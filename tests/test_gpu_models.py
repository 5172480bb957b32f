"""GPU parity of the generated layered Gaussian velocity models: the device
generator against the oracle's (same hash, filter and normalisation; the
device log/cos and the reduction order differ from the host's by
rounding only, so agreement is to 1e-12 relative), and a solver fed the
device-generated model directly against one fed the same model from host
memory (bit-identical)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,d,layers,amp,l,seed", [
    ((23, 17, 11), (10.0, 10.0, 10.0), [3000.0, 4500.0], 0.05, 60.0, 1),
    ((20, 9, 13), (5.0, 12.0, 7.0), [2500.0, 3500.0, 5000.0], 0.1, 40.0, 2),
    ((16, 8, 4), (2.0, 3.0, 1.0), [4000.0], 0.02, 30.0, 3),
    ((12, 10, 9), (10.0, 10.0, 10.0), [3000.0, 6000.0], 0.05, 4.0, 4),
    ((33, 31, 29), (10.0, 10.0, 10.0), [3000.0] * 5, 0.0, 80.0, 5),
    ((64, 48, 40), (25.0, 25.0, 25.0), [5000.0, 6000.0, 7000.0], 0.08,
     250.0, (1 << 64) - 3),
])
def test_generated_model_matches_oracle(ekx, oracle, dims, d, layers, amp, l,
                                        seed):
    nx, ny, nz = dims
    geom = ekx.Geometry3D(nx, ny, nz, *d)
    og = oracle.geometry3d(nx, ny, nz, *d)
    got = ekx.LayeredGaussianModel(layers, amp, l, seed).generate(geom)
    want = oracle.layered_gaussian3d(og, layers, amp, l, seed)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=0)
    # the same seed gives the same field; another seed another one
    again = ekx.LayeredGaussianModel(layers, amp, l, seed).generate(geom)
    np.testing.assert_array_equal(got, again)
    if amp > 0:
        other = ekx.LayeredGaussianModel(layers, amp, l,
                                         seed ^ 1).generate(geom)
        assert not np.array_equal(got, other)


def test_solver_with_device_model_equals_host_model(ekx):
    geom = ekx.Geometry3D(41, 37, 33, 20.0)
    model = ekx.LayeredGaussianModel([3000.0, 4000.0, 5000.0], 0.06, 120.0,
                                     11)
    v = model.generate(geom)
    src = (412.0, 333.0, 217.0)
    opts = ekx.SolverOptions(number_of_sweeps=2, factored_radius=3)
    a = ekx.Solver3D()
    a.initialize(opts, geom)
    a.set_layered_gaussian_model(model)
    a.set_source(src)
    a.solve()
    b = ekx.Solver3D()
    b.initialize(opts, geom)
    b.set_nodal_velocities(v)
    b.set_source(src)
    b.solve()
    np.testing.assert_array_equal(a.get_travel_time_field(), b.get_travel_time_field())


def test_generated_model_errors(ekx):
    geom = ekx.Geometry3D(12, 10, 9, 10.0)
    with pytest.raises(ValueError, match="Non-positive velocities"):
        ekx.LayeredGaussianModel([3000.0], 2.0, 40.0, 1).generate(geom)
    with pytest.raises(ValueError, match="Non-positive velocities"):
        ekx.LayeredGaussianModel([3000.0, -1.0], 0.0, 40.0, 1).generate(geom)
    with pytest.raises(ValueError, match="at least one layer"):
        ekx.LayeredGaussianModel([], 0.0, 40.0, 1).generate(geom)
    with pytest.raises(ValueError, match="amplitude"):
        ekx.LayeredGaussianModel([3000.0], -0.1, 40.0, 1).generate(geom)
    s = ekx.Solver3D()
    s.initialize(ekx.SolverOptions(), geom)
    with pytest.raises(ValueError, match="Non-positive velocities"):
        s.set_layered_gaussian_model(
            ekx.LayeredGaussianModel([3000.0], 2.0, 40.0, 1))
    assert not s.have_velocity_model()
    s.set_layered_gaussian_model(ekx.LayeredGaussianModel([3000.0], 0.01,
                                                          40.0, 1))
    assert s.have_velocity_model()
