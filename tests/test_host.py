"""Host-side logic of the drop-in models (no GPU needed)."""

from __future__ import annotations

import numpy as np
import pytest

from somax_amd import BaroclinicQG, BarotropicQG, ModalTransform, SomaxModel, StratificationProfile
from somax_amd import testcases as tc


def test_models_are_somax_models():
    m = BaroclinicQG.create(nx=16, ny=16, n_layers=2, H=(500.0, 4500.0), g_prime=(9.81, 0.025))
    assert isinstance(m, SomaxModel)
    assert m.consts.n_layers == 2 and m.helmholtz_lambdas.shape == (2,)
    assert isinstance(BarotropicQG.create(nx=8, ny=8), SomaxModel)


def test_create_default_three_layers():
    assert BaroclinicQG.create().consts.n_layers == 3


def test_create_length_mismatch_raises():
    """baroclinic.py:284-289: same ValueError text."""
    with pytest.raises(ValueError, match="must all be equal"):
        BaroclinicQG.create(n_layers=3, H=(1.0, 2.0), g_prime=(9.81, 0.02))


def test_create_from_stratification():
    strat = StratificationProfile.from_N2_constant(N2=1e-5, depth=4000.0, n_layers=4)
    m = BaroclinicQG.create(nx=16, ny=16, stratification=strat)
    assert m.consts.n_layers == 4


def test_modal_roundtrip_and_inverse():
    t = ModalTransform.from_physics(H=(500.0, 4500.0), g_prime=(9.81, 0.025), f0=1e-4)
    x = np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]])
    np.testing.assert_allclose(t.to_layer(t.to_modal(x)), x, atol=1e-12)
    np.testing.assert_allclose(t.Cl2m @ t.Cm2l, np.eye(2), atol=1e-12)
    assert np.all(t.rossby_radii > 0)
    assert np.all(np.diff(t.eigenvalues) > 0)


def test_lambda_choice():
    m_c = BaroclinicQG.create(nx=8, ny=8)
    m_r = BaroclinicQG.create(nx=8, ny=8, helmholtz_eigenvalues="reference")
    f0 = m_c.consts.f0
    np.testing.assert_allclose(m_c.helmholtz_lambdas, f0**2 * m_c.modal.eigenvalues)
    np.testing.assert_allclose(m_r.helmholtz_lambdas, f0**2 * m_c.modal.reference_eigenvalues)


def test_forcing_profiles():
    m = BaroclinicQG.create(nx=16, ny=16, Lx=4e6, Ly=4e6)
    y = np.arange(18) * (4e6 / 16)
    np.testing.assert_allclose(m.beta_y[:, 3], m.consts.beta * (y - 2e6))
    np.testing.assert_allclose(m.wind_forcing[:, 0], -np.sin(2 * np.pi * y / 4e6))


def test_bad_flags_raise():
    with pytest.raises(ValueError):
        BaroclinicQG.create(nx=8, ny=8, advection="upwind")
    with pytest.raises(ValueError):
        BaroclinicQG.create(nx=8, ny=8, precision="fp16")


def test_random_streamfunction_properties():
    psi = tc.random_streamfunction(32, 32, tc.BASIN, tc.BASIN, 2, seed=1)
    assert psi.shape == (2, 34, 34)
    assert np.all(psi[:, 0] == 0) and np.all(psi[:, :, -1] == 0)
    dx = tc.BASIN / 32
    u = -(psi[:, 2:, 1:-1] - psi[:, 1:-1, 1:-1]) / dx
    v = (psi[:, 1:-1, 2:] - psi[:, 1:-1, 1:-1]) / dx
    np.testing.assert_allclose(np.sqrt(np.mean(u**2 + v**2, axis=(1, 2))), 0.05, rtol=1e-12)
    assert not np.allclose(psi[0], psi[1])
    np.testing.assert_array_equal(psi, tc.random_streamfunction(32, 32, tc.BASIN, tc.BASIN, 2, seed=1))


def test_config_parameters():
    m = tc.config_baroclinic(1024)
    assert m.params.biharmonic_viscosity == pytest.approx(2e11 * (1 / 8) ** 4)
    assert m.params.wind_amplitude == pytest.approx(9.8175e-11, rel=1e-4)
    b = tc.config_barotropic(64)
    assert b.params.wind_amplitude == pytest.approx(2.4544e-14, rel=1e-4)


def test_swm_create_and_validation():
    """tests/models/test_swm_multilayer.py:55-83 on the host side."""
    from somax_amd import MultilayerShallowWater2D

    assert MultilayerShallowWater2D.create().consts.n_layers == 3
    m = MultilayerShallowWater2D.create(nx=16, ny=16, n_layers=2, H=(500.0, 500.0), g_prime=(9.81, 0.02))
    assert isinstance(m, SomaxModel) and m.strat.H.shape == (2,)
    strat = StratificationProfile.from_N2_constant(N2=1e-5, depth=4000.0, n_layers=4)
    assert MultilayerShallowWater2D.create(nx=16, ny=16, stratification=strat).consts.n_layers == 4
    with pytest.raises(ValueError, match="must all be equal"):
        MultilayerShallowWater2D.create(nx=16, ny=16, n_layers=3, H=(500.0, 500.0), g_prime=(9.81, 0.02))
    with pytest.raises(ValueError):
        MultilayerShallowWater2D.create(nx=16, ny=16, bc="open")
    y = np.arange(18) * (4e6 / 16)
    np.testing.assert_allclose(m.wind_stress_x[:, 0], -np.cos(2 * np.pi * y / 4e6))
    np.testing.assert_allclose(m.f_field[:, 0], 9.375e-5 + 1.754e-11 * (y - 2e6))


def test_baroclinic_jet_balance():
    """The jet's layer thicknesses satisfy the discrete-in-y geostrophic
    balance f u_0 = -g'_0 dh_0/dy to the accuracy of the Gaussian's sampling."""
    from somax_amd import testcases as tc

    model, (h, u, v) = tc.baroclinic_jet_swm(512, noise=0.0)
    dy = model.grid.dy
    dh0 = np.gradient(h[0, :, 5], dy)
    dh1 = np.gradient(h[1, :, 5], dy)
    np.testing.assert_allclose(tc.JET_F0 * u[0, :, 5], -9.81 * dh0, atol=1e-3 * tc.JET_F0 * 0.5)
    np.testing.assert_allclose(9.81 * dh0 + 0.02 * dh1, 0.0, atol=1e-12)
    assert h.min() > 0 and 20.0 < tc.jet_dt(model) < 500.0
    # the host-only constructors the long-run fixture uses give the same state and step
    h2, u2, v2 = tc.baroclinic_jet_fields(512, noise=0.0)
    assert np.array_equal(h, h2) and np.array_equal(u, u2) and np.array_equal(v, v2)
    assert tc.cfl_dt(tc.JET_H, tc.JET_GP, model.grid.dx, model.grid.dy) == tc.jet_dt(model)


def test_solver_names():
    """integrate's solver argument: names or solver objects (by class name)."""
    from somax_amd.core import solver_name

    class Tsit5:  # stands for diffrax.Tsit5, which is not installed here
        pass

    assert solver_name("tsit5") == "tsit5" and solver_name(Tsit5()) == "tsit5"
    assert solver_name("SSP-RK3") == "ssprk3" and solver_name("ssprk3") == "ssprk3"
    with pytest.raises(NotImplementedError):
        solver_name("dopri5")


def test_save_times_are_checked_like_diffrax():
    """integrate's time span: positive dt, save times at or after t0 and not
    decreasing (a negative dt or a backward span raises instead of taking
    one step over the whole span or returning the state unchanged)."""
    import pytest

    from somax_amd.core import save_times

    assert save_times(0.0, 3600.0, 600.0, None) == [3600.0]
    assert save_times(0.0, 1.0, 0.5, [0.0, 0.5, 0.5, 2.0]) == [0.0, 0.5, 0.5, 2.0]
    for t0, t1, dt, saveat in ((0.0, 3600.0, -600.0, None), (0.0, 3600.0, 0.0, None), (3600.0, 0.0, 600.0, None),
                               (0.0, 1.0, 0.5, [1.0, 0.5]), (0.0, 1.0, 0.5, [])):
        with pytest.raises(ValueError):
            save_times(t0, t1, dt, saveat)
