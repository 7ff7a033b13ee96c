"""Pins of the FP64 shallow-water oracle (oracle/swm.py), no GPU.

* The reference's own qualitative tests (tests/models/test_swm_multilayer.py:94-175):
  rest state has zero tendency, wind acts on the top layer only, a height
  bump couples into the lower layer's momentum, wall BC values.
* Textbook C-grid theory: the eigenvalues of the tendency's Jacobian about
  rest on a periodic f-plane are 0 (geostrophic modes) and +-i*omega with
  omega^2 = f^2 cos^2(k dx/2) cos^2(l dy/2) + c^2 [(2/dx sin(k dx/2))^2 +
  (2/dy sin(l dy/2))^2] (Arakawa & Lamb 1977, grid C), where c^2 runs over
  the eigenvalues of diag(H) M for the pressure matrix M of the cumulative
  g' h coupling.  A wrong sign, a transposed average or an index shift in
  any of the linear operators moves these eigenvalues.
* Flux-form mass conservation for every boundary condition, and the sign of
  the Coriolis term on uniform flow.
* The energy-enstrophy-conserving vorticity flux (extension): on a doubly
  periodic grid its work sum U F_u + V F_v vanishes for any PV and fluxes,
  and its potential-enstrophy tendency sum q curl(F) vanishes for any PV and
  a non-divergent mass flux, both to round-off; the reference's average form
  fails the first, which makes the check discriminating.  Uniform PV gives the
  average form, and the linear dispersion relation is unchanged.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import swm


def _setup(bc="periodic", nl=2, n=16, **kw):
    H = (500.0, 500.0, 1000.0)[:nl]
    gp = (9.81, 0.02, 0.01)[:nl]
    args = dict(nx=n, ny=n, Lx=1e6, Ly=1e6, f0=1e-4, beta=0.0, H=H, g_prime=gp, bc=bc)
    args.update(kw)
    return swm.setup(**args)


def _rest(s):
    h = np.ones(s.shape) * s.H[:, None, None]
    return h, np.zeros(s.shape), np.zeros(s.shape)


@pytest.mark.parametrize("bc", ["periodic", "wall", "channel"])
def test_rest_state_zero_tendency(bc):
    s = _setup(bc)
    dh, du, dv = swm.rhs(s, *swm.apply_bc(s, *_rest(s)))
    assert max(np.abs(dh).max(), np.abs(du).max(), np.abs(dv).max()) < 1e-10


def test_wind_top_layer_only():
    s = _setup(wind_amplitude=1e-5)
    dh, du, dv = swm.rhs(s, *swm.apply_bc(s, *_rest(s)))
    assert np.abs(du[0]).max() > 0
    assert np.abs(du[-1]).max() < 1e-15


def test_pressure_couples_layers():
    s = _setup()
    h, u, v = _rest(s)
    h[0, s.Ny // 2, s.Nx // 2] += 1.0
    dh, du, dv = swm.rhs(s, *swm.apply_bc(s, h, u, v))
    assert np.abs(du[1]).max() > 0


def test_wall_bc_values():
    s = _setup("wall")
    h, u, v = _rest(s)
    u[:] = 1.0
    v[:] = 1.0
    hb, ub, vb = swm.apply_bc(s, h, u, v)
    assert np.all(ub[:, :, 0] == 0) and np.all(ub[:, :, -1] == 0) and np.all(ub[:, :, -2] == 0)
    assert np.all(vb[:, 0, :] == 0) and np.all(vb[:, -1, :] == 0) and np.all(vb[:, -2, :] == 0)
    assert np.all(ub[:, 0, 1:-2] == 1.0)  # free slip copies the first interior row


def _jacobian_eigs(s):
    """Eigenvalues of d(rhs)/d(interior state) about rest, by central differences."""
    h0, u0, v0 = _rest(s)
    ny, nx = s.ny, s.nx
    n1 = s.nl * ny * nx
    idx = [(f, l, j, i) for f in range(3) for l in range(s.nl) for j in range(1, ny + 1) for i in range(1, nx + 1)]
    J = np.zeros((3 * n1, 3 * n1))
    eps = 1e-6
    for c, (f, l, j, i) in enumerate(idx):
        cols = []
        for sign in (1.0, -1.0):
            st = [h0.copy(), u0.copy(), v0.copy()]
            st[f][l, j, i] += sign * eps * (s.H[l] if f == 0 else 1.0)
            d = swm.rhs(s, *swm.apply_bc(s, *st))
            cols.append(np.concatenate([x[:, 1:-1, 1:-1].ravel() for x in d]))
        col = (cols[0] - cols[1]) / (2 * eps * (s.H[l] if f == 0 else 1.0))
        J[:, c] = col
    return np.linalg.eigvals(J)


@pytest.mark.parametrize("flux", ["average", "energy_enstrophy"])
@pytest.mark.parametrize("nl", [1, 2])
def test_linear_dispersion_matches_c_grid_theory(nl, flux):
    nx = ny = 6
    s = _setup("periodic", nl=nl, n=nx, Lx=6e5, Ly=6e5, vorticity_flux=flux)
    lam = _jacobian_eigs(s)
    assert np.abs(lam.real).max() < 1e-9 * np.abs(lam).max()
    # pressure matrix of the cumulative coupling: P_k = sum_{j<=k} g'_j h_j
    M = np.tril(np.ones((nl, nl))) * s.g_prime[None, :]
    c2 = np.sort(np.linalg.eigvals(np.diag(s.H) @ M).real)
    expect = []
    for kk in range(nx):
        for ll in range(ny):
            k = 2 * np.pi * kk / s.Lx
            l = 2 * np.pi * ll / s.Ly
            C = np.cos(k * s.dx / 2) * np.cos(l * s.dy / 2)
            S2 = (2 / s.dx * np.sin(k * s.dx / 2)) ** 2 + (2 / s.dy * np.sin(l * s.dy / 2)) ** 2
            for c in c2:
                w = np.sqrt(1e-4**2 * C**2 + c * S2)
                expect += [w, -w]
            expect += [0.0] * nl  # geostrophic modes
    got = np.sort(lam.imag)
    np.testing.assert_allclose(got, np.sort(expect), atol=1e-7 * np.abs(expect).max())


@pytest.mark.parametrize("bc", ["periodic", "wall", "channel"])
def test_mass_conserved(bc):
    s = _setup(bc, nl=2, n=12)
    rng = np.random.default_rng(3)
    h, u, v = _rest(s)
    h = h + rng.standard_normal(h.shape)
    u = 0.1 * rng.standard_normal(u.shape)
    v = 0.1 * rng.standard_normal(v.shape)
    dh, _, _ = swm.rhs(s, *swm.apply_bc(s, h, u, v))
    scale = np.abs(dh).max()
    assert np.abs(dh[:, 1:-1, 1:-1].sum(axis=(-2, -1))).max() < 1e-12 * scale * dh[0].size


def test_coriolis_turns_uniform_flow_right():
    """Uniform eastward flow on an f-plane: dv/dt = -f U, du/dt = dh/dt = 0."""
    s = _setup("periodic", nl=1, n=8)
    h, u, v = _rest(s)
    u[:] = 0.3
    dh, du, dv = swm.rhs(s, *swm.apply_bc(s, h, u, v))
    np.testing.assert_allclose(dv[0, 1:-1, 1:-1], -1e-4 * 0.3, rtol=1e-12)
    assert np.abs(du[0, 1:-1, 1:-1]).max() < 1e-15 and np.abs(dh).max() < 1e-9


def test_diagnose_rest_energy():
    s = _setup("wall")
    d = swm.diagnose(s, *swm.apply_bc(s, *_rest(s)))
    expect = 0.5 * s.g_prime * s.H**2 * s.nx * s.ny * s.dx * s.dy
    np.testing.assert_allclose(d["energy"], expect, rtol=1e-12)
    assert np.all(d["energy"] >= 0)


def test_nonlinear_single_layer_wind_not_divided():
    """nonlinear_2d.py:180-182: du/dt gains tau0 * tau_x (no 1/H0), so on a
    resting layer du = tau0 * (-cos(2 pi y / Ly)) in closed form."""
    s = swm.nonlinear_setup(nx=12, ny=10, H0=250.0, wind_amplitude=3e-6)
    h = np.full((1, s.Ny, s.Nx), 250.0)
    z = np.zeros_like(h)
    dh, du, dv = swm.rhs(s, *swm.apply_bc(s, h, z, z))
    y = np.arange(s.Ny) * s.dy
    want = 3e-6 * -np.cos(2 * np.pi * y / s.Ly)
    np.testing.assert_allclose(du[0, 1:-1, 1:-1], np.broadcast_to(want[1:-1, None], (s.ny, s.nx)), rtol=1e-12)
    assert np.abs(dh).max() < 1e-12 and np.abs(dv).max() < 1e-18


# ---------------------------------------------------------------------------
# energy-enstrophy-conserving vorticity flux (extension)
# ---------------------------------------------------------------------------
def _torus_fields(s, seed):
    """Random PV on X points and mass fluxes on U / V points of a doubly
    periodic grid (ghost ring = periodic images)."""
    rng = np.random.default_rng(seed)
    nl, ny, nx = s.nl, s.ny, s.nx

    def x_field():
        a = np.zeros(s.shape)
        a[:, 1:-1, 1:-1] = rng.standard_normal((nl, ny, nx))
        a[:, 0, :] = a[:, -2, :]
        a[:, :, 0] = a[:, :, -2]
        return a

    def t_field():
        a = np.zeros(s.shape)
        a[:, 1:-1, 1:-1] = rng.standard_normal((nl, ny, nx))
        return swm.enforce_periodic(a)

    return x_field, t_field


@pytest.mark.parametrize("flux,conserves", [("energy_enstrophy", True), ("average", False)])
def test_vorticity_flux_does_no_work(flux, conserves):
    s = _setup("periodic", nl=2, n=12, ny=10, Ly=1e6 * 10 / 12, vorticity_flux=flux)
    x_field, t_field = _torus_fields(s, 4)
    q, uh, vh = x_field(), t_field(), t_field()
    fu, fv = swm.vorticity_flux_terms(s, q, uh, vh)
    work = (uh[:, 1:-1, 1:-1] * fu).sum() + (vh[:, 1:-1, 1:-1] * fv).sum()
    scale = np.abs(uh[:, 1:-1, 1:-1] * fu).sum()
    assert (abs(work) < 1e-14 * scale) == conserves


@pytest.mark.parametrize("flux", ["energy_enstrophy", "average"])
def test_vorticity_flux_keeps_enstrophy_of_nondivergent_flow(flux):
    """sum_X q curl(F) = 0 when div(uh, vh) = 0 (then h_X is steady and the
    potential enstrophy sum h_X q^2 / 2 changes only through curl(F))."""
    s = _setup("periodic", nl=2, n=12, ny=10, Ly=1e6 * 10 / 12, vorticity_flux=flux)
    x_field, _ = _torus_fields(s, 5)
    q, psi = x_field(), x_field()
    uh = np.zeros(s.shape)
    vh = np.zeros(s.shape)
    uh[:, 1:, :] = -(psi[:, 1:, :] - psi[:, :-1, :]) / s.dy
    vh[:, :, 1:] = (psi[:, :, 1:] - psi[:, :, :-1]) / s.dx
    uh, vh = swm.enforce_periodic(uh), swm.enforce_periodic(vh)
    div = (uh[:, 1:-1, 1:-1] - uh[:, 1:-1, :-2]) / s.dx + (vh[:, 1:-1, 1:-1] - vh[:, :-2, 1:-1]) / s.dy
    assert np.abs(div).max() < 1e-12 * np.abs(uh).max() / s.dx
    fu, fv = swm.vorticity_flux_terms(s, q, uh, vh)
    FU, FV = np.zeros(s.shape), np.zeros(s.shape)
    FU[:, 1:-1, 1:-1], FV[:, 1:-1, 1:-1] = fu, fv
    FU, FV = swm.enforce_periodic(FU), swm.enforce_periodic(FV)
    curl = (FV[:, 1:-1, 2:] - FV[:, 1:-1, 1:-1]) / s.dx - (FU[:, 2:, 1:-1] - FU[:, 1:-1, 1:-1]) / s.dy
    dz = (q[:, 1:-1, 1:-1] * curl).sum()
    assert abs(dz) < 1e-14 * np.abs(q[:, 1:-1, 1:-1] * curl).sum()


def test_energy_enstrophy_flux_with_uniform_pv_is_the_average_form():
    s_een = _setup("periodic", nl=1, n=8, vorticity_flux="energy_enstrophy")
    s_avg = _setup("periodic", nl=1, n=8)
    _, t_field = _torus_fields(s_een, 6)
    q = np.full(s_een.shape, 2.5e-7)
    uh, vh = t_field(), t_field()
    for a, b in zip(swm.vorticity_flux_terms(s_een, q, uh, vh), swm.vorticity_flux_terms(s_avg, q, uh, vh)):
        np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-22)


@pytest.mark.parametrize("bc", ["periodic", "wall", "channel"])
def test_energy_enstrophy_flux_rest_and_mass(bc):
    s = _setup(bc, nl=2, n=12, vorticity_flux="energy_enstrophy")
    dh, du, dv = swm.rhs(s, *swm.apply_bc(s, *_rest(s)))
    assert max(np.abs(dh).max(), np.abs(du).max(), np.abs(dv).max()) < 1e-10
    rng = np.random.default_rng(8)
    h, u, v = _rest(s)
    h = h + rng.standard_normal(h.shape)
    u, v = 0.1 * rng.standard_normal(u.shape), 0.1 * rng.standard_normal(v.shape)
    dh, du, _ = swm.rhs(s, *swm.apply_bc(s, h, u, v))
    assert np.abs(dh[:, 1:-1, 1:-1].sum(axis=(-2, -1))).max() < 1e-12 * np.abs(dh).max() * dh[0].size
    s_avg = _setup(bc, nl=2, n=12)
    _, du_avg, _ = swm.rhs(s_avg, *swm.apply_bc(s_avg, h, u, v))
    assert np.abs(du - du_avg).max() > 1e-6 * np.abs(du).max()  # the two forms differ on a real flow
