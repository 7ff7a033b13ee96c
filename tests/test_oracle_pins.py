"""Pins of the FP64 oracle against the reference's published numbers, closed
forms and invariants.  CPU only.

Each test names the reference artefact or the mathematical fact it checks and
is chosen so that a plausible slip (wrong DST size or ghost handling, dropped
term, transposed operand, wrong sign) fails it.
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.fft

from oracle import dst, modes, qg, stencils

# ---------------------------------------------------------------------------
# DST-I Dirichlet solve: executed tutorial outputs
# ---------------------------------------------------------------------------

# content/tutorials/step09_laplace_2d.ipynb cells 7 / 11 / 13 and
# step10_poisson_2d.ipynb cell 4 (FP32 runs of the reference; 1e-5 relative).
POISSON_L2 = {16: 2.046900e-01, 32: 1.088572e-01, 64: 5.585729e-02, 128: 2.825762e-02}
POISSON_LINF_64 = 9.197066e-02
POISSON_HIGH_L2_64 = 7.211602e-02
# step11_helmholtz_2d.ipynb cell 6
HELMHOLTZ_MAX = {0.0: 0.053551, 1.0: 0.050761, 10.0: 0.034598, 100.0: 0.008362}


def _unit_grid(n):
    N = n + 2
    d = 1.0 / n
    x = np.arange(N) * d
    X, Y = np.meshgrid(x, x)
    return d, X, Y


@pytest.mark.parametrize("n", sorted(POISSON_L2))
def test_dst_poisson_tutorial_l2(n):
    d, X, Y = _unit_grid(n)
    rhs = -2.0 * np.pi**2 * np.sin(np.pi * X) * np.sin(np.pi * Y)
    phi = dst.helmholtz_dst(rhs, d, d)
    err = phi[1:-1, 1:-1] - (np.sin(np.pi * X) * np.sin(np.pi * Y))[1:-1, 1:-1]
    assert np.sqrt(np.mean(err**2)) == pytest.approx(POISSON_L2[n], rel=1e-5)
    if n == 64:
        assert np.max(np.abs(err)) == pytest.approx(POISSON_LINF_64, rel=1e-5)


def test_dst_poisson_tutorial_higher_mode():
    d, X, Y = _unit_grid(64)
    rhs = -13.0 * np.pi**2 * np.sin(2 * np.pi * X) * np.sin(3 * np.pi * Y)
    phi = dst.helmholtz_dst(rhs, d, d)
    err = phi[1:-1, 1:-1] - (np.sin(2 * np.pi * X) * np.sin(3 * np.pi * Y))[1:-1, 1:-1]
    assert np.sqrt(np.mean(err**2)) == pytest.approx(POISSON_HIGH_L2_64, rel=1e-5)


@pytest.mark.parametrize("lam", sorted(HELMHOLTZ_MAX))
def test_dst_helmholtz_tutorial_max(lam):
    d, X, Y = _unit_grid(64)
    phi = dst.helmholtz_dst(np.sin(np.pi * X) * np.sin(np.pi * Y), d, d, lam)
    assert np.max(np.abs(phi[1:-1, 1:-1])) == pytest.approx(HELMHOLTZ_MAX[lam], abs=1.5e-6)


def test_dst_ghost_ring_enters():
    """The tutorial numbers need the ghost ring of the RHS inside the transform:
    a ghost-zeroed RHS gives a visibly different L2 (2.1029e-1 vs 2.0469e-1)."""
    d, X, Y = _unit_grid(16)
    rhs = -2.0 * np.pi**2 * np.sin(np.pi * X) * np.sin(np.pi * Y)
    phi = dst.helmholtz_dst(stencils.zero_boundaries(rhs), d, d)
    err = phi[1:-1, 1:-1] - (np.sin(np.pi * X) * np.sin(np.pi * Y))[1:-1, 1:-1]
    assert abs(np.sqrt(np.mean(err**2)) - POISSON_L2[16]) > 1e-3


def test_dst_matrix_is_scipy_dst1():
    x = np.random.default_rng(0).standard_normal((7, 9))
    S7, S9 = dst.dst1_matrix(7), dst.dst1_matrix(9)
    np.testing.assert_allclose(S7 @ x @ S9, scipy.fft.dstn(x, type=1, norm="ortho"), atol=1e-13)
    np.testing.assert_allclose(S9 @ S9, np.eye(9), atol=1e-13)


@pytest.mark.parametrize("lam", [0.0, 3.0, 40.0])
def test_dst_is_brute_force_inverse(lam):
    """On a 6x5 grid the DST solve equals the dense inverse of the 5-point
    operator with zero values just outside the array, for non-square spacing."""
    ny, nx, dx, dy = 6, 5, 0.3, 0.2
    n = ny * nx
    M = np.zeros((n, n))
    for j in range(ny):
        for i in range(nx):
            r = j * nx + i
            M[r, r] = -2 / dx**2 - 2 / dy**2 - lam
            if i > 0:
                M[r, r - 1] = 1 / dx**2
            if i < nx - 1:
                M[r, r + 1] = 1 / dx**2
            if j > 0:
                M[r, r - nx] = 1 / dy**2
            if j < ny - 1:
                M[r, r + nx] = 1 / dy**2
    f = np.random.default_rng(1).standard_normal((ny, nx))
    np.testing.assert_allclose(dst.helmholtz_dst(f, dx, dy, lam), np.linalg.solve(M, f.ravel()).reshape(ny, nx),
                               rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(dst.helmholtz_apply(dst.helmholtz_dst(f, dx, dy, lam), dx, dy, lam), f, atol=1e-10)


def test_dst_per_mode_lambda_vector():
    """pv_inversion takes one lambda per leading index (baroclinic.py:146-152)."""
    f = np.random.default_rng(2).standard_normal((3, 8, 8))
    lam = np.array([0.0, 2.0, 9.0])
    out = dst.helmholtz_dst(f, 0.1, 0.1, lam)
    for m in range(3):
        np.testing.assert_allclose(out[m], dst.helmholtz_dst(f[m], 0.1, 0.1, lam[m]), atol=1e-12)


# ---------------------------------------------------------------------------
# vertical modes and the homogeneous solution
# ---------------------------------------------------------------------------
def test_modes_closed_forms():
    A = modes.coupling_matrix([400.0, 1100.0, 2600.0], [9.81, 0.025, 0.0125])
    w, Cl2m, Cm2l, radii = modes.vertical_modes(A, 9.375e-5)
    np.testing.assert_allclose(Cl2m @ Cm2l, np.eye(3), atol=1e-12)
    D = Cl2m @ A @ Cm2l
    np.testing.assert_allclose(D, np.diag(w), atol=1e-12 * np.abs(w).max())
    assert np.all(np.diff(w) > 0)
    # free-surface term: the barotropic eigenvalue is ~ 1/(g * total depth)
    assert w[0] == pytest.approx(1.0 / (9.81 * 4100.0), rel=5e-3)
    assert 40e3 < radii[1] < 43e3 and 25e3 < radii[2] < 26.5e3


def test_reference_eigenvalues_are_symmetrised():
    """What jnp.linalg.eigh returns for the non-symmetric A (transforms.py:194):
    eigenvalues of (A + A^T)/2, which are not those of A and include a
    negative one for this stratification."""
    A = modes.coupling_matrix([400.0, 1100.0, 2600.0], [9.81, 0.025, 0.0125])
    ref = modes.reference_eigenvalues(A)
    np.testing.assert_allclose(ref, np.linalg.eigvalsh(0.5 * (A + A.T)))
    assert ref[0] < 0.0
    assert not np.allclose(ref, np.sort(np.linalg.eigvals(A).real))


def test_reference_eigenvalues_make_the_barotropic_solve_indefinite():
    """With the reference's negative barotropic parameter (f0^2 * -0.00644) the
    mode's operator lap_h - lambda_0 is indefinite: Lam_x[i] + Lam_y[j] +
    |lambda_0| changes sign over the spectrum.  How close a discrete eigenvalue
    comes to resonance depends on the grid: on the 5120 km basin at 64^2 the
    smallest |denominator| is 2.8e-15 (condition 4.2e5, 40x a Poisson mode's),
    at 128^2 it is 4.6e-13 (condition 1.1e4, twice a Poisson mode's 5.3e3)."""
    A = modes.coupling_matrix([400.0, 1100.0, 2600.0], [9.81, 0.025, 0.0125])
    f0, L = 9.375e-5, 5120e3
    lam_ref = f0 * f0 * modes.reference_eigenvalues(A)[0]
    lam_poisson = f0 * f0 * modes.vertical_modes(A, f0)[0][0]
    cond = {}
    for n in (64, 128):
        lam_n = dst.dst1_eigenvalues(n + 2, L / n)
        den = lam_n[:, None] + lam_n[None, :]
        assert (den - lam_ref).max() > 0 > (den - lam_ref).min()  # indefinite
        assert (den - lam_poisson).max() < 0  # the consistent mode stays definite
        c = lambda d: np.abs(d).max() / np.abs(d).min()  # noqa: E731
        cond[n] = (c(den - lam_ref), c(den - lam_poisson))
    assert cond[64][0] > 1e5 and cond[64][0] > 30 * cond[64][1]
    assert cond[128][0] < 2e4 and cond[128][0] < 2.5 * cond[128][1]


def test_homsol_notebook_known_answers():
    """notebooks/dev/QG_Multilayer/mqg_demo_retangular.ipynb cell 18 (x64):
    rectangle of 256 nodes, L = 4000 km, dx = L/N, f0 = 9.4e-5, rigid-lid A.
    HOMSOL mean 3.611293e-01, max 1.0; HOMSOL MEAN min 2.656387e-02,
    mean 3.561087e-01, max 1.0."""
    N, L, f0 = 256, 4000e3, 9.4e-5
    d = L / N
    A = modes.coupling_matrix([400.0, 1100.0, 2600.0], [9.81, 0.025, 0.0125], free_surface=False)
    lam = f0**2 * np.sort(np.linalg.eigvals(A).real)
    hs = []
    for lm in lam:
        h = np.ones((N, N))
        h[1:-1, 1:-1] += dst.helmholtz_dst(lm * np.ones((N - 2, N - 2)), d, d, lm)
        hs.append(h)
    hs = np.array(hs)
    assert hs.mean() == pytest.approx(3.611293e-01, rel=2e-5)
    assert hs.max() == pytest.approx(1.0, abs=1e-12)
    assert abs(hs.min()) < 1e-6
    cell = 0.25 * (hs[:, 1:, 1:] + hs[:, :-1, 1:] + hs[:, 1:, :-1] + hs[:, :-1, :-1])
    means = cell.mean(axis=(1, 2))
    assert means.max() == pytest.approx(1.0, abs=1e-12)
    assert means.min() == pytest.approx(2.656387e-02, rel=2e-5)
    assert means.mean() == pytest.approx(3.561087e-01, rel=2e-5)


def test_homsol_full_array_properties():
    """The model-grid homogeneous solution: (D - lam) h = 0 away from the walls
    and equals 1 for lam = 0."""
    lam = np.array([0.0, 3e-10, 1.5e-9])
    h, hm = qg.homogeneous_solution(18, 18, 3e4, 3e4, lam)
    np.testing.assert_allclose(h[0], 1.0, atol=1e-12)
    r = dst.helmholtz_apply(h, 3e4, 3e4, lam)
    # the residual is (lap_h 1) at the array edge only
    np.testing.assert_allclose(r[:, 1:-1, 1:-1], 0.0, atol=1e-22)
    assert np.all(hm[1:] < 1.0) and np.all(hm > 0)


# ---------------------------------------------------------------------------
# stencils
# ---------------------------------------------------------------------------
def test_laplacian_exact_on_quadratics():
    ny, nx, dx, dy = 9, 11, 0.5, 0.25
    x = np.arange(nx) * dx
    y = np.arange(ny) * dy
    X, Y = np.meshgrid(x, y)
    f = 3 * X**2 - 2 * X * Y + 5 * Y**2 + X - 7
    lap = stencils.laplacian(f, dx, dy)
    np.testing.assert_allclose(lap[1:-1, 1:-1], 6 + 10, atol=1e-10)
    assert np.all(lap[0] == 0) and np.all(lap[:, -1] == 0)


def test_arakawa_invariants():
    rng = np.random.default_rng(3)
    p = stencils.zero_boundaries(rng.standard_normal((2, 14, 17)))
    q = stencils.zero_boundaries(rng.standard_normal((2, 14, 17)))
    J = stencils.arakawa_jacobian(p, q, 0.7, 1.3)
    # energy and enstrophy conservation with zero ghost rings (Arakawa 1966)
    assert abs(np.sum(p[:, 1:-1, 1:-1] * J)) < 1e-12 * np.sum(np.abs(J))
    assert abs(np.sum(q[:, 1:-1, 1:-1] * J)) < 1e-12 * np.sum(np.abs(J))
    # antisymmetry
    np.testing.assert_allclose(stencils.arakawa_jacobian(q, p, 0.7, 1.3), -J, atol=1e-13)
    # J(x, y) = 1 for the coordinate fields
    X, Y = np.meshgrid(np.arange(17) * 0.7, np.arange(14) * 1.3)
    np.testing.assert_allclose(stencils.arakawa_jacobian(X, Y, 0.7, 1.3), 1.0, atol=1e-12)


def test_arakawa_continuum_limit():
    """J(sin x sin y, cos x + y^2) against the analytic Jacobian: 2nd-order convergence."""
    errs = []
    for n in (32, 64):
        h = 1.0 / n
        x = np.arange(n + 2) * h
        X, Y = np.meshgrid(x, x)
        p = np.sin(X) * np.sin(Y)
        q = np.cos(X) + Y**2
        exact = (np.cos(X) * np.sin(Y)) * (2 * Y) - (np.sin(X) * np.cos(Y)) * (-np.sin(X))
        errs.append(np.max(np.abs(stencils.arakawa_jacobian(p, q, h, h) - exact[1:-1, 1:-1])))
    assert errs[0] / errs[1] == pytest.approx(4.0, rel=0.1)


def test_weno_exact_for_constants_and_conservative():
    rng = np.random.default_rng(4)
    psi = stencils.zero_boundaries(rng.standard_normal((2, 16, 21)))
    t = stencils.weno_advection(psi, np.full((2, 16, 21), 2.5), 1.0, 1.0)
    np.testing.assert_allclose(t, 0.0, atol=1e-13)
    q = rng.standard_normal((2, 16, 21))
    t = stencils.weno_advection(psi, q, 1.0, 1.0)
    np.testing.assert_allclose(t[:, 1:-1, 1:-1].sum(axis=(-2, -1)), 0.0, atol=1e-12)
    assert np.all(t[:, 0, :] == 0) and np.all(t[:, :, -1] == 0)


def test_weno_face_velocities_divergence_free():
    psi = stencils.zero_boundaries(np.random.default_rng(5).standard_normal((12, 15)))
    u, v = stencils.weno_velocities(psi, 0.5, 0.8, psi_wall=0.3)
    ny, nx = 10, 13
    Fx = np.zeros((ny, nx + 1))
    Fx[:, 1:-1] = u
    Fy = np.zeros((ny + 1, nx))
    Fy[1:-1, :] = v
    div = (Fx[:, 1:] - Fx[:, :-1]) / 0.5 + (Fy[1:] - Fy[:-1]) / 0.8
    np.testing.assert_allclose(div, 0.0, atol=1e-12)


def test_weno5z_fifth_order_smooth():
    """Interior WENO-Z reconstruction of a cell-average-free smooth profile
    converges at 5th order (point values of sin at faces)."""
    errs = []
    for n in (40, 80):
        h = 1.0 / n
        xc = (np.arange(-2, n + 3)) * h
        f = np.sin(2 * np.pi * xc) + 2.0
        # point-value reconstruction: use the finite-difference form (cells -> face)
        face = stencils.weno5z_left(f[:-4], f[1:-3], f[2:-2], f[3:-1], f[4:])
        xf = xc[2:-2] + 0.5 * h
        # compare against the face value of the primitive-consistent flux function:
        # for point data, WENO5 is 5th order for the interpolant of the flux
        # function h(x) with cell averages equal to f; evaluate via the
        # identity f(x) = (1/h) int_{x-h/2}^{x+h/2} h(s) ds  =>  h = f - h^2/24 f'' + ...
        hf = (np.sin(2 * np.pi * xf) + 2.0) * 1.0
        corr = -(h**2) / 24 * (-(4 * np.pi**2) * np.sin(2 * np.pi * xf))
        corr4 = 7 * h**4 / 5760 * (16 * np.pi**4 * np.sin(2 * np.pi * xf))
        errs.append(np.max(np.abs(face - (hf + corr + corr4))))
    assert np.log2(errs[0] / errs[1]) > 4.5


def test_ssprk3_third_order():
    """SSP-RK3 (Shu-Osher form used by oracle.qg.ssprk3_step) on y' = lam y."""
    lam = -1.3
    errs = []
    for n in (20, 40, 80):
        dt = 1.0 / n
        y = 1.0
        for _ in range(n):
            k0 = lam * y
            y1 = y + dt * k0
            y2 = 0.75 * y + 0.25 * (y1 + dt * lam * y1)
            y = y / 3.0 + 2.0 / 3.0 * (y2 + dt * lam * y2)
        errs.append(abs(y - np.exp(lam)))
    assert np.log2(errs[0] / errs[1]) == pytest.approx(3.0, abs=0.1)
    assert np.log2(errs[1] / errs[2]) == pytest.approx(3.0, abs=0.1)


def test_ssprk3_matches_notebook_stage_form():
    """Shu-Osher stages equal the ssrk3_update_stage_{1,2,3} increments of
    mqg_demo_retangular.ipynb cell 31 for an arbitrary nonlinear RHS."""
    f = lambda y: np.sin(y) * y  # noqa: E731
    y0, dt = np.array([0.3, -1.1, 2.0]), 0.17
    k0 = f(y0)
    a = y0 + dt * k0
    k1 = f(a)
    b = a + dt / 4 * (k1 - 3 * k0)
    k2 = f(b)
    c = b + dt / 12 * (8 * k2 - k1 - k0)
    y1 = y0 + dt * f(y0)
    y2 = 0.75 * y0 + 0.25 * (y1 + dt * f(y1))
    y3 = y0 / 3 + 2 / 3 * (y2 + dt * f(y2))
    np.testing.assert_allclose(y3, c, atol=1e-14)


# ---------------------------------------------------------------------------
# QG model: behaviours the reference tests assert
# ---------------------------------------------------------------------------
def _small_model(**kw):
    base = dict(nx=16, ny=16, H=(500.0, 4500.0), g_prime=(9.81, 0.025))
    base.update(kw)
    return qg.baroclinic_setup(**base)


def test_qg_zero_state_zero_tendency():
    """test_qg_baroclinic.py:79-94."""
    s = _small_model()
    assert np.max(np.abs(qg.rhs(s, np.zeros((2, 18, 18))))) < 1e-12


def test_qg_wind_top_layer_only():
    """test_qg_baroclinic.py:96-112."""
    s = _small_model(wind_amplitude=1e-10)
    dq = qg.rhs(s, np.zeros((2, 18, 18)))
    assert np.max(np.abs(dq[0])) > 0 and np.max(np.abs(dq[1])) < 1e-15


def test_qg_inversion_inverts_stretching_operator():
    """With consistent eigenvalues the layer-space inversion solves
    (lap_h - f0^2 A) psi = q on the full array (then zeroes psi's ghost ring)."""
    s = _small_model()
    s.zero_psi_ghost = False
    q = stencils.zero_boundaries(np.random.default_rng(6).standard_normal((2, 18, 18)))
    psi = qg.invert_pv(s, q)
    A = modes.coupling_matrix(s.H, s.g_prime)
    lhs = dst.helmholtz_apply(psi, s.dx, s.dy) - s.f0**2 * np.einsum("lk,kyx->lyx", A, psi)
    np.testing.assert_allclose(lhs, q, atol=1e-9 * np.abs(q).max())


def test_qg_bc_semantics():
    """test_qg_baroclinic.py:166-183."""
    q = np.ones((2, 18, 18))
    b = qg.apply_bc(q)
    assert np.all(b[:, 0] == 0) and np.all(b[:, -1] == 0) and np.all(b[:, :, 0] == 0)
    assert np.all(b[:, 1:-1, 1:-1] == 1)


def test_qg_mass_correction_zero_mean_modes():
    s = _small_model(mass_correction=True)
    q = stencils.zero_boundaries(np.random.default_rng(7).standard_normal((2, 18, 18)))
    _, pm, alpha = qg.invert_pv(s, q, return_modal=True)
    np.testing.assert_allclose(pm.mean(axis=(-2, -1)), 0.0, atol=1e-14 * np.abs(pm).max())
    assert np.all(alpha != 0)


def test_qg_weno_rhs_conserves_pv_without_forcing():
    s = _small_model(advection="weno")
    q = stencils.zero_boundaries(np.random.default_rng(8).standard_normal((2, 18, 18)) * 1e-6)
    dq = qg.rhs(s, q)
    np.testing.assert_allclose(dq[:, 1:-1, 1:-1].sum(axis=(-2, -1)), 0.0, atol=1e-18)


def test_qg_diagnose_closed_form():
    """Quadratic psi = c j^2 (barotropic: the DST solve returns it exactly from
    q = its zero-padded Laplacian): the T-point velocity is the centred
    difference -2 c j / dy, so KE = 0.5 sum_j (2 c j / dy)^2 nx dx dy over the
    interior rows; a forward or backward difference would give (2j +- 1).
    The wall-face velocity u[0] = -c/dy is included; u on the last row is 0."""
    from oracle import dst
    from oracle import qg as oqg

    n = 12
    s = oqg.barotropic_setup(nx=n, ny=n, Lx=1.2e5, Ly=1.2e5)
    j = np.arange(s.Ny, dtype=np.float64)
    c = 1e3
    psi = np.broadcast_to((c * j**2)[:, None], (s.Ny, s.Nx)).copy()
    q = dst.helmholtz_apply(psi, s.dx, s.dy)
    d = oqg.diagnose(s, q)
    np.testing.assert_allclose(d["psi"][0], psi, rtol=1e-9, atol=1e-9 * psi.max())
    jj = np.arange(1, n + 1, dtype=np.float64)
    ke = 0.5 * np.sum((2 * c * jj / s.dy) ** 2) * n * s.dx * s.dy
    np.testing.assert_allclose(d["kinetic_energy"], ke, rtol=1e-9)
    np.testing.assert_allclose(d["u"][0, 0, 3], -c / s.dy, rtol=1e-9)
    assert np.all(d["u"][0, -1] == 0) and np.abs(d["v"]).max() < 1e-6 * c / s.dx
    np.testing.assert_allclose(d["enstrophy"], 0.5 * np.sum(q[1:-1, 1:-1] ** 2) * s.dx * s.dy)
    # reflection symmetry of the summed energy
    d2 = oqg.diagnose(s, q[::-1, ::-1].copy())
    np.testing.assert_allclose(d2["kinetic_energy"], d["kinetic_energy"], rtol=1e-9)


# ---------------------------------------------------------------------------
# Tsit5 (oracle/rk.py): order conditions and convergence
# ---------------------------------------------------------------------------
def test_tsit5_order_conditions():
    """The 17 Butcher order conditions of orders 1-5 (one per rooted tree) and
    c_i = sum_j a_ij: a wrong digit in any coefficient breaks one of them."""
    from oracle.rk import TSIT5_A as A, TSIT5_B as b, TSIT5_C as c

    np.testing.assert_allclose(A.sum(axis=1), c, atol=1e-14)
    Ac = A @ c
    Ac2, Ac3 = A @ c**2, A @ c**3
    AAc = A @ Ac
    conds = [
        (b.sum(), 1.0),
        (b @ c, 1 / 2),
        (b @ c**2, 1 / 3), (b @ Ac, 1 / 6),
        (b @ c**3, 1 / 4), (b @ (c * Ac), 1 / 8), (b @ Ac2, 1 / 12), (b @ AAc, 1 / 24),
        (b @ c**4, 1 / 5), (b @ (c**2 * Ac), 1 / 10), (b @ (c * Ac2), 1 / 15), (b @ (c * AAc), 1 / 30),
        (b @ Ac**2, 1 / 20), (b @ Ac3, 1 / 20), (b @ (A @ (c * Ac)), 1 / 40), (b @ (A @ Ac2), 1 / 60),
        (b @ (A @ AAc), 1 / 120),
    ]
    assert len(conds) == 17
    for k, (got, want) in enumerate(conds):
        assert abs(got - want) < 1e-12, (k, got, want)


def test_tsit5_fifth_order_convergence():
    """Global error at t = 1 of y' = lambda y (exact exp(lambda t)) falls by
    ~2^5 per halving of dt."""
    from oracle import rk

    lam = -1.3 + 0.7j
    errs = []
    for n in (10, 20, 40):
        y = (np.array([1.0 + 0j]),)
        for _ in range(n):
            y = rk.explicit_step(lambda v: (lam * v[0],), y, 1.0 / n)
        errs.append(abs(y[0][0] - np.exp(lam)))
    rates = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(4.7 < r < 5.6 for r in rates), rates


# ---------------------------------------------------------------------------
# ReparameterizedQG (oracle/reparam.py)
# ---------------------------------------------------------------------------
def _rqg(n=16, **kw):
    from oracle import reparam

    args = dict(nx=n, ny=n, H=(500.0, 500.0), g_prime=(9.81, 0.02))
    args.update(kw)
    return reparam.setup(**args)


def test_reparam_q_of_g_is_the_helmholtz_operator():
    """Q(G(psi)) = lap(psi) - f0^2 M psi wherever both stencils lie in the
    array (all but the last row and column): pins the vorticity and grad_perp
    index senses against each other and the f0, H factors of Q and G."""
    from oracle import reparam

    r = _rqg(12, H=(400.0, 1100.0, 2600.0), g_prime=(9.81, 0.025, 0.0125), Lx=1.2e6, Ly=0.9e6)
    rng = np.random.default_rng(5)
    psi = rng.standard_normal((3, r.swm.Ny, r.swm.Nx)) * 1e4
    psi[:, 0, :] = psi[:, -1, :] = psi[:, :, 0] = psi[:, :, -1] = 0.0
    q = reparam.pv(r, *reparam.geostrophic(r, psi))
    pad = np.pad(psi, ((0, 0), (1, 1), (1, 1)))
    lap = (pad[:, 1:-1, 2:] - 2 * psi + pad[:, 1:-1, :-2]) / r.swm.dx**2 + (
        pad[:, 2:, 1:-1] - 2 * psi + pad[:, :-2, 1:-1]) / r.swm.dy**2
    want = lap - r.f0**2 * np.einsum("kl,lyx->kyx", r.M, psi)
    np.testing.assert_allclose(q[:, :-1, :-1], want[:, :-1, :-1], rtol=0, atol=1e-12 * np.abs(want).max())


def test_reparam_geostrophic_velocity_divergence_free():
    from oracle import reparam

    r = _rqg(10)
    rng = np.random.default_rng(6)
    psi = rng.standard_normal((2, r.swm.Ny, r.swm.Nx))
    _, u, v = reparam.geostrophic(r, psi)
    div = (u[:, 1:, 1:] - u[:, 1:, :-1]) / r.swm.dx + (v[:, 1:, 1:] - v[:, :-1, 1:]) / r.swm.dy
    assert np.abs(div).max() < 1e-12 * np.abs(psi).max() / min(r.swm.dx, r.swm.dy) ** 2


def test_reparam_reference_behaviours():
    """tests/models/test_qg_reparameterized.py:60-87, 117-125: rest state is
    a fixed point, the projection is idempotent to O(dx^2) (atol 0.1 for a 1 m
    bump), and the rest state has zero tendency."""
    from oracle import reparam

    r = _rqg(16)
    H = r.swm.H[:, None, None]
    h0 = np.ones((2, r.swm.Ny, r.swm.Nx)) * H
    z = np.zeros_like(h0)
    hp, up, vp = reparam.project(r, h0, z, z)
    assert np.abs(hp - h0).max() < 1e-6 and np.abs(up).max() < 1e-10 and np.abs(vp).max() < 1e-10
    h1 = h0.copy()
    h1[0, r.swm.Ny // 2, r.swm.Nx // 2] += 1.0
    p1 = reparam.project(r, h1, z, z)
    p2 = reparam.project(r, *p1)
    for a, b in zip(p2, p1):
        assert np.abs(a - b).max() < 0.1
    assert np.abs(p1[1]).max() > 0 and np.abs(p1[2]).max() > 0
    t = reparam.rhs_of_bc(r, (h0, z, z))
    assert max(np.abs(x).max() for x in t) < 1e-10
