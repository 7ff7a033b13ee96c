"""Multilayer / barotropic QG right-hand side and SSP-RK3 step (FP64).

Restates ``BaroclinicQG`` (``somax/_src/models/qg/baroclinic.py:135-195,
231-332``) and ``BarotropicQG`` (``somax/_src/models/qg/barotropic.py:113-157,
186-248``), plus the extensions the benchmark configurations ask for and that
the current reference does not contain (each marked *extension*):

* SSP-RK3 (``notebooks/dev/QG_Multilayer/mqg_demo_retangular.ipynb`` cell 31,
  the ``ssrk3_update_stage_{1,2,3}`` functions; here in Shu-Osher form, which
  is algebraically the same update);
* WENO-Z flux-form PV advection (removed API, ``method="wenoz"``);
* biharmonic viscosity ``-nu4 del^4 q`` (absent everywhere);
* homogeneous-solution mass correction (removed
  ``compute_homogeneous_solution``; MQGeometry formula, pinned by the numbers
  the notebook prints in cell 18).

Stage semantics follow ``SomaxModel.build_terms`` (``somax/_src/core/model.py:
39-51``): the boundary condition (ghost ring of q set to 0) is applied to the
stage state before each RHS evaluation, while the stage states themselves keep
their ghost values (they accumulate the wind forcing, which the reference adds
over the full array at baroclinic.py:181).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import dst, modes, stencils


@dataclass
class QGSetup:
    """Everything ``create()`` precomputes, in FP64."""

    nx: int
    ny: int
    Lx: float
    Ly: float
    nl: int
    f0: float
    beta: float
    H: np.ndarray
    g_prime: np.ndarray
    Cl2m: np.ndarray
    Cm2l: np.ndarray
    lambdas: np.ndarray  # Helmholtz parameter per mode (f0^2 * eigenvalue)
    beta_y: np.ndarray  # (Ny, Nx)
    wind_forcing: np.ndarray  # (Ny, Nx), normalised pattern
    wind_scale: np.ndarray  # per layer multiplier of tau0 * pattern
    zero_psi_ghost: bool
    nu: float = 0.0  # harmonic viscosity (reference)
    nu4: float = 0.0  # biharmonic viscosity (extension)
    kappa: float = 0.0  # bottom drag
    tau0: float = 0.0  # wind amplitude
    advection: str = "arakawa"  # or "weno" (extension)
    mass_correction: bool = False  # extension
    homsol: np.ndarray | None = None
    homsol_mean: np.ndarray | None = None
    extra: dict = field(default_factory=dict)

    @property
    def Nx(self) -> int:
        return self.nx + 2

    @property
    def Ny(self) -> int:
        return self.ny + 2

    @property
    def dx(self) -> float:
        return self.Lx / self.nx

    @property
    def dy(self) -> float:
        return self.Ly / self.ny


def _row_fields(ny, nx, Lx, Ly, beta, wind_profile):
    """beta*(y - Ly/2) and the wind-curl pattern over T rows incl. ghosts (baroclinic.py:307-318)."""
    dy = Ly / ny
    y = np.arange(ny + 2, dtype=np.float64) * dy
    Y = np.broadcast_to(y[:, None], (ny + 2, nx + 2))
    beta_y = beta * (Y - Ly / 2.0)
    if wind_profile == "single":
        wind = np.sin(np.pi * Y / Ly)
    else:
        wind = -np.sin(2.0 * np.pi * Y / Ly)
    return np.array(beta_y), np.array(wind)


def baroclinic_setup(
    nx=64,
    ny=64,
    Lx=4e6,
    Ly=4e6,
    f0=9.375e-5,
    beta=1.754e-11,
    H=(400.0, 1100.0, 2600.0),
    g_prime=(9.81, 0.025, 0.0125),
    lateral_viscosity=0.0,
    bottom_drag=0.0,
    wind_amplitude=0.0,
    wind_profile="doublegyre",
    biharmonic_viscosity=0.0,
    advection="arakawa",
    mass_correction=False,
    eigenvalues="consistent",
    free_surface=True,
) -> QGSetup:
    """``BaroclinicQG.create`` (baroclinic.py:231-332) in FP64."""
    H = np.asarray(H, dtype=np.float64)
    gp = np.asarray(g_prime, dtype=np.float64)
    A = modes.coupling_matrix(H, gp, free_surface=free_surface)
    w, Cl2m, Cm2l, _ = modes.vertical_modes(A, f0)
    if eigenvalues == "reference":
        w = modes.reference_eigenvalues(A)
    lambdas = f0 * f0 * w
    beta_y, wind = _row_fields(ny, nx, Lx, Ly, beta, wind_profile)
    nl = H.size
    scale = np.zeros(nl)
    scale[0] = 1.0 / H[0]
    s = QGSetup(
        nx=nx, ny=ny, Lx=Lx, Ly=Ly, nl=nl, f0=f0, beta=beta, H=H, g_prime=gp,
        Cl2m=Cl2m, Cm2l=Cm2l, lambdas=lambdas, beta_y=beta_y, wind_forcing=wind,
        wind_scale=scale, zero_psi_ghost=not mass_correction,
        nu=lateral_viscosity, nu4=biharmonic_viscosity, kappa=bottom_drag,
        tau0=wind_amplitude, advection=advection, mass_correction=mass_correction,
    )
    if mass_correction:
        s.homsol, s.homsol_mean = homogeneous_solution(s.Ny, s.Nx, s.dx, s.dy, lambdas)
    return s


def barotropic_setup(
    nx=64, ny=64, Lx=1e6, Ly=1e6, f0=1e-4, beta=1.6e-11,
    lateral_viscosity=0.0, bottom_drag=0.0, wind_amplitude=0.0,
    wind_profile="doublegyre", biharmonic_viscosity=0.0, advection="arakawa",
) -> QGSetup:
    """``BarotropicQG.create`` (barotropic.py:186-248): one layer, Poisson
    inversion, no 1/H on the wind, drag on the only layer, psi ghost ring kept
    (no ``zero_boundaries`` after the solve, barotropic.py:113-121)."""
    beta_y, wind = _row_fields(ny, nx, Lx, Ly, beta, wind_profile)
    return QGSetup(
        nx=nx, ny=ny, Lx=Lx, Ly=Ly, nl=1, f0=f0, beta=beta,
        H=np.array([1.0]), g_prime=np.array([0.0]),
        Cl2m=np.eye(1), Cm2l=np.eye(1), lambdas=np.zeros(1),
        beta_y=beta_y, wind_forcing=wind, wind_scale=np.ones(1),
        zero_psi_ghost=False, nu=lateral_viscosity, nu4=biharmonic_viscosity,
        kappa=bottom_drag, tau0=wind_amplitude, advection=advection,
    )


def homogeneous_solution(Ny, Nx, dx, dy, lambdas):
    """Per-mode homogeneous solution ``h_m = 1 + (D - lam_m)^-1 [lam_m * 1]``
    (value 1 on the DST walls, ``(D - lam_m) h_m = 0`` inside) and its mean
    over the full array.  MQGeometry formula behind the removed
    ``compute_homogeneous_solution`` (mqg_demo_retangular.ipynb cell 18)."""
    lam = np.asarray(lambdas, dtype=np.float64)
    ones = np.ones((lam.size, Ny, Nx))
    h = 1.0 + dst.helmholtz_dst(lam[:, None, None] * ones, dx, dy, lam)
    return h, h.mean(axis=(-2, -1))


def apply_bc(q: np.ndarray) -> np.ndarray:
    """``apply_boundary_conditions`` (baroclinic.py:192-195): q ghost ring = 0."""
    return stencils.zero_boundaries(q)


def invert_pv(s: QGSetup, q: np.ndarray, return_modal=False, workers: int | None = None):
    """``_invert_pv`` (baroclinic.py:135-159): to_modal, per-mode DST-I
    Helmholtz solve, [mass correction], to_layer, [ghost ring of psi = 0].
    ``workers``: threads of the sine transforms (results unchanged)."""
    q = np.asarray(q, dtype=np.float64).reshape(s.nl, s.Ny, s.Nx)
    qm = np.einsum("lm,m...->l...", s.Cl2m, q)
    pm = dst.helmholtz_dst(qm, s.dx, s.dy, s.lambdas, workers=workers)
    alpha = np.zeros(s.nl)
    if s.mass_correction:
        alpha = -pm.mean(axis=(-2, -1)) / s.homsol_mean
        pm = pm + alpha[:, None, None] * s.homsol
    psi = np.einsum("lm,m...->l...", s.Cm2l, pm)
    if s.zero_psi_ghost:
        psi = stencils.zero_boundaries(psi)
    if return_modal:
        return psi, pm, alpha
    return psi


def psi_wall(s: QGSetup, alpha: np.ndarray) -> np.ndarray:
    """Wall value of psi per layer used by the WENO corner streamfunction:
    ``Cm2l @ alpha`` (0 without mass correction)."""
    return s.Cm2l @ alpha


def rhs(s: QGSetup, q: np.ndarray, return_psi=False, fast=False):
    """``vector_field`` (baroclinic.py:161-190 / barotropic.py:123-152) on a
    state whose boundary condition has already been applied.

    ``fast``: the same arithmetic with threaded sine transforms and the
    stencil terms from ``oracle/qgrhs.c`` (FP64, OpenMP), which restates the
    NumPy expressions below operation for operation (same bits; pinned in
    ``tests/test_oracle_pins.py``).  Used where the grid is large."""
    q = np.asarray(q, dtype=np.float64).reshape(s.nl, s.Ny, s.Nx)
    psi, _, alpha = invert_pv(s, q, return_modal=True, workers=-1 if fast else None)
    qt = q + s.beta_y[None]
    if fast:
        from . import cpu

        dq = cpu.qg_stencil_rhs(s, psi, q, qt, psi_wall(s, alpha))
        return (dq, psi) if return_psi else dq
    if s.advection == "weno":
        dq = stencils.weno_advection(psi, qt, s.dx, s.dy, psi_wall(s, alpha))
    else:
        dq = np.zeros_like(q)
        dq[:, 1:-1, 1:-1] = -stencils.arakawa_jacobian(psi, qt, s.dx, s.dy)
    dq = dq + (s.tau0 * s.wind_scale)[:, None, None] * s.wind_forcing[None]
    if s.kappa != 0.0:
        dq[-1] += -s.kappa * stencils.laplacian(psi[-1], s.dx, s.dy)
    if s.nu != 0.0:
        dq = dq + s.nu * stencils.laplacian(q, s.dx, s.dy)
    if s.nu4 != 0.0:
        dq = dq - s.nu4 * stencils.biharmonic(q, s.dx, s.dy)
    if return_psi:
        return dq, psi
    return dq


def ssprk3_step(s: QGSetup, q: np.ndarray, dt: float, fast=False) -> np.ndarray:
    """One SSP-RK3 step (Shu-Osher form) with the BC applied before each RHS."""
    q0 = np.asarray(q, dtype=np.float64)
    q1 = q0 + dt * rhs(s, apply_bc(q0), fast=fast)
    q2 = 0.75 * q0 + 0.25 * (q1 + dt * rhs(s, apply_bc(q1), fast=fast))
    return q0 / 3.0 + 2.0 / 3.0 * (q2 + dt * rhs(s, apply_bc(q2), fast=fast))


def tsit5_step(s: QGSetup, q: np.ndarray, dt: float, fast=False) -> np.ndarray:
    """One Tsit5 step (``oracle/rk.py``; ``integrate``'s default solver,
    core/model.py:75) with the BC applied before each RHS (core/model.py:47-49);
    ``fast`` as in ``rhs``."""
    from . import rk

    return rk.explicit_step(lambda y: (rhs(s, apply_bc(y[0]), fast=fast),), (q,), dt)[0]


def run(s: QGSetup, q: np.ndarray, dt: float, nsteps: int, fast=False) -> np.ndarray:
    for _ in range(nsteps):
        q = ssprk3_step(s, q, dt, fast=fast)
    return q


def diagnose(s: QGSetup, q: np.ndarray):
    """``diagnose`` (baroclinic.py:197-228, barotropic.py:159-183): psi,
    velocities, kinetic energy and enstrophy per layer, relative vorticity.

    Convention (finitevolX's index ranges are parity unpinned): every operator
    is evaluated wherever its stencil lies inside the array.
    u = -diff_y_T_to_V(psi) on rows 0..Ny-2, v = diff_x_T_to_U(psi) on columns
    0..Nx-2 (forward differences); V_to_T / U_to_T average the two faces of a
    T cell, so the T-point velocities are centred differences of psi; KE and
    enstrophy sum the interior T cells times dx dy; zeta = 5-point Laplacian on
    the interior."""
    q = np.asarray(q, dtype=np.float64).reshape(s.nl, s.Ny, s.Nx)
    psi = invert_pv(s, q)
    dx, dy = s.dx, s.dy
    u = np.zeros_like(psi)
    v = np.zeros_like(psi)
    u[..., :-1, :] = -(psi[..., 1:, :] - psi[..., :-1, :]) / dy
    v[..., :, :-1] = (psi[..., :, 1:] - psi[..., :, :-1]) / dx
    uT = 0.5 * (u[..., 1:-1, 1:-1] + u[..., :-2, 1:-1])
    vT = 0.5 * (v[..., 1:-1, 1:-1] + v[..., 1:-1, :-2])
    area = dx * dy
    ke = 0.5 * np.sum(uT**2 + vT**2, axis=(-2, -1)) * area
    ens = 0.5 * np.sum(q[..., 1:-1, 1:-1] ** 2, axis=(-2, -1)) * area
    zeta = np.zeros_like(psi)
    c = psi[..., 1:-1, 1:-1]
    zeta[..., 1:-1, 1:-1] = (psi[..., 1:-1, 2:] - 2 * c + psi[..., 1:-1, :-2]) / dx**2 + (
        psi[..., 2:, 1:-1] - 2 * c + psi[..., :-2, 1:-1]
    ) / dy**2
    return {"psi": psi, "u": u, "v": v, "kinetic_energy": ke, "total_kinetic_energy": ke.sum(),
            "enstrophy": ens, "total_enstrophy": ens.sum(), "relative_vorticity": zeta}
