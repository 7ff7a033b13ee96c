"""ReparameterizedQG restated in FP64 (test infrastructure).

Follows ``somax/_src/models/qg/reparameterized.py``: ``project`` (:142-178),
``_solve_helmholtz`` (:125-140), ``apply_boundary_conditions`` (:186-189:
the shallow-water boundary condition, then the projection), ``diagnose``
(:191-222) and ``create`` (:224-315; wall boundaries only).

Stated conventions (finitevolX v0.0.39 is absent; *parity unpinned* against
the reference): the relative vorticity is the shallow-water oracle's
(``oracle/swm.py``: X point [j, i] = north-east corner of cell [j, i],
forward differences, 0 on the last row and column); ``grad_perp`` treats psi
as living on those X points, so u on the east face of cell [j, i] is
``-(psi[j, i] - psi[j-1, i]) / dy`` and v on its north face
``(psi[j, i] - psi[j, i-1]) / dx`` (psi read as 0 outside the array).  With
these, Q(G(psi)) = (lap - f0^2 A) psi wherever both stencils lie inside the
array, which ``tests/test_oracle_pins.py`` checks, together with the
reference's own qualitative tests (tests/models/test_qg_reparameterized.py).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import dst, modes, rk, swm


@dataclass
class RQGSetup:
    swm: swm.SWMSetup
    f0: float
    Cl2m: np.ndarray
    Cm2l: np.ndarray
    eigenvalues: np.ndarray
    lambdas: np.ndarray  # f0^2 * eigenvalues

    @property
    def M(self) -> np.ndarray:
        """Cm2l diag(eigenvalues) Cl2m: A psi in layer space (reparameterized.py:174-175)."""
        return self.Cm2l @ np.diag(self.eigenvalues) @ self.Cl2m


def setup(nx=64, ny=64, Lx=4e6, Ly=4e6, f0=9.375e-5, beta=1.754e-11, H=(400.0, 1100.0, 2600.0),
          g_prime=(9.81, 0.025, 0.0125), lateral_viscosity=0.0, bottom_drag=0.0, wind_amplitude=0.0,
          wind_profile="doublegyre", bc="wall", eigenvalues="consistent", free_surface=True) -> RQGSetup:
    """``ReparameterizedQG.create``: the shallow-water model plus the modal
    tables; helmholtz_lambdas = f0^2 * eigenvalues (reparameterized.py:310)."""
    if bc != "wall":
        raise ValueError(f"ReparameterizedQG requires wall BCs (got bc={bc!r}).")
    s = swm.setup(nx=nx, ny=ny, Lx=Lx, Ly=Ly, f0=f0, beta=beta, H=H, g_prime=g_prime,
                  lateral_viscosity=lateral_viscosity, bottom_drag=bottom_drag, wind_amplitude=wind_amplitude,
                  wind_profile=wind_profile, bc=bc)
    A = modes.coupling_matrix(np.asarray(H, dtype=np.float64), np.asarray(g_prime, dtype=np.float64),
                              free_surface=free_surface)
    w, Cl2m, Cm2l, _ = modes.vertical_modes(A, f0)
    if eigenvalues == "reference":
        w = modes.reference_eigenvalues(A)
    return RQGSetup(swm=s, f0=f0, Cl2m=Cl2m, Cm2l=Cm2l, eigenvalues=w, lambdas=f0 * f0 * w)


def solve_helmholtz(r: RQGSetup, q):
    """(Q G)^-1: modal DST-I Helmholtz solve, ghost ring of psi set to 0."""
    qm = np.einsum("ml,l...->m...", r.Cl2m, np.asarray(q, dtype=np.float64))
    pm = dst.helmholtz_dst(qm, r.swm.dx, r.swm.dy, r.lambdas)
    p = np.einsum("lm,m...->l...", r.Cm2l, pm)
    p[..., 0, :] = p[..., -1, :] = 0.0
    p[..., :, 0] = p[..., :, -1] = 0.0
    return p


def pv(r: RQGSetup, h, u, v):
    """Q: q = zeta(u, v) - f0 (h - H) / H."""
    H = r.swm.H[:, None, None]
    return swm.relative_vorticity(r.swm, u, v) - r.f0 * (np.asarray(h, dtype=np.float64) - H) / H


def grad_perp(r: RQGSetup, psi):
    """(-dpsi/dy on U faces, dpsi/dx on V faces), backward differences."""
    psi = np.asarray(psi, dtype=np.float64)
    u = -psi.copy()
    u[..., 1:, :] += psi[..., :-1, :]
    v = psi.copy()
    v[..., :, 1:] -= psi[..., :, :-1]
    return u / r.swm.dy, v / r.swm.dx


def geostrophic(r: RQGSetup, psi):
    """G: psi -> (h_g, u_g, v_g), h_g = H (1 + f0 M psi)."""
    u, v = grad_perp(r, psi)
    H = r.swm.H[:, None, None]
    h = H * (1.0 + r.f0 * np.einsum("kl,l...->k...", r.M, psi))
    return h, u, v


def project(r: RQGSetup, h, u, v, return_psi=False):
    """P = G (Q G)^-1 Q (reparameterized.py:142-178)."""
    psi = solve_helmholtz(r, pv(r, h, u, v))
    out = geostrophic(r, psi)
    return (out, psi) if return_psi else out


def apply_bc(r: RQGSetup, h, u, v):
    """Shallow-water boundary condition, then the projection (:186-189)."""
    return project(r, *swm.apply_bc(r.swm, h, u, v))


def rhs_of_bc(r: RQGSetup, y):
    """f(y) = vector_field(BC(y)) as ``build_terms`` composes them."""
    return swm.rhs(r.swm, *apply_bc(r, *y))


def step(r: RQGSetup, h, u, v, dt, solver="tsit5"):
    """One Tsit5 (default) or SSP-RK3 step of the projected system."""
    f = lambda y: rhs_of_bc(r, y)  # noqa: E731
    if solver == "tsit5":
        return rk.explicit_step(f, (h, u, v), dt)
    a = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.25, 0.25, 0.0]])
    return rk.explicit_step(f, (h, u, v), dt, a=a, b=np.array([1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0]))


def diagnose(r: RQGSetup, h, u, v):
    """diagnose (:191-222): the shallow-water diagnostics plus psi and the
    ageostrophic velocity (u - u_g, v - v_g)."""
    d = swm.diagnose(r.swm, h, u, v)
    psi = solve_helmholtz(r, pv(r, h, u, v))
    ug, vg = grad_perp(r, psi)
    d.update(psi=psi, u_ageostrophic=np.asarray(u) - ug, v_ageostrophic=np.asarray(v) - vg)
    return d
