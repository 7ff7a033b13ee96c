"""FP64 restatement of the multilayer shallow-water model (test infrastructure).

Follows ``somax/_src/models/swm/multilayer.py``: ``vector_field`` (:150-201),
``apply_boundary_conditions`` (:203-223) with the wall helpers (:383-410),
``diagnose`` (:225-255) and ``create`` (:259-377).  The arithmetic lives in
finitevolX v0.0.39 (an un-vendored git dependency that is not installed), so
the staggering and stencils below are this restatement's stated convention:

* C-grid by index offset on arrays of one shape ``(Ny, Nx)``: ``h[j, i]`` at
  the T point (cell centre), ``u[j, i]`` on its east face, ``v[j, i]`` on its
  north face, X (corner) point ``[j, i]`` at its north-east corner.  T row j
  sits at ``y = j*dy`` (grid pin of the QG models).
* Every intermediate field (PV at X, mass fluxes, kinetic energy, averages)
  is evaluated wherever its stencil lies inside the array; tendencies are
  written on the interior ``[1:-1, 1:-1]`` of each field and are zero on the
  ghost ring, except the wind (top layer) and bottom drag (bottom layer)
  terms, which the reference adds to the full arrays (:190-199).
* Averages are arithmetic means of the nearest points; kinetic energy at T
  is ``(u_w^2 + u_e^2 + v_s^2 + v_n^2) / 4`` (mean of squares); PV at X is
  ``(zeta + f_X) / h_X`` with ``f_X`` and ``h_X`` the 4-point T means;
  ``zeta = dv/dx - du/dy`` by one-sided differences around the corner.
* Mass flux ``Advection2D(method="upwind1")``: face value of h from the
  upwind cell (left/lower for positive velocity).
* ``Diffusion2D(f, nu)``: ``nu`` times the 5-point Laplacian.
* Periodic ghost copy (``enforce_periodic``): ghost row/column 0 takes the
  last interior row/column and ghost N-1 the first.

These conventions are "parity unpinned" against the reference: the reference's
own tests pin only the rest state (zero tendency < 1e-10), forcing locality
and the wall values (tests/models/test_swm_multilayer.py:94-175), which the
checks in tests/test_oracle_swm.py reproduce; invariants (mass conservation,
free-stream preservation, antisymmetry of the Coriolis/PV term's work) pin
the arithmetic further.

``bc="channel"`` (periodic in x, walls in y) is an extension for the
baroclinic-jet benchmark: the periodic copy on columns, the wall rules on
rows.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class SWMSetup:
    nx: int
    ny: int
    Lx: float
    Ly: float
    nl: int
    g_prime: np.ndarray
    H: np.ndarray
    f_row: np.ndarray  # (Ny,) Coriolis parameter at T rows
    wind_row: np.ndarray  # (Ny,) normalised x wind-stress pattern at T rows
    nu: float = 0.0
    kappa: float = 0.0
    tau0: float = 0.0
    bc: str = "periodic"
    vorticity_flux: str = "average"  # or "energy_enstrophy" (extension)

    @property
    def Nx(self):
        return self.nx + 2

    @property
    def Ny(self):
        return self.ny + 2

    @property
    def dx(self):
        return self.Lx / self.nx

    @property
    def dy(self):
        return self.Ly / self.ny

    @property
    def shape(self):
        return (self.nl, self.Ny, self.Nx)


def setup(nx=64, ny=64, Lx=4e6, Ly=4e6, f0=9.375e-5, beta=1.754e-11, H=(400.0, 1100.0, 2600.0),
          g_prime=(9.81, 0.025, 0.0125), lateral_viscosity=0.0, bottom_drag=0.0, wind_amplitude=0.0,
          wind_profile="doublegyre", bc="periodic", vorticity_flux="average") -> SWMSetup:
    """``MultilayerShallowWater2D.create`` (multilayer.py:259-377) in FP64:
    ``f = f0 + beta (y - Ly/2)`` (:345-349), ``tau_x = -cos(2 pi y / Ly)``
    (doublegyre) or ``-cos(pi y / Ly)`` (single) (:352-356).
    ``vorticity_flux``: "average" (the reference's, multilayer.py:172-187) or
    "energy_enstrophy" (extension, ``vorticity_flux_terms``)."""
    H = np.asarray(H, dtype=np.float64)
    gp = np.asarray(g_prime, dtype=np.float64)
    if H.size != gp.size:
        raise ValueError("H and g_prime must have the same length")
    Ny = ny + 2
    dy = Ly / ny
    y = np.arange(Ny, dtype=np.float64) * dy
    f_row = f0 + beta * (y - Ly / 2.0)
    if wind_profile == "single":
        wind = -np.cos(np.pi * y / Ly)
    else:
        wind = -np.cos(2.0 * np.pi * y / Ly)
    if bc not in ("periodic", "wall", "channel"):
        raise ValueError(f"unknown bc {bc!r}")
    if vorticity_flux not in ("average", "energy_enstrophy"):
        raise ValueError(f"unknown vorticity_flux {vorticity_flux!r}")
    return SWMSetup(nx=nx, ny=ny, Lx=Lx, Ly=Ly, nl=H.size, g_prime=gp, H=H, f_row=f_row, wind_row=wind,
                    nu=lateral_viscosity, kappa=bottom_drag, tau0=wind_amplitude, bc=bc,
                    vorticity_flux=vorticity_flux)


# ---------------------------------------------------------------------------
# boundary conditions (multilayer.py:203-223, 383-410)
# ---------------------------------------------------------------------------
def _periodic_cols(f):
    f = f.copy()
    f[..., :, 0] = f[..., :, -2]
    f[..., :, -1] = f[..., :, 1]
    return f


def _periodic_rows(f):
    f = f.copy()
    f[..., 0, :] = f[..., -2, :]
    f[..., -1, :] = f[..., 1, :]
    return f


def enforce_periodic(f):
    """Ghost ring from the opposite interior row/column."""
    return _periodic_cols(_periodic_rows(np.asarray(f, dtype=np.float64)))


def wall_bc_u(u):
    """multilayer.py:386-392: u = 0 on columns 0, -2, -1; rows 0/-1 copy rows 1/-2."""
    u = np.array(u, dtype=np.float64)
    u[..., :, 0] = 0.0
    u[..., :, -2] = 0.0
    u[..., :, -1] = 0.0
    u[..., 0, :] = u[..., 1, :]
    u[..., -1, :] = u[..., -2, :]
    return u


def wall_bc_v(v):
    """multilayer.py:395-401: v = 0 on rows 0, -2, -1; columns 0/-1 copy 1/-2."""
    v = np.array(v, dtype=np.float64)
    v[..., 0, :] = 0.0
    v[..., -2, :] = 0.0
    v[..., -1, :] = 0.0
    v[..., :, 0] = v[..., :, 1]
    v[..., :, -1] = v[..., :, -2]
    return v


def wall_bc_h(h):
    """multilayer.py:404-408: zero gradient on all walls."""
    h = np.array(h, dtype=np.float64)
    h[..., 0, :] = h[..., 1, :]
    h[..., -1, :] = h[..., -2, :]
    h[..., :, 0] = h[..., :, 1]
    h[..., :, -1] = h[..., :, -2]
    return h


def apply_bc(s: SWMSetup, h, u, v):
    if s.bc == "periodic":
        return enforce_periodic(h), enforce_periodic(u), enforce_periodic(v)
    if s.bc == "wall":
        return wall_bc_h(h), wall_bc_u(u), wall_bc_v(v)
    # channel: periodic columns, wall rows
    h = np.array(h, dtype=np.float64)
    h[..., 0, :] = h[..., 1, :]
    h[..., -1, :] = h[..., -2, :]
    u = np.array(u, dtype=np.float64)
    u[..., 0, :] = u[..., 1, :]
    u[..., -1, :] = u[..., -2, :]
    v = np.array(v, dtype=np.float64)
    v[..., 0, :] = 0.0
    v[..., -2, :] = 0.0
    v[..., -1, :] = 0.0
    return _periodic_cols(h), _periodic_cols(u), _periodic_cols(v)


# ---------------------------------------------------------------------------
# intermediate fields (each where its stencil lies inside the array)
# ---------------------------------------------------------------------------
def relative_vorticity(s: SWMSetup, u, v):
    """zeta at X points [j, i], j < Ny-1, i < Nx-1; 0 elsewhere."""
    z = np.zeros(np.shape(u))
    z[..., :-1, :-1] = (v[..., :-1, 1:] - v[..., :-1, :-1]) / s.dx - (u[..., 1:, :-1] - u[..., :-1, :-1]) / s.dy
    return z


def t_to_x(f):
    out = np.zeros(np.shape(f))
    out[..., :-1, :-1] = 0.25 * (f[..., :-1, :-1] + f[..., :-1, 1:] + f[..., 1:, :-1] + f[..., 1:, 1:])
    return out


def potential_vorticity(s: SWMSetup, u, v, h):
    """q = (zeta + f_X) / h_X at X points (multilayer.py:164)."""
    q = np.zeros(np.shape(u))
    fX = 0.5 * (s.f_row[:-1] + s.f_row[1:])[:, None]
    hX = t_to_x(h)[..., :-1, :-1]
    zeta = relative_vorticity(s, u, v)[..., :-1, :-1]
    q[..., :-1, :-1] = (zeta + fX) / hX
    return q


def kinetic_energy(u, v):
    """KE at T points [j, i], j >= 1, i >= 1 (mean of the four face squares)."""
    ke = np.zeros(np.shape(u))
    ke[..., 1:, 1:] = 0.25 * (u[..., 1:, 1:] ** 2 + u[..., 1:, :-1] ** 2 + v[..., 1:, 1:] ** 2 + v[..., :-1, 1:] ** 2)
    return ke


def _interior(a):
    return a[..., 1:-1, 1:-1]


def extend_pv(s: SWMSetup, q):
    """PV on the X points of the whole array.  ``potential_vorticity`` fills
    rows and columns 0..N-2; the last X column (row) is the periodic image of
    column (row) 1 where x (y) is periodic, and 0 beside a wall."""
    q = np.array(q, dtype=np.float64)
    if s.bc in ("periodic", "channel"):
        q[..., :, -1] = q[..., :, 1]
    if s.bc == "periodic":
        q[..., -1, :] = q[..., 1, :]
    return q


def vorticity_flux_terms(s: SWMSetup, q, uh, vh):
    """The PV-flux part of the momentum tendency on interior U and V points:
    ``(+ q (vh) at U, - q (uh) at V)``.

    * ``"average"`` (the reference, multilayer.py:172-187): the two X values
      beside the face times the four-point average of the other flux.
    * ``"energy_enstrophy"`` (extension): the triad form of Arakawa & Lamb
      (1981) as written for a C grid with PV at the corners (the "EEN" form):
      each T cell carries four triads, the sum of its corner values except the
      one opposite a corner, /12; a U face takes each of its four neighbouring
      V fluxes with the triad of the cell that holds both, at the corner they
      share, and a V face the U fluxes likewise, with the opposite sign.  Both
      equations use the same triad for each (U, V) pair, so the term does no
      work (sum U F_u + V F_v = 0); for a non-divergent mass flux it also
      leaves the potential enstrophy sum h_X q^2 / 2 unchanged.  With uniform
      q it equals the average form.

    T cell [j, i] has the corners NE = X[j, i], NW = X[j, i-1], SE = X[j-1, i],
    SW = X[j-1, i-1]; U[j, i] is its east face and V[j, i] its north face."""
    if s.vorticity_flux == "average":
        qU = 0.5 * (q[..., 1:-1, 1:-1] + q[..., :-2, 1:-1])
        vhU = 0.25 * (vh[..., 1:-1, 1:-1] + vh[..., 1:-1, 2:] + vh[..., :-2, 1:-1] + vh[..., :-2, 2:])
        qV = 0.5 * (q[..., 1:-1, 1:-1] + q[..., 1:-1, :-2])
        uhV = 0.25 * (uh[..., 1:-1, 1:-1] + uh[..., 1:-1, :-2] + uh[..., 2:, 1:-1] + uh[..., 2:, :-2])
        return qU * vhU, -qV * uhV
    q = extend_pv(s, q)
    # triads of every T cell [j, i], j, i >= 1 (array index [j-1, i-1])
    ne, nw, se, sw = q[..., 1:, 1:], q[..., 1:, :-1], q[..., :-1, 1:], q[..., :-1, :-1]
    t_ne = (nw + ne + se) / 12.0
    t_nw = (sw + nw + ne) / 12.0
    t_se = (ne + se + sw) / 12.0
    t_sw = (se + sw + nw) / 12.0

    def T(t, dj, di):  # triad of cell [j+dj, i+di] for interior [j, i]
        ny, nx = t.shape[-2:]
        return t[..., dj : ny - 1 + dj, di : nx - 1 + di]

    def F(f, dj, di):  # flux at [j+dj, i+di] for interior [j, i]
        ny, nx = f.shape[-2:]
        return f[..., 1 + dj : ny - 1 + dj, 1 + di : nx - 1 + di]

    fu = (T(t_ne, 0, 0) * F(vh, 0, 0) + T(t_nw, 0, 1) * F(vh, 0, 1) + T(t_se, 0, 0) * F(vh, -1, 0)
          + T(t_sw, 0, 1) * F(vh, -1, 1))
    fv = -(T(t_sw, 1, 0) * F(uh, 1, -1) + T(t_se, 1, 0) * F(uh, 1, 0) + T(t_nw, 0, 0) * F(uh, 0, -1)
           + T(t_ne, 0, 0) * F(uh, 0, 0))
    return fu, fv


def rhs(s: SWMSetup, h, u, v):
    """``vector_field`` (multilayer.py:150-201) on the state as given."""
    h = np.asarray(h, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    dx, dy = s.dx, s.dy
    # 1. mass: upwind fluxes on every face between two cells of the array
    F = np.zeros_like(u)
    F[..., :, :-1] = u[..., :, :-1] * np.where(u[..., :, :-1] > 0, h[..., :, :-1], h[..., :, 1:])
    G = np.zeros_like(v)
    G[..., :-1, :] = v[..., :-1, :] * np.where(v[..., :-1, :] > 0, h[..., :-1, :], h[..., 1:, :])
    dh = np.zeros_like(h)
    dh[..., 1:-1, 1:-1] = -((F[..., 1:-1, 1:-1] - F[..., 1:-1, :-2]) / dx + (G[..., 1:-1, 1:-1] - G[..., :-2, 1:-1]) / dy)
    # 2-3. PV at X and mass fluxes at U / V
    q = potential_vorticity(s, u, v, h)
    uh = np.zeros_like(u)
    uh[..., :, :-1] = 0.5 * (h[..., :, :-1] + h[..., :, 1:]) * u[..., :, :-1]
    vh = np.zeros_like(v)
    vh[..., :-1, :] = 0.5 * (h[..., :-1, :] + h[..., 1:, :]) * v[..., :-1, :]
    # 4. Bernoulli potential: KE + cumulative g' h (multilayer.py:178-183)
    P = kinetic_energy(u, v) + np.cumsum(s.g_prime[:, None, None] * h, axis=0)
    # 5. momentum on interior U / V points
    fu, fv = vorticity_flux_terms(s, q, uh, vh)
    du = np.zeros_like(u)
    dv = np.zeros_like(v)
    du[..., 1:-1, 1:-1] = fu - (P[..., 1:-1, 2:] - P[..., 1:-1, 1:-1]) / dx
    dv[..., 1:-1, 1:-1] = fv - (P[..., 2:, 1:-1] - P[..., 1:-1, 1:-1]) / dy
    # 6. wind on the top layer, full array (multilayer.py:190-191)
    du[0] += s.tau0 * s.wind_row[:, None] / s.H[0]
    # 7. diffusion (all layers)
    if s.nu != 0.0:
        for f, d in ((u, du), (v, dv)):
            d[..., 1:-1, 1:-1] += s.nu * ((f[..., 1:-1, 2:] - 2 * _interior(f) + f[..., 1:-1, :-2]) / dx**2
                                          + (f[..., 2:, 1:-1] - 2 * _interior(f) + f[..., :-2, 1:-1]) / dy**2)
    # 8. bottom drag on the bottom layer, full array (multilayer.py:198-199)
    du[-1] += -s.kappa * u[-1]
    dv[-1] += -s.kappa * v[-1]
    return dh, du, dv


def ssprk3_step(s: SWMSetup, h, u, v, dt):
    """One SSP-RK3 step (Shu-Osher form) with the BC applied before each RHS,
    as ``SomaxModel.build_terms`` composes them (core/model.py:47-49)."""
    y0 = [np.asarray(a, dtype=np.float64) for a in (h, u, v)]

    def f(y):
        return rhs(s, *apply_bc(s, *y))

    k = f(y0)
    y1 = [a + dt * b for a, b in zip(y0, k)]
    k = f(y1)
    y2 = [0.75 * a + 0.25 * (b + dt * c) for a, b, c in zip(y0, y1, k)]
    k = f(y2)
    return tuple(a / 3.0 + 2.0 / 3.0 * (b + dt * c) for a, b, c in zip(y0, y2, k))


def tsit5_step(s: SWMSetup, h, u, v, dt):
    """One Tsit5 step (``oracle/rk.py``; ``integrate``'s default solver,
    core/model.py:75) with the BC applied before each RHS."""
    from . import rk

    return rk.explicit_step(lambda y: rhs(s, *apply_bc(s, *y)), (h, u, v), dt)


def run(s: SWMSetup, h, u, v, dt, nsteps):
    y = (h, u, v)
    for _ in range(nsteps):
        y = ssprk3_step(s, *y, dt)
    return y


def diagnose(s: SWMSetup, h, u, v):
    """``diagnose`` (multilayer.py:225-255): per-layer energy (KE + 0.5 g' h^2)
    and potential enstrophy 0.5 q^2 h_X, summed over the interior x dx dy."""
    h = np.asarray(h, dtype=np.float64)
    ke = kinetic_energy(u, v)
    q = potential_vorticity(s, u, v, h)
    hX = t_to_x(h)
    area = s.dx * s.dy
    energy = (_interior(ke).sum(axis=(-2, -1)) + 0.5 * s.g_prime * (_interior(h) ** 2).sum(axis=(-2, -1))) * area
    enstrophy = 0.5 * (_interior(q) ** 2 * _interior(hX)).sum(axis=(-2, -1)) * area
    return {"energy": energy, "total_energy": energy.sum(), "enstrophy": enstrophy,
            "total_enstrophy": enstrophy.sum(), "potential_vorticity": q,
            "relative_vorticity": relative_vorticity(s, u, v), "kinetic_energy_field": ke}


# ---------------------------------------------------------------------------
# single-layer NonlinearShallowWater2D (swm/nonlinear_2d.py)
# ---------------------------------------------------------------------------
def nonlinear_setup(nx=64, ny=64, Lx=1e6, Ly=1e6, g=9.81, f0=1e-4, beta=0.0, H0=100.0, lateral_viscosity=0.0,
                    bottom_drag=0.0, wind_amplitude=0.0, wind_profile="doublegyre", bc="periodic") -> SWMSetup:
    """``NonlinearShallowWater2D.create`` (nonlinear_2d.py:243-318): the
    vector-invariant equations of ``vector_field`` (:132-194) are the
    multilayer ones at one layer with g' = g (P = KE + g h, drag on the only
    layer), except that the wind stress is *not* divided by the depth
    (:180-182 against multilayer.py:190-191).  The setup therefore holds the
    one-layer model with the wind amplitude multiplied by H0, which the
    multilayer RHS divides out again."""
    if bc not in ("periodic", "wall"):
        raise ValueError(f"unknown bc {bc!r}")
    return setup(nx=nx, ny=ny, Lx=Lx, Ly=Ly, f0=f0, beta=beta, H=(H0,), g_prime=(g,),
                 lateral_viscosity=lateral_viscosity, bottom_drag=bottom_drag, wind_amplitude=wind_amplitude * H0,
                 wind_profile=wind_profile, bc=bc)


def nonlinear_diagnose(s: SWMSetup, h, u, v):
    """``diagnose`` (nonlinear_2d.py:219-241): scalar energy and enstrophy."""
    d = diagnose(s, *(np.asarray(a, dtype=np.float64)[None] for a in (h, u, v)))
    return {"energy": d["energy"][0], "enstrophy": d["enstrophy"][0],
            "potential_vorticity": d["potential_vorticity"][0], "relative_vorticity": d["relative_vorticity"][0],
            "kinetic_energy_field": d["kinetic_energy_field"][0]}
