"""Test-case factories and synthetic initial conditions.

``doublegyre_qg`` and ``doublegyre_baroclinic_qg`` mirror
``somax/_src/models/gfd_testcases.py:184-287`` (same defaults; the reference
starts from rest).  ``random_streamfunction`` and the ``config*`` builders are
extensions for the benchmark workloads: seeded isotropic k^-3 random fields
and the basin/forcing parameters of the benchmark configurations.  All
set-up arithmetic is FP64 NumPy on the host; the returned states are float32
device tensors.
"""

from __future__ import annotations

import numpy as np

from .qg import BaroclinicQG, BaroclinicQGState, BarotropicQG, BarotropicQGState

RHO0 = 1000.0
BASIN = 5.12e6
F0 = 9.375e-5
BETA = 1.754e-11
TAU0 = 0.08
DRAG = 3.6e-8
NU4_128 = 2e11  # biharmonic viscosity at 128^2, scaled by (128/n)^4


def _torch():
    import torch

    return torch


def doublegyre_qg(nx=64, ny=64, Lx=1e6, Ly=1e6, f0=1e-4, beta=1.6e-11, lateral_viscosity=500.0,
                  bottom_drag=1e-7, wind_amplitude=1e-12, **kw):
    """gfd_testcases.py:184-228: barotropic double gyre from rest."""
    model = BarotropicQG.create(nx=nx, ny=ny, Lx=Lx, Ly=Ly, f0=f0, beta=beta, lateral_viscosity=lateral_viscosity,
                                bottom_drag=bottom_drag, wind_amplitude=wind_amplitude, wind_profile="doublegyre",
                                **kw)
    torch = _torch()
    q0 = torch.zeros((model.grid.Ny, model.grid.Nx), dtype=torch.float32, device=model.device)
    return model, BarotropicQGState(q=q0)


def doublegyre_baroclinic_qg(nx=128, ny=128, Lx=4e6, Ly=4e6, f0=9.375e-5, beta=1.754e-11, n_layers=3,
                             H=(400.0, 1100.0, 2600.0), g_prime=(9.81, 0.025, 0.0125), lateral_viscosity=15.0,
                             bottom_drag=1e-7, wind_amplitude=1.3e-10, **kw):
    """gfd_testcases.py:231-287: multilayer double gyre from rest."""
    model = BaroclinicQG.create(nx=nx, ny=ny, Lx=Lx, Ly=Ly, f0=f0, beta=beta, n_layers=n_layers, H=H,
                                g_prime=g_prime, lateral_viscosity=lateral_viscosity, bottom_drag=bottom_drag,
                                wind_amplitude=wind_amplitude, wind_profile="doublegyre", **kw)
    torch = _torch()
    q0 = torch.zeros((model.consts.n_layers, model.grid.Ny, model.grid.Nx), dtype=torch.float32,
                     device=model.device)
    return model, BaroclinicQGState(q=q0)


# ---------------------------------------------------------------------------
# synthetic initial conditions (extension)
# ---------------------------------------------------------------------------
def _sine(n: int) -> np.ndarray:
    j = np.arange(1, n + 1, dtype=np.float64)
    return np.sqrt(2.0 / (n + 1)) * np.sin(np.pi * np.outer(j, j) / (n + 1))


def random_streamfunction(nx: int, ny: int, Lx: float, Ly: float, nl: int, seed: int, urms: float = 0.05,
                          slope: float = -3.0, zero_mean: bool = False) -> np.ndarray:
    """Seeded Gaussian random psi, ``(nl, ny+2, nx+2)`` FP64, zero ghost ring.

    Built in the interior DST-I basis (so psi = 0 on the ghost ring) with an
    isotropic kinetic-energy spectrum ``E(k) ~ k**slope``: the coefficient of
    mode ``(ky, kx)`` has standard deviation ``k**((slope - 3)/2)``.  Each
    layer is scaled to an rms speed ``urms`` of the forward-difference
    geostrophic velocity ``u = -dpsi/dy``, ``v = dpsi/dx`` on the interior.

    ``zero_mean`` removes each layer's mean over the array with the gravest
    basin mode ``sin(pi x / Lx) sin(pi y / Ly)`` (also 0 on the ghost ring): a
    state the mass-conservation correction leaves unchanged, as a spun-up
    basin is, instead of one it turns into a grid-scale wall current.
    """
    rng = np.random.default_rng(seed)
    dx, dy = Lx / nx, Ly / ny
    kx = np.pi * np.arange(1, nx + 1) / ((nx + 1) * dx)
    ky = np.pi * np.arange(1, ny + 1) / ((ny + 1) * dy)
    K = np.sqrt(ky[:, None] ** 2 + kx[None, :] ** 2)
    amp = K ** ((slope - 3.0) / 2.0)
    Sx, Sy = _sine(nx), _sine(ny)
    psi = np.zeros((nl, ny + 2, nx + 2))
    for l in range(nl):
        c = rng.standard_normal((ny, nx)) * amp
        p = np.zeros((ny + 2, nx + 2))
        p[1:-1, 1:-1] = Sy @ c @ Sx
        if zero_mean:
            g = np.zeros((ny + 2, nx + 2))
            g[1:-1, 1:-1] = np.outer(Sy[:, 0], Sx[0, :])
            p -= (p.mean() / g.mean()) * g
        u = -(p[2:, 1:-1] - p[1:-1, 1:-1]) / dy
        v = (p[1:-1, 2:] - p[1:-1, 1:-1]) / dx
        s = np.sqrt(np.mean(u * u + v * v))
        psi[l] = p * (urms / s)
    return psi


def pv_from_streamfunction(psi: np.ndarray, dx: float, dy: float, lambdas, Cl2m, Cm2l) -> np.ndarray:
    """``q = lap(psi) - B psi`` with ``B = Cm2l diag(lambdas) Cl2m`` (= f0^2 A for
    consistent eigenvalues); 5-point Laplacian on the interior, ghost ring 0."""
    psi = np.asarray(psi, dtype=np.float64)
    q = np.zeros_like(psi)
    c = psi[:, 1:-1, 1:-1]
    q[:, 1:-1, 1:-1] = (psi[:, 1:-1, 2:] - 2 * c + psi[:, 1:-1, :-2]) / dx**2 + (
        psi[:, 2:, 1:-1] - 2 * c + psi[:, :-2, 1:-1]
    ) / dy**2
    B = np.asarray(Cm2l) @ np.diag(np.asarray(lambdas)) @ np.asarray(Cl2m)
    q[:, 1:-1, 1:-1] -= np.einsum("lk,kyx->lyx", B, c)
    return q


def random_pv(model, seed: int, urms: float = 0.05, zero_mean: bool = False) -> np.ndarray:
    """FP64 PV field whose inversion is a seeded k^-3 random streamfunction."""
    g = model.grid
    psi = random_streamfunction(g.nx, g.ny, g.Lx, g.Ly, model._nl, seed, urms, zero_mean=zero_mean)
    return pv_from_streamfunction(psi, g.dx, g.dy, model._lambdas, model._Cl2m, model._Cm2l)


# ---------------------------------------------------------------------------
# benchmark configurations (basin 5120 km, tau0 = 0.08 N/m^2, drag 3.6e-8 /s)
# ---------------------------------------------------------------------------
def wind_amplitude(tau0: float = TAU0, Ly: float = BASIN) -> float:
    """``tau0 * 2 pi / (rho0 Ly)`` (gfd_testcases.py:250-251)."""
    return tau0 * 2.0 * np.pi / (RHO0 * Ly)


def config_barotropic(n: int = 64, depth: float = 4000.0, wind_tau0: float = TAU0, **kw) -> BarotropicQG:
    """1-layer double gyre: DST-I Poisson inversion, biharmonic 2e11 m^4/s."""
    return BarotropicQG.create(nx=n, ny=n, Lx=BASIN, Ly=BASIN, f0=F0, beta=BETA, bottom_drag=DRAG,
                               wind_amplitude=wind_amplitude(wind_tau0) / depth, biharmonic_viscosity=NU4_128, **kw)


def config_baroclinic(n: int = 128, nl: int = 3, tau0: float = TAU0, **kw) -> BaroclinicQG:
    """Multilayer double gyre: H = 400/1100/2600 m, g' = 9.81/0.025/0.0125,
    biharmonic 2e11 * (128/n)^4 m^4/s, bottom drag on the lowest layer."""
    H = (400.0, 1100.0, 2600.0)[:nl]
    gp = (9.81, 0.025, 0.0125)[:nl]
    return BaroclinicQG.create(nx=n, ny=n, Lx=BASIN, Ly=BASIN, f0=F0, beta=BETA, n_layers=nl, H=H, g_prime=gp,
                               bottom_drag=DRAG, wind_amplitude=wind_amplitude(tau0),
                               biharmonic_viscosity=NU4_128 * (128.0 / n) ** 4, **kw)


# ---------------------------------------------------------------------------
# ensemble of uneven members (extension)
# ---------------------------------------------------------------------------
# Layer thicknesses and reduced gravities per layer count, each 4000 m deep;
# the 3-layer set is the multilayer double gyre's.
ENSEMBLE_LAYERS = {
    1: ((4000.0,), (9.81,)),
    2: ((400.0, 3600.0), (9.81, 0.025)),
    3: ((400.0, 1100.0, 2600.0), (9.81, 0.025, 0.0125)),
}
# (grid size, members) of the 256-member ensemble
ENSEMBLE_COUNTS = ((64, 96), (128, 96), (256, 48), (512, 16))


def ensemble_member_args(nx: int, nl: int, tau0: float, ny: int | None = None) -> dict:
    """``BaroclinicQG.create`` keyword arguments of one ensemble member: the
    double-gyre basin (square cells, Ly scaled with ny), biharmonic
    2e11 * (128/nx)^4 m^4/s, bottom drag 3.6e-8 /s, wind stress tau0."""
    ny = nx if ny is None else ny
    H, gp = ENSEMBLE_LAYERS[nl]
    return dict(nx=nx, ny=ny, Lx=BASIN, Ly=BASIN * ny / nx, f0=F0, beta=BETA, n_layers=nl, H=H, g_prime=gp,
                bottom_drag=DRAG, wind_amplitude=wind_amplitude(tau0, BASIN * ny / nx),
                biharmonic_viscosity=NU4_128 * (128.0 / nx) ** 4)


def ensemble_member(nx: int, nl: int, tau0: float, ny: int | None = None, **kw) -> BaroclinicQG:
    return BaroclinicQG.create(**ensemble_member_args(nx, nl, tau0, ny), **kw)


def config_ensemble(seed: int = 0, counts=ENSEMBLE_COUNTS, precision: str = "mixed", advection: str = "weno",
                    mass_correction: bool = True, **kw):
    """The uneven ensemble: ``counts`` = (grid size, members) pairs; layer
    counts cycle 1/2/3 over the members in order; tau0 ~ U[0.04, 0.12] N/m^2
    and a seeded k^-3 random initial streamfunction per member (zero layer
    means when the mass constraint is on).  The physics defaults to the
    multilayer double gyre's: WENO-Z PV flux and the FP64 mass constraint.
    Returns ``(models, member PV fields as float32 host arrays)``."""
    rng = np.random.default_rng(seed)
    models, qs = [], []
    i = 0
    for n, count in counts:
        for _ in range(count):
            nl = 1 + i % 3
            tau0 = float(rng.uniform(0.04, 0.12))
            m = ensemble_member(n, nl, tau0, precision=precision, advection=advection,
                                mass_correction=mass_correction, **kw)
            models.append(m)
            qs.append(random_pv(m, seed * 100003 + i, zero_mean=mass_correction).astype(np.float32))
            i += 1
    return models, qs


# ---------------------------------------------------------------------------
# multilayer shallow water
# ---------------------------------------------------------------------------
def baroclinic_instability_swm(nx=64, ny=64, Lx=1e6, Ly=1e6, f0=1e-4, beta=1.6e-11, H=(500.0, 4500.0),
                               g_prime=(9.81, 0.025), lateral_viscosity=100.0, bottom_drag=1e-7, jet_speed=0.5,
                               jet_width=5e4, perturbation=0.01, **kw):
    """gfd_testcases.py:290-362: opposite Gaussian jets in the two layers of a
    periodic domain, ``v`` perturbed by ``sin(4 pi x / Lx)`` on the jet."""
    from .swm import MultilayerShallowWater2D, MultilayerSW2DState

    model = MultilayerShallowWater2D.create(nx=nx, ny=ny, Lx=Lx, Ly=Ly, f0=f0, beta=beta, n_layers=2, H=H,
                                            g_prime=g_prime, lateral_viscosity=lateral_viscosity,
                                            bottom_drag=bottom_drag, bc="periodic", **kw)
    g = model.grid
    x = np.arange(g.Nx) * g.dx
    y = np.arange(g.Ny) * g.dy
    X, Y = np.meshgrid(x, y)
    jet = np.exp(-0.5 * ((Y - Ly / 2.0) / jet_width) ** 2)
    signs = np.array([1.0, -1.0])
    u0 = signs[:, None, None] * jet_speed * jet[None]
    v0 = np.broadcast_to(perturbation * np.sin(4.0 * np.pi * X / Lx)[None] * jet[None], (2, g.Ny, g.Nx)).copy()
    h0 = np.ones((2, g.Ny, g.Nx)) * np.asarray(model.strat.H)[:, None, None]
    torch = _torch()
    dev = model.device
    state = MultilayerSW2DState(h=torch.tensor(h0, dtype=torch.float32, device=dev),
                                u=torch.tensor(u0, dtype=torch.float32, device=dev),
                                v=torch.tensor(v0, dtype=torch.float32, device=dev))
    return model, state


# The 2-layer baroclinic jet: f-plane, H = 500/3500 m, g' = 9.81/0.02, a
# square channel of 2048 km (periodic in x, walls in y).
JET_L = 2.048e6
JET_F0 = 1e-4


JET_H = (500.0, 3500.0)
JET_GP = (9.81, 0.02)


def baroclinic_jet_fields(n: int = 512, seed: int = 0, jet_speed: float = 0.5, jet_width: float = 2e5,
                          noise: float = 1.0, L: float = JET_L):
    """Host (h, u, v) FP64 arrays (2, n+2, n+2) of the baroclinic jet on an
    n x n channel of side L: a geostrophically balanced zonal jet in the
    upper layer, lower layer at rest, plus seeded white noise of amplitude
    ``noise`` metres on the interface (added to h_0, taken from h_1).  Balance
    with the cumulative pressure ``P_k = sum_{j<=k} g'_j h_j``:
    ``f u_0 = -g'_0 dh_0/dy`` and ``g'_1 dh_1/dy = -g'_0 dh_0/dy`` (u_1 = 0),
    integrated from the channel centre."""
    from math import erf, pi, sqrt

    H, gp = JET_H, JET_GP
    Ny = Nx = n + 2
    y = np.arange(Ny) * (L / n)
    yc = L / 2.0
    s = (y - yc) / jet_width
    ujet = jet_speed * np.exp(-0.5 * s**2)
    # integral of the jet from the centre: U w sqrt(pi/2) erf(s / sqrt 2)
    integ = jet_speed * jet_width * sqrt(pi / 2.0) * np.array([erf(v / sqrt(2.0)) for v in s])
    dh0 = -JET_F0 / gp[0] * integ
    dh1 = -gp[0] / gp[1] * dh0
    rng = np.random.default_rng(seed)
    eta = noise * (2.0 * rng.random((Ny, Nx)) - 1.0)
    h = np.empty((2, Ny, Nx))
    h[0] = H[0] + dh0[:, None] + eta
    h[1] = H[1] + dh1[:, None] - eta
    u = np.zeros((2, Ny, Nx))
    u[0] = ujet[:, None]
    v = np.zeros((2, Ny, Nx))
    return h, u, v


def baroclinic_jet_swm(n: int = 512, seed: int = 0, jet_speed: float = 0.5, jet_width: float = 2e5,
                       noise: float = 1.0, L: float = JET_L, **kw):
    """The baroclinic jet's channel model (f-plane, H = 500/3500 m,
    g' = 9.81/0.02) and its initial state (``baroclinic_jet_fields``).
    Returns ``(model, host (h, u, v) FP64 arrays)``."""
    from .swm import MultilayerShallowWater2D

    model = MultilayerShallowWater2D.create(nx=n, ny=n, Lx=L, Ly=L, f0=JET_F0, beta=0.0, n_layers=2, H=JET_H,
                                            g_prime=JET_GP, bc="channel", **kw)
    return model, baroclinic_jet_fields(n, seed, jet_speed, jet_width, noise, L)


def cfl_dt(H, g_prime, dx: float, dy: float, cfl: float = 0.5) -> float:
    """CFL-limited step: dt = cfl * min(dx, dy) / (max gravity-wave speed + 0.5 m/s),
    wave speeds from the eigenvalues of diag(H) times the cumulative pressure matrix."""
    H = np.asarray(H, dtype=np.float64)
    gp = np.asarray(g_prime, dtype=np.float64)
    M = np.tril(np.ones((H.size, H.size))) * gp[None, :]
    c = float(np.sqrt(np.max(np.linalg.eigvals(np.diag(H) @ M).real)))
    return cfl * min(dx, dy) / (c + 0.5)


def jet_dt(model, cfl: float = 0.5) -> float:
    """``cfl_dt`` of a shallow-water model."""
    return cfl_dt(model.strat.H, model.strat.g_prime, model.grid.dx, model.grid.dy, cfl)


def doublegyre_reparameterized_qg(nx: int = 128, ny: int = 128, Lx: float = 4e6, Ly: float = 4e6,
                                  f0: float = 9.375e-5, beta: float = 1.754e-11, n_layers: int = 3,
                                  H=(400.0, 1100.0, 2600.0), g_prime=(9.81, 0.025, 0.0125),
                                  lateral_viscosity: float = 15.0, bottom_drag: float = 3.6e-8,
                                  wind_amplitude: float = 8e-5, **kw):
    """``doublegyre_reparameterized_qg`` (gfd_testcases.py:365-430): the wind-driven
    double gyre in (u, v, h) with the geostrophic projection, from rest."""
    import torch

    from .reparam import ReparameterizedQG
    from .swm import MultilayerSW2DState

    model = ReparameterizedQG.create(nx=nx, ny=ny, Lx=Lx, Ly=Ly, f0=f0, beta=beta, n_layers=n_layers, H=H,
                                     g_prime=g_prime, lateral_viscosity=lateral_viscosity, bottom_drag=bottom_drag,
                                     wind_amplitude=wind_amplitude, wind_profile="doublegyre", bc="wall", **kw)
    nl, Ny, Nx = model.consts.n_layers, model.grid.Ny, model.grid.Nx
    dev = model.swm.device
    h0 = torch.ones((nl, Ny, Nx), device=dev) * torch.as_tensor(np.asarray(model.strat.H), dtype=torch.float32,
                                                               device=dev)[:, None, None]
    z = torch.zeros((nl, Ny, Nx), device=dev)
    return model, MultilayerSW2DState(h=h0, u=z, v=z.clone())
