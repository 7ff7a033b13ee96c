"""Multilayer and barotropic QG models on the HIP path.

Drop-in mirrors of ``somax/_src/models/qg/baroclinic.py`` and
``somax/_src/models/qg/barotropic.py``: same class names, state/param/const
field names, ``create()`` keyword arguments and defaults, and the
``vector_field`` / ``apply_boundary_conditions`` / ``integrate`` /
``diagnose`` contract of ``SomaxModel``.  Every evaluation runs in
``libsomaxhip.so`` (gfx950); there is no fallback.

Extensions beyond the reference (keyword-only, defaults reproduce the
reference): ``biharmonic_viscosity``, ``advection="weno"``,
``precision`` (``"fp32"``: FP32 MFMA transforms; ``"mixed"``: bf16 hi/lo
three-product transforms plus one refinement pass on the Helmholtz residual;
``"bf16x6"``: bf16 hi/mid/lo six-product transforms with FP64 accumulation,
one direct solve; ``"fp64"``: the whole step in FP64 on float64 state, as the
reference runs on the CPU with x64 enabled), ``mass_correction`` and
``helmholtz_eigenvalues``.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from .core import (
    ArakawaCGrid2D,
    Diagnostics,
    ModalTransform,
    Params,
    PhysConsts,
    Solution,
    SomaxModel,
    State,
    StratificationProfile,
    save_times,
    solver_name,
)

_ADV = {"arakawa": _lib.ADV_ARAKAWA, "weno": _lib.ADV_WENO}
_PREC = {"fp32": _lib.PREC_FP32, "mixed": _lib.PREC_MIXED, "bf16x6": _lib.PREC_BF16X6, "fp64": _lib.PREC_FP64}


def _torch():
    import torch

    return torch


# ---------------------------------------------------------------------------
# pytrees (baroclinic.py:23-88, barotropic.py)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class BaroclinicQGState(State):
    """Layer PV anomaly ``q`` on T points, ``(nl, Ny, Nx)`` (baroclinic.py:23-34)."""

    q: object


@dataclass(frozen=True)
class BaroclinicQGParams(Params):
    """baroclinic.py:37-48, plus the biharmonic coefficient (extension)."""

    lateral_viscosity: float
    bottom_drag: float
    wind_amplitude: float
    biharmonic_viscosity: float = 0.0


@dataclass(frozen=True)
class BaroclinicQGPhysConsts(PhysConsts):
    """baroclinic.py:51-62."""

    f0: float = 1e-4
    beta: float = 1.6e-11
    n_layers: int = 2


@dataclass(frozen=True)
class BaroclinicQGDiagnostics(Diagnostics):
    """baroclinic.py:65-88."""

    psi: object
    u: object
    v: object
    kinetic_energy: object
    total_kinetic_energy: object
    enstrophy: object
    total_enstrophy: object
    relative_vorticity: object
    rossby_radii: object


@dataclass(frozen=True)
class BarotropicQGState(State):
    """PV anomaly ``q`` on T points, ``(Ny, Nx)`` (barotropic.py)."""

    q: object


@dataclass(frozen=True)
class BarotropicQGParams(Params):
    lateral_viscosity: float
    bottom_drag: float
    wind_amplitude: float
    biharmonic_viscosity: float = 0.0


@dataclass(frozen=True)
class BarotropicQGPhysConsts(PhysConsts):
    f0: float = 1e-4
    beta: float = 1.6e-11


@dataclass(frozen=True)
class BarotropicQGDiagnostics(Diagnostics):
    psi: object
    u: object
    v: object
    kinetic_energy: object
    enstrophy: object
    relative_vorticity: object


def _row_profiles(grid: ArakawaCGrid2D, beta: float, Ly: float, wind_profile: str):
    """beta*(y - Ly/2) and the wind-curl pattern per T row, y_j = j*dy
    (baroclinic.py:307-318; the fields are functions of y only)."""
    y = np.arange(grid.Ny, dtype=np.float64) * grid.dy
    beta_y = beta * (y - Ly / 2.0)
    if wind_profile == "single":
        wind = np.sin(np.pi * y / Ly)
    else:
        wind = -np.sin(2.0 * np.pi * y / Ly)
    return beta_y, wind


def homogeneous_solution(Ny: int, Nx: int, dx: float, dy: float, lambdas) -> tuple[np.ndarray, np.ndarray]:
    """Per-mode homogeneous solution ``h_m = 1 + (D - lam_m)^-1 [lam_m]`` in FP64
    by dense DST-I on the host (setup only); returns ``(h, mean(h))``."""

    def sine(n):
        j = np.arange(1, n + 1, dtype=np.float64)
        return np.sqrt(2.0 / (n + 1)) * np.sin(np.pi * np.outer(j, j) / (n + 1))

    def eig(n, d):
        return (2.0 * np.cos(np.pi * np.arange(1, n + 1) / (n + 1)) - 2.0) / (d * d)

    Sy, Sx = sine(Ny), sine(Nx)
    D = eig(Ny, dy)[:, None] + eig(Nx, dx)[None, :]
    ones_hat = Sy @ np.ones((Ny, Nx)) @ Sx
    out = []
    for lam in np.asarray(lambdas, dtype=np.float64):
        out.append(1.0 + Sy @ (lam * ones_hat / (D - lam)) @ Sx)
    h = np.stack(out)
    return h, h.mean(axis=(-2, -1))


class _QGBase(SomaxModel):
    """Shared HIP plumbing of the two QG models."""

    def _setup(self, *, nl, grid, Cl2m, Cm2l, lambdas, beta_y_row, wind_row, wind_scale, zero_psi_ghost,
               advection, precision, mass_correction, device):
        if advection not in _ADV:
            raise ValueError(f"advection must be one of {sorted(_ADV)}, got {advection!r}")
        if precision not in _PREC:
            raise ValueError(f"precision must be one of {sorted(_PREC)}, got {precision!r}")
        self._nl = nl
        self._Cl2m = np.ascontiguousarray(Cl2m, dtype=np.float64)
        self._Cm2l = np.ascontiguousarray(Cm2l, dtype=np.float64)
        self._lambdas = np.ascontiguousarray(lambdas, dtype=np.float64)
        self._beta_row = np.ascontiguousarray(beta_y_row, dtype=np.float64)
        self._wind_row = np.ascontiguousarray(wind_row, dtype=np.float64)
        self._wind_scale = np.ascontiguousarray(wind_scale, dtype=np.float64)
        self._zero_psi_ghost = bool(zero_psi_ghost)
        self.advection = advection
        self.precision = precision
        self._f64 = precision == "fp64"
        self.mass_correction = bool(mass_correction)
        self.device = device
        self._plan = None
        self._ws = None
        self._tws = None
        self._homsol = None
        self._homsol_mean = None
        if self.mass_correction:
            self._homsol, self._homsol_mean = homogeneous_solution(grid.Ny, grid.Nx, grid.dx, grid.dy, lambdas)

    # -- plan and workspace -------------------------------------------------
    def _desc(self):
        """The C description of this model and the host arrays it points into
        (keep them alive while the description is in use)."""
        g = self.grid
        keep = []

        def dp(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return a.ctypes.data_as(_lib.c_double_p)

        d = _lib.QGDesc()
        d.nx, d.ny, d.nl = g.nx, g.ny, self._nl
        d.dx, d.dy = g.dx, g.dy
        d.cl2m, d.cm2l, d.lambdas = dp(self._Cl2m), dp(self._Cm2l), dp(self._lambdas)
        d.beta_y_row, d.wind_row, d.wind_scale = dp(self._beta_row), dp(self._wind_row), dp(self._wind_scale)
        d.zero_psi_ghost = int(self._zero_psi_ghost)
        d.advection = _ADV[self.advection]
        d.precision = _PREC[self.precision]
        d.mass_correction = int(self.mass_correction)
        if self.mass_correction:
            hs = np.ascontiguousarray(self._homsol, dtype=np.float32)
            keep.append(hs)
            d.homsol = hs.ctypes.data_as(_lib.c_float_p)
            d.homsol_mean = dp(self._homsol_mean)
        return d, keep

    def _ensure_plan(self):
        if self._plan is not None:
            return self._plan
        torch = _torch()
        lib = _lib.load()
        _lib.init_device()
        d, keep = self._desc()
        plan = ctypes.c_void_p()
        _lib.check(lib.somaxhip_qg_plan_create(ctypes.byref(d), ctypes.byref(plan)), "somaxhip_qg_plan_create")
        self._plan = plan
        nbytes = lib.somaxhip_qg_workspace_bytes(plan)
        self._ws = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=self.device)
        return plan

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None:
            try:
                _lib.load().somaxhip_qg_plan_destroy(plan)
            except Exception:
                pass

    def _c_params(self) -> _lib.QGParams:
        p = self.params
        return _lib.QGParams(
            float(p.lateral_viscosity), float(p.biharmonic_viscosity), float(p.bottom_drag), float(p.wind_amplitude)
        )

    def _field(self, q):
        torch = _torch()
        shape = (self._nl, self.grid.Ny, self.grid.Nx)
        dtype = torch.float64 if self._f64 else torch.float32
        if not isinstance(q, torch.Tensor) or q.device.type != "cuda":
            q = torch.as_tensor(np.asarray(q), dtype=dtype, device=self.device)
        if q.dtype != dtype:
            raise TypeError(f"state must be {str(dtype).replace('torch.', '')} for precision "
                            f"{self.precision!r}, got {q.dtype}")
        if q.numel() != math.prod(shape):
            raise ValueError(f"state has {tuple(q.shape)}, expected {shape}")
        return q.contiguous()

    @staticmethod
    def _stream():
        return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)

    def _fn(self, name):
        """The entry point ``name``, or its ``_f64`` form for an FP64 model."""
        name = name + "_f64" if self._f64 else name
        return name, getattr(_lib.load(), name)

    # -- kernels ------------------------------------------------------------
    def _invert(self, q, apply_bc: bool):
        torch = _torch()
        plan = self._ensure_plan()
        q = self._field(q)
        psi = torch.empty_like(q)
        name, fn = self._fn("somaxhip_qg_invert")
        _lib.check(fn(plan, q.data_ptr(), psi.data_ptr(), int(apply_bc), self._ws.data_ptr(), self._stream()), name)
        return psi

    def _rhs(self, q, apply_bc: bool, want_psi: bool = False):
        torch = _torch()
        plan = self._ensure_plan()
        q = self._field(q)
        dq = torch.empty_like(q)
        psi = torch.empty_like(q) if want_psi else None
        prm = self._c_params()
        name, fn = self._fn("somaxhip_qg_rhs")
        _lib.check(
            fn(plan, q.data_ptr(), dq.data_ptr(), psi.data_ptr() if want_psi else None, ctypes.byref(prm),
               int(apply_bc), self._ws.data_ptr(), self._stream()),
            name,
        )
        return (dq, psi) if want_psi else dq

    def _step_inplace(self, q, dt: float, nsteps: int):
        plan = self._ensure_plan()
        prm = self._c_params()
        name, fn = self._fn("somaxhip_qg_step")
        _lib.check(fn(plan, q.data_ptr(), None, int(nsteps), float(dt), ctypes.byref(prm), self._ws.data_ptr(),
                      self._stream()), name)

    def _step_tsit5_inplace(self, q, dt: float, nsteps: int):
        plan = self._ensure_plan()
        if self._tws is None:
            nbytes = _lib.load().somaxhip_qg_tsit5_workspace_bytes(plan)
            self._tws = _torch().empty(max(int(nbytes), 256), dtype=_torch().uint8, device=self.device)
        prm = self._c_params()
        name, fn = self._fn("somaxhip_qg_step_tsit5")
        _lib.check(fn(plan, q.data_ptr(), int(nsteps), float(dt), ctypes.byref(prm), self._ws.data_ptr(),
                      self._tws.data_ptr(), self._stream()), name)

    def _zero_ghost(self, q):
        q = self._field(q).clone()
        name, fn = self._fn("somaxhip_zero_ghost")
        _lib.check(fn(q.data_ptr(), self._nl, self.grid.ny, self.grid.nx, self._stream()), name)
        return q

    # -- time stepping (SomaxModel.integrate, model.py:53-88) ----------------
    def step(self, state, dt: float, nsteps: int = 1, solver="ssprk3"):
        """``nsteps`` steps of ``dt`` (BC applied before each RHS) with SSP-RK3
        (the benchmark scheme) or Tsit5; returns a new state."""
        q = self._field(state.q).clone()
        if solver_name(solver) == "tsit5":
            self._step_tsit5_inplace(q, dt, nsteps)
        else:
            self._step_inplace(q, dt, nsteps)
        return self._state_cls(q=q.reshape(state.q.shape))

    def integrate(self, state0, t0: float, t1: float, dt: float, **kw) -> Solution:
        """Constant-step integration from ``t0`` to ``t1``.

        Like the reference (model.py:53-88) the boundary condition is applied
        to ``state0`` first and before every RHS; the last step is shortened
        to land on ``t1``.  ``solver``: Tsit5 by default, as the reference
        (``"tsit5"`` or a ``Tsit5`` object), or ``"ssprk3"``.  ``saveat`` may
        be ``None`` (save ``t1`` only) or a sequence of save times.
        """
        torch = _torch()
        advance = self._step_tsit5_inplace if solver_name(kw.pop("solver", "tsit5")) == "tsit5" else self._step_inplace
        saveat = kw.pop("saveat", None)
        if kw:
            raise TypeError(f"unexpected keyword arguments {sorted(kw)}")
        ts = save_times(t0, t1, dt, saveat)
        q = self._zero_ghost(state0.q)
        t = float(t0)
        out = []
        for ts_k in ts:
            span = ts_k - t
            n = int(math.floor(span / dt + 1e-9))
            if n > 0:
                advance(q, dt, n)
                t += n * dt
            rem = ts_k - t
            if rem > 1e-12 * max(1.0, abs(ts_k)):
                advance(q, rem, 1)
                t = ts_k
            out.append(q.clone().reshape(state0.q.shape))
        ys = self._state_cls(q=torch.stack(out))
        return Solution(ts=torch.tensor(ts, dtype=torch.float64), ys=ys)

    def apply_boundary_conditions(self, state):
        """Ghost ring of q set to 0 (baroclinic.py:192-195)."""
        return self._state_cls(q=self._zero_ghost(state.q).reshape(state.q.shape))

    def vector_field(self, t: float, state, args=None):
        """``dq/dt`` on the given state as is (baroclinic.py:161-190)."""
        return self._state_cls(q=self._rhs(state.q, apply_bc=False).reshape(state.q.shape))

    def build_terms(self):
        def _rhs(t, state, args=None):
            return self._state_cls(q=self._rhs(state.q, apply_bc=True).reshape(state.q.shape))

        return _rhs

    def _diagnose(self, q):
        """psi, u, v, zeta fields and per-layer (KE, enstrophy) on the device."""
        torch = _torch()
        plan = self._ensure_plan()
        q = self._field(q)
        psi, u, v, zeta = (torch.empty_like(q) for _ in range(4))
        sums = torch.empty((self._nl, 2), dtype=torch.float64, device=q.device)
        name, fn = self._fn("somaxhip_qg_diagnose")
        _lib.check(
            fn(plan, q.data_ptr(), psi.data_ptr(), u.data_ptr(), v.data_ptr(), zeta.data_ptr(), sums.data_ptr(),
               self._ws.data_ptr(), self._stream()),
            name,
        )
        return psi, u, v, zeta, sums[:, 0], sums[:, 1]

    def _invert_pv(self, q):
        """``_invert_pv`` (baroclinic.py:135-159 / barotropic.py:113-121) on q as given."""
        return self._invert(q, apply_bc=False).reshape(q.shape)


class BaroclinicQG(_QGBase):
    """Multilayer QG model (baroclinic.py:91-332) on the HIP path."""

    _state_cls = BaroclinicQGState

    def __init__(self, params, consts, grid, modal, strat, beta_y_row, wind_row, helmholtz_lambdas,
                 poisson_bc="dst", *, advection="arakawa", precision="fp32", mass_correction=False,
                 device="cuda"):
        if poisson_bc != "dst":
            raise ValueError("only poisson_bc='dst' is implemented on the HIP path")
        self.params = params
        self.consts = consts
        self.grid = grid
        self.modal = modal
        self.strat = strat
        self.helmholtz_lambdas = np.asarray(helmholtz_lambdas, dtype=np.float64)
        self.poisson_bc = poisson_bc
        nl = strat.nl
        scale = np.zeros(nl)
        scale[0] = 1.0 / float(strat.H[0])
        self._setup(nl=nl, grid=grid, Cl2m=modal.Cl2m, Cm2l=modal.Cm2l, lambdas=self.helmholtz_lambdas,
                    beta_y_row=beta_y_row, wind_row=wind_row, wind_scale=scale, zero_psi_ghost=not mass_correction,
                    advection=advection, precision=precision, mass_correction=mass_correction, device=device)

    @property
    def beta_y(self):
        """``beta*(y - y0)`` broadcast to ``(Ny, Nx)`` (baroclinic.py:130), host array."""
        return np.broadcast_to(self._beta_row[:, None], (self.grid.Ny, self.grid.Nx))

    @property
    def wind_forcing(self):
        """Wind-curl pattern broadcast to ``(Ny, Nx)`` (baroclinic.py:131), host array."""
        return np.broadcast_to(self._wind_row[:, None], (self.grid.Ny, self.grid.Nx))

    @staticmethod
    def create(nx: int = 64, ny: int = 64, Lx: float = 4e6, Ly: float = 4e6, f0: float = 9.375e-5,
               beta: float = 1.754e-11, n_layers: int = 3, H=(400.0, 1100.0, 2600.0),
               g_prime=(9.81, 0.025, 0.0125), stratification: StratificationProfile | None = None,
               lateral_viscosity: float = 0.0, bottom_drag: float = 0.0, wind_amplitude: float = 0.0,
               wind_profile: str = "doublegyre", poisson_bc: str = "dst", *, biharmonic_viscosity: float = 0.0,
               advection: str = "arakawa", precision: str = "fp32", mass_correction: bool = False,
               helmholtz_eigenvalues: str = "consistent", device="cuda") -> "BaroclinicQG":
        """``BaroclinicQG.create`` (baroclinic.py:231-332), same arguments and errors.

        One default differs from the reference: the Helmholtz parameters are
        ``f0^2`` times the eigenvalues of the stretching matrix ``A`` that
        belong to ``Cm2l`` (``helmholtz_eigenvalues="consistent"``).  The
        reference takes ``jnp.linalg.eigh(A)`` of the non-symmetric ``A``
        (transforms.py:192-194), i.e. the eigenvalues of ``(A + A^T)/2``, whose
        barotropic one is negative and makes that mode's operator indefinite;
        ``helmholtz_eigenvalues="reference"`` reproduces them.  Keyword-only
        extensions: ``biharmonic_viscosity``, ``advection`` ("arakawa" or
        "weno"), ``precision`` ("fp32", "mixed", "bf16x6" or "fp64"),
        ``mass_correction``."""
        grid = ArakawaCGrid2D.from_interior(nx, ny, Lx, Ly)
        if stratification is not None:
            strat = stratification
        else:
            if len(H) != n_layers or len(g_prime) != n_layers:
                raise ValueError(
                    f"n_layers ({n_layers}), len(H) ({len(H)}), and "
                    f"len(g_prime) ({len(g_prime)}) must all be equal"
                )
            strat = StratificationProfile.from_layers(H=list(H), g_prime=list(g_prime))
        nl = strat.nl
        modal = ModalTransform.from_stratification(strat, f0)
        if helmholtz_eigenvalues == "reference":
            lambdas = f0**2 * modal.reference_eigenvalues
        elif helmholtz_eigenvalues == "consistent":
            lambdas = f0**2 * modal.eigenvalues
        else:
            raise ValueError("helmholtz_eigenvalues must be 'consistent' or 'reference'")
        params = BaroclinicQGParams(
            lateral_viscosity=float(lateral_viscosity), bottom_drag=float(bottom_drag),
            wind_amplitude=float(wind_amplitude), biharmonic_viscosity=float(biharmonic_viscosity),
        )
        consts = BaroclinicQGPhysConsts(f0=f0, beta=beta, n_layers=nl)
        beta_row, wind_row = _row_profiles(grid, beta, Ly, wind_profile)
        return BaroclinicQG(params, consts, grid, modal, strat, beta_row, wind_row, lambdas, poisson_bc,
                            advection=advection, precision=precision, mass_correction=mass_correction,
                            device=device)

    def diagnose(self, state) -> BaroclinicQGDiagnostics:
        """baroclinic.py:197-228: psi, geostrophic velocities on V / U faces,
        kinetic energy and enstrophy per layer (FP64 sums), relative vorticity."""
        torch = _torch()
        psi, u, v, zeta, ke, ens = self._diagnose(state.q)
        shape = (self._nl, self.grid.Ny, self.grid.Nx)
        return BaroclinicQGDiagnostics(
            psi=psi.reshape(shape), u=u.reshape(shape), v=v.reshape(shape), kinetic_energy=ke,
            total_kinetic_energy=ke.sum(), enstrophy=ens, total_enstrophy=ens.sum(),
            relative_vorticity=zeta.reshape(shape),
            rossby_radii=torch.as_tensor(np.asarray(self.modal.rossby_radii), dtype=torch.float64),
        )


class BarotropicQG(_QGBase):
    """Barotropic QG model (barotropic.py:21-248) on the HIP path: one layer,
    Poisson inversion whose ghost ring is kept (barotropic.py:113-121), wind
    not divided by a depth (barotropic.py:143), drag on the only layer."""

    _state_cls = BarotropicQGState

    def __init__(self, params, consts, grid, beta_y_row, wind_row, poisson_bc="dst", *, advection="arakawa",
                 precision="fp32", device="cuda"):
        if poisson_bc != "dst":
            raise ValueError("only poisson_bc='dst' is implemented on the HIP path")
        self.params = params
        self.consts = consts
        self.grid = grid
        self.poisson_bc = poisson_bc
        self._setup(nl=1, grid=grid, Cl2m=np.eye(1), Cm2l=np.eye(1), lambdas=np.zeros(1), beta_y_row=beta_y_row,
                    wind_row=wind_row, wind_scale=np.ones(1), zero_psi_ghost=False, advection=advection,
                    precision=precision, mass_correction=False, device=device)

    @property
    def beta_y(self):
        return np.broadcast_to(self._beta_row[:, None], (self.grid.Ny, self.grid.Nx))

    @property
    def wind_forcing(self):
        return np.broadcast_to(self._wind_row[:, None], (self.grid.Ny, self.grid.Nx))

    @staticmethod
    def create(nx: int = 64, ny: int = 64, Lx: float = 1e6, Ly: float = 1e6, f0: float = 1e-4,
               beta: float = 1.6e-11, lateral_viscosity: float = 0.0, bottom_drag: float = 0.0,
               wind_amplitude: float = 0.0, wind_profile: str = "doublegyre", *, biharmonic_viscosity: float = 0.0,
               advection: str = "arakawa", precision: str = "fp32", device="cuda") -> "BarotropicQG":
        """``BarotropicQG.create`` (barotropic.py:186-248)."""
        grid = ArakawaCGrid2D.from_interior(nx, ny, Lx, Ly)
        params = BarotropicQGParams(
            lateral_viscosity=float(lateral_viscosity), bottom_drag=float(bottom_drag),
            wind_amplitude=float(wind_amplitude), biharmonic_viscosity=float(biharmonic_viscosity),
        )
        consts = BarotropicQGPhysConsts(f0=f0, beta=beta)
        beta_row, wind_row = _row_profiles(grid, beta, Ly, wind_profile)
        return BarotropicQG(params, consts, grid, beta_row, wind_row, "dst", advection=advection,
                            precision=precision, device=device)

    def diagnose(self, state) -> BarotropicQGDiagnostics:
        """barotropic.py:159-183 (fields 2-D, energies scalar)."""
        psi, u, v, zeta, ke, ens = self._diagnose(state.q)
        shape = (self.grid.Ny, self.grid.Nx)
        return BarotropicQGDiagnostics(psi=psi.reshape(shape), u=u.reshape(shape), v=v.reshape(shape),
                                       kinetic_energy=ke[0], enstrophy=ens[0], relative_vorticity=zeta.reshape(shape))
