"""Multilayer shallow-water model on the HIP path.

Drop-in mirror of ``somax/_src/models/swm/multilayer.py``: class names,
state/param/const fields, ``create()`` arguments and defaults, and the
``SomaxModel`` contract.  Every evaluation runs in ``libsomaxhip.so``
(gfx950); there is no fallback.  Extensions (keyword-compatible):
``bc="channel"`` -- periodic in x, free-slip walls in y;
``vorticity_flux="energy_enstrophy"`` -- the Arakawa & Lamb (1981) triad form
of the PV flux (no work done, potential enstrophy of non-divergent flow
kept) instead of the reference's average form (the default, "average").
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

_BC = {"periodic": _lib.BC_PERIODIC, "wall": _lib.BC_WALL, "channel": _lib.BC_CHANNEL}
_VFLUX = {"average": 0, "energy_enstrophy": 1}


def _torch():
    import torch

    return torch


@dataclass(frozen=True)
class MultilayerSW2DState(State):
    """``h``, ``u``, ``v``, each ``(nl, Ny, Nx)`` (multilayer.py:29-40)."""

    h: object
    u: object
    v: object


@dataclass(frozen=True)
class MultilayerSW2DParams(Params):
    """multilayer.py:43-55."""

    lateral_viscosity: float
    bottom_drag: float
    wind_amplitude: float


@dataclass(frozen=True)
class MultilayerSW2DPhysConsts(PhysConsts):
    """multilayer.py:58-70."""

    gravity: float = 9.81
    f0: float = 1e-4
    beta: float = 0.0
    n_layers: int = 2


@dataclass(frozen=True)
class MultilayerSW2DDiagnostics(Diagnostics):
    """multilayer.py:73-92."""

    energy: object
    total_energy: object
    enstrophy: object
    total_enstrophy: object
    potential_vorticity: object
    relative_vorticity: object
    kinetic_energy_field: object


class MultilayerShallowWater2D(SomaxModel):
    """Multilayer vector-invariant shallow-water model (multilayer.py:95-377)."""

    def __init__(self, params, consts, grid, strat, modal, f_row, wind_row, bc_type="periodic", method="upwind1",
                 device="cuda", vorticity_flux="average"):
        if bc_type not in _BC:
            raise ValueError(f"bc must be one of {sorted(_BC)}, got {bc_type!r}")
        if vorticity_flux not in _VFLUX:
            raise ValueError(f"vorticity_flux must be one of {sorted(_VFLUX)}, got {vorticity_flux!r}")
        self.vorticity_flux = vorticity_flux
        if method != "upwind1":
            raise ValueError("only method='upwind1' is implemented on the HIP path")
        self.params = params
        self.consts = consts
        self.grid = grid
        self.strat = strat
        self.modal = modal
        self._f_row = np.ascontiguousarray(f_row, dtype=np.float64)
        self._wind_row = np.ascontiguousarray(wind_row, dtype=np.float64)
        self.bc_type = bc_type
        self.method = method
        self.device = device
        self._nl = strat.nl
        self._plan = None
        self._ws = None
        self._tws = None

    # -- reference fields (host arrays) ---------------------------------------
    @property
    def f_field(self):
        """Coriolis parameter at T points, ``(Ny, Nx)`` (multilayer.py:345-349)."""
        return np.broadcast_to(self._f_row[:, None], (self.grid.Ny, self.grid.Nx))

    @property
    def wind_stress_x(self):
        return np.broadcast_to(self._wind_row[:, None], (self.grid.Ny, self.grid.Nx))

    @property
    def wind_stress_y(self):
        return np.zeros((self.grid.Ny, self.grid.Nx))

    # -- plan -----------------------------------------------------------------
    def _ensure_plan(self):
        if self._plan is not None:
            return self._plan
        torch = _torch()
        lib = _lib.load()
        _lib.init_device()
        g = self.grid
        keep = [np.ascontiguousarray(self.strat.g_prime, dtype=np.float64),
                np.ascontiguousarray(self.strat.H, dtype=np.float64), self._f_row, self._wind_row]
        d = _lib.SWMDesc()
        d.nx, d.ny, d.nl = g.nx, g.ny, self._nl
        d.dx, d.dy = g.dx, g.dy
        d.g_prime, d.H, d.f_row, d.wind_row = (a.ctypes.data_as(_lib.c_double_p) for a in keep)
        d.bc = _BC[self.bc_type]
        d.vorticity_flux = _VFLUX[self.vorticity_flux]
        plan = ctypes.c_void_p()
        _lib.check(lib.somaxhip_swm_plan_create(ctypes.byref(d), ctypes.byref(plan)), "somaxhip_swm_plan_create")
        self._plan = plan
        self._ws = torch.empty(max(int(lib.somaxhip_swm_workspace_bytes(plan)), 256), dtype=torch.uint8,
                               device=self.device)
        return plan

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None:
            try:
                _lib.load().somaxhip_swm_plan_destroy(plan)
            except Exception:
                pass

    def _c_params(self) -> _lib.SWMParams:
        p = self.params
        return _lib.SWMParams(float(p.lateral_viscosity), float(p.bottom_drag), float(p.wind_amplitude))

    def _field(self, a):
        torch = _torch()
        shape = (self._nl, self.grid.Ny, self.grid.Nx)
        if not isinstance(a, torch.Tensor) or a.device.type != "cuda":
            a = torch.as_tensor(np.asarray(a), dtype=torch.float32, device=self.device)
        if a.dtype != torch.float32:
            raise TypeError(f"state fields must be float32, got {a.dtype}")
        if a.numel() != math.prod(shape):
            raise ValueError(f"state field has {tuple(a.shape)}, expected {shape}")
        return a.reshape(shape).contiguous()

    @staticmethod
    def _stream():
        return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)

    def _fields(self, state):
        return self._field(state.h), self._field(state.u), self._field(state.v)

    def _rhs(self, state, apply_bc: bool):
        torch = _torch()
        plan = self._ensure_plan()
        h, u, v = self._fields(state)
        dh, du, dv = torch.empty_like(h), torch.empty_like(u), torch.empty_like(v)
        prm = self._c_params()
        _lib.check(
            _lib.load().somaxhip_swm_rhs(plan, h.data_ptr(), u.data_ptr(), v.data_ptr(), dh.data_ptr(),
                                        du.data_ptr(), dv.data_ptr(), ctypes.byref(prm), int(apply_bc),
                                        self._stream()),
            "somaxhip_swm_rhs",
        )
        return MultilayerSW2DState(h=dh, u=du, v=dv)

    def _step_inplace(self, h, u, v, dt: float, nsteps: int):
        plan = self._ensure_plan()
        prm = self._c_params()
        _lib.check(
            _lib.load().somaxhip_swm_step(plan, h.data_ptr(), u.data_ptr(), v.data_ptr(), int(nsteps), float(dt),
                                         ctypes.byref(prm), self._ws.data_ptr(), self._stream()),
            "somaxhip_swm_step",
        )

    def _step_tsit5_inplace(self, h, u, v, dt: float, nsteps: int):
        plan = self._ensure_plan()
        if self._tws is None:
            nbytes = _lib.load().somaxhip_swm_tsit5_workspace_bytes(plan)
            self._tws = _torch().empty(max(int(nbytes), 256), dtype=_torch().uint8, device=self.device)
        prm = self._c_params()
        _lib.check(
            _lib.load().somaxhip_swm_step_tsit5(plan, h.data_ptr(), u.data_ptr(), v.data_ptr(), int(nsteps),
                                               float(dt), ctypes.byref(prm), self._tws.data_ptr(), self._stream()),
            "somaxhip_swm_step_tsit5",
        )

    # -- SomaxModel -----------------------------------------------------------
    def vector_field(self, t: float, state, args=None):
        """Tendencies on the state as given (multilayer.py:150-201)."""
        return self._rhs(state, apply_bc=False)

    def build_terms(self):
        def _rhs(t, state, args=None):
            return self._rhs(state, apply_bc=True)

        return _rhs

    def apply_boundary_conditions(self, state):
        """multilayer.py:203-223 (and the channel extension)."""
        torch = _torch()
        plan = self._ensure_plan()
        h, u, v = self._fields(state)
        ho, uo, vo = torch.empty_like(h), torch.empty_like(u), torch.empty_like(v)
        _lib.check(
            _lib.load().somaxhip_swm_apply_bc(plan, h.data_ptr(), u.data_ptr(), v.data_ptr(), ho.data_ptr(),
                                             uo.data_ptr(), vo.data_ptr(), self._stream()),
            "somaxhip_swm_apply_bc",
        )
        return MultilayerSW2DState(h=ho, u=uo, v=vo)

    def step(self, state, dt: float, nsteps: int = 1, solver="ssprk3"):
        """``nsteps`` steps (BC before every RHS) with SSP-RK3 (the benchmark
        scheme) or Tsit5; returns a new state."""
        h, u, v = (f.clone() for f in self._fields(state))
        if solver_name(solver) == "tsit5":
            self._step_tsit5_inplace(h, u, v, dt, nsteps)
        else:
            self._step_inplace(h, u, v, dt, nsteps)
        return MultilayerSW2DState(h=h, u=u, v=v)

    def integrate(self, state0, t0: float, t1: float, dt: float, **kw) -> Solution:
        """Constant-step integration from ``t0`` to ``t1`` (the last step
        shortened to land on ``t1``); ``solver`` (Tsit5 by default, as the
        reference, or ``"ssprk3"``) and ``saveat`` as for the QG models."""
        torch = _torch()
        advance = self._step_tsit5_inplace if solver_name(kw.pop("solver", "tsit5")) == "tsit5" else self._step_inplace
        saveat = kw.pop("saveat", None)
        if kw:
            raise TypeError(f"unexpected keyword arguments {sorted(kw)}")
        ts = save_times(t0, t1, dt, saveat)
        h, u, v = self._fields(self.apply_boundary_conditions(state0))  # model.py:71
        t = float(t0)
        out = []
        for ts_k in ts:
            n = int(math.floor((ts_k - t) / dt + 1e-9))
            if n > 0:
                advance(h, u, v, dt, n)
                t += n * dt
            rem = ts_k - t
            if rem > 1e-12 * max(1.0, abs(ts_k)):
                advance(h, u, v, rem, 1)
                t = ts_k
            out.append((h.clone(), u.clone(), v.clone()))
        ys = MultilayerSW2DState(h=torch.stack([o[0] for o in out]), u=torch.stack([o[1] for o in out]),
                                 v=torch.stack([o[2] for o in out]))
        return Solution(ts=torch.tensor(ts, dtype=torch.float64), ys=ys)

    def diagnose(self, state) -> MultilayerSW2DDiagnostics:
        """multilayer.py:225-255: energy and potential enstrophy per layer
        (FP64 sums), PV and relative vorticity at X points, KE at T points."""
        torch = _torch()
        plan = self._ensure_plan()
        h, u, v = self._fields(state)
        q, zeta, ke = (torch.empty_like(h) for _ in range(3))
        sums = torch.empty((self._nl, 2), dtype=torch.float64, device=h.device)
        _lib.check(
            _lib.load().somaxhip_swm_diagnose(plan, h.data_ptr(), u.data_ptr(), v.data_ptr(), q.data_ptr(),
                                             zeta.data_ptr(), ke.data_ptr(), sums.data_ptr(), self._ws.data_ptr(),
                                             self._stream()),
            "somaxhip_swm_diagnose",
        )
        energy, enstrophy = sums[:, 0], sums[:, 1]
        return MultilayerSW2DDiagnostics(energy=energy, total_energy=energy.sum(), enstrophy=enstrophy,
                                         total_enstrophy=enstrophy.sum(), potential_vorticity=q,
                                         relative_vorticity=zeta, kinetic_energy_field=ke)

    @staticmethod
    def create(nx: int = 64, ny: int = 64, Lx: float = 4e6, Ly: float = 4e6, g: float = 9.81, f0: float = 9.375e-5,
               beta: float = 1.754e-11, n_layers: int = 3, H=(400.0, 1100.0, 2600.0),
               g_prime=(9.81, 0.025, 0.0125), stratification: StratificationProfile | None = None,
               lateral_viscosity: float = 0.0, bottom_drag: float = 0.0, wind_amplitude: float = 0.0,
               wind_profile: str = "doublegyre", bc: str = "periodic", method: str = "upwind1",
               *, vorticity_flux: str = "average", device="cuda") -> "MultilayerShallowWater2D":
        """``MultilayerShallowWater2D.create`` (multilayer.py:259-377), same arguments and errors;
        ``vorticity_flux`` (extension): "average" (the reference's) or "energy_enstrophy"."""
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
        modal = ModalTransform.from_stratification(strat, f0)
        params = MultilayerSW2DParams(lateral_viscosity=float(lateral_viscosity), bottom_drag=float(bottom_drag),
                                      wind_amplitude=float(wind_amplitude))
        consts = MultilayerSW2DPhysConsts(gravity=g, f0=f0, beta=beta, n_layers=strat.nl)
        y = np.arange(grid.Ny, dtype=np.float64) * grid.dy
        f_row = f0 + beta * (y - Ly / 2.0)
        if wind_profile == "single":
            wind = -np.cos(np.pi * y / Ly)
        else:
            wind = -np.cos(2.0 * np.pi * y / Ly)
        return MultilayerShallowWater2D(params, consts, grid, strat, modal, f_row, wind, bc_type=bc, method=method,
                                        device=device, vorticity_flux=vorticity_flux)


# ---------------------------------------------------------------------------
# single-layer NonlinearShallowWater2D (swm/nonlinear_2d.py)
# ---------------------------------------------------------------------------
@dataclass
class NonlinearSW2DState(State):
    """``h``, ``u``, ``v``, each ``(Ny, Nx)`` (nonlinear_2d.py:25-36)."""

    h: object
    u: object
    v: object


@dataclass
class NonlinearSW2DParams(Params):
    """nonlinear_2d.py:39-50."""

    lateral_viscosity: float = 0.0
    bottom_drag: float = 0.0
    wind_amplitude: float = 0.0


@dataclass(frozen=True)
class NonlinearSW2DPhysConsts(PhysConsts):
    """nonlinear_2d.py:53-66."""

    gravity: float = 9.81
    f0: float = 1e-4
    beta: float = 0.0
    H0: float = 100.0


@dataclass(frozen=True)
class NonlinearSW2DDiagnostics(Diagnostics):
    """nonlinear_2d.py:69-85."""

    energy: object
    enstrophy: object
    potential_vorticity: object
    relative_vorticity: object
    kinetic_energy_field: object


class NonlinearShallowWater2D(SomaxModel):
    """2-D nonlinear (vector-invariant) shallow water, one layer
    (nonlinear_2d.py:88-318).  Its equations are the multilayer model's at one
    layer with g' = g, except that the wind stress is not divided by the depth
    (nonlinear_2d.py:180-182); it runs on a one-layer multilayer plan whose
    wind amplitude is scaled by H0."""

    def __init__(self, params: NonlinearSW2DParams, consts: NonlinearSW2DPhysConsts, ml: MultilayerShallowWater2D):
        self.params = params
        self.consts = consts
        self._ml = ml

    @property
    def grid(self):
        return self._ml.grid

    @property
    def bc_type(self):
        return self._ml.bc_type

    @property
    def method(self):
        return self._ml.method

    @property
    def f_field(self):
        return self._ml.f_field

    @property
    def wind_stress_x(self):
        return self._ml.wind_stress_x

    @property
    def wind_stress_y(self):
        return self._ml.wind_stress_y

    def _sync(self) -> MultilayerShallowWater2D:
        p = self.params
        self._ml.params = MultilayerSW2DParams(lateral_viscosity=float(p.lateral_viscosity),
                                               bottom_drag=float(p.bottom_drag),
                                               wind_amplitude=float(p.wind_amplitude) * self.consts.H0)
        return self._ml

    def _to2d(self, st) -> NonlinearSW2DState:
        shape = (self.grid.Ny, self.grid.Nx)
        return NonlinearSW2DState(h=st.h.reshape(shape), u=st.u.reshape(shape), v=st.v.reshape(shape))

    def vector_field(self, t: float, state, args=None) -> NonlinearSW2DState:
        """nonlinear_2d.py:132-194 on the state as given."""
        return self._to2d(self._sync().vector_field(t, state, args))

    def apply_boundary_conditions(self, state) -> NonlinearSW2DState:
        """Periodic or free-slip wall (nonlinear_2d.py:196-217)."""
        return self._to2d(self._sync().apply_boundary_conditions(state))

    def build_terms(self):
        def _rhs(t, state, args=None):
            return self._to2d(self._sync()._rhs(state, apply_bc=True))

        return _rhs

    def step(self, state, dt: float, nsteps: int = 1, solver="ssprk3") -> NonlinearSW2DState:
        return self._to2d(self._sync().step(state, dt, nsteps, solver=solver))

    def integrate(self, state0, t0: float, t1: float, dt: float, **kw) -> Solution:
        """As the multilayer model's (Tsit5 by default, BC applied to state0)."""
        sol = self._sync().integrate(state0, t0, t1, dt, **kw)
        shape = (-1, self.grid.Ny, self.grid.Nx)
        ys = NonlinearSW2DState(h=sol.ys.h.reshape(shape), u=sol.ys.u.reshape(shape), v=sol.ys.v.reshape(shape))
        return Solution(ts=sol.ts, ys=ys)

    def diagnose(self, state) -> NonlinearSW2DDiagnostics:
        """Scalar energy and potential enstrophy, PV, vorticity, KE fields
        (nonlinear_2d.py:219-241)."""
        d = self._sync().diagnose(state)
        shape = (self.grid.Ny, self.grid.Nx)
        return NonlinearSW2DDiagnostics(energy=d.energy[0], enstrophy=d.enstrophy[0],
                                        potential_vorticity=d.potential_vorticity.reshape(shape),
                                        relative_vorticity=d.relative_vorticity.reshape(shape),
                                        kinetic_energy_field=d.kinetic_energy_field.reshape(shape))

    @staticmethod
    def create(nx: int = 64, ny: int = 64, Lx: float = 1e6, Ly: float = 1e6, g: float = 9.81, f0: float = 1e-4,
               beta: float = 0.0, H0: float = 100.0, lateral_viscosity: float = 0.0, bottom_drag: float = 0.0,
               wind_amplitude: float = 0.0, wind_profile: str = "doublegyre", bc: str = "periodic",
               method: str = "upwind1", *, device="cuda") -> "NonlinearShallowWater2D":
        """``NonlinearShallowWater2D.create`` (nonlinear_2d.py:243-318), same arguments."""
        if bc not in ("periodic", "wall"):
            raise ValueError(f"bc must be 'periodic' or 'wall', got {bc!r}")
        ml = MultilayerShallowWater2D.create(nx=nx, ny=ny, Lx=Lx, Ly=Ly, g=g, f0=f0, beta=beta, n_layers=1,
                                             H=(H0,), g_prime=(g,), wind_profile=wind_profile, bc=bc, method=method,
                                             device=device)
        params = NonlinearSW2DParams(lateral_viscosity=float(lateral_viscosity), bottom_drag=float(bottom_drag),
                                     wind_amplitude=float(wind_amplitude))
        return NonlinearShallowWater2D(params, NonlinearSW2DPhysConsts(gravity=g, f0=f0, beta=beta, H0=H0), ml)
