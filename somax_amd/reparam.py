"""Reparameterized QG on the HIP path: multilayer shallow water plus the
geostrophic projection P = G (Q G)^-1 Q.

Drop-in mirror of ``somax/_src/models/qg/reparameterized.py``: class names,
``create()`` arguments, defaults and errors, ``project`` / ``vector_field``
/ ``apply_boundary_conditions`` / ``diagnose``.  The projection runs in
``libsomaxhip.so`` (``somaxhip_qg_project``: a PV kernel, the QG Helmholtz
inversion, a geostrophic-reconstruction kernel), and so does the time
stepping (``somaxhip_rqg_step``: Tsit5 or SSP-RK3 on RHS(P(BC(y)))).
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from .core import Diagnostics, SomaxModel, Solution, save_times, solver_name
from .qg import BaroclinicQG
from .swm import MultilayerShallowWater2D, MultilayerSW2DState


def _torch():
    import torch

    return torch


@dataclass(frozen=True)
class ReparamQGDiagnostics(Diagnostics):
    """reparameterized.py:37-62."""

    energy: object
    total_energy: object
    enstrophy: object
    total_enstrophy: object
    potential_vorticity: object
    relative_vorticity: object
    kinetic_energy_field: object
    psi: object
    u_ageostrophic: object
    v_ageostrophic: object


class ReparameterizedQG(SomaxModel):
    """Multilayer shallow water with the geostrophic projection applied in
    ``apply_boundary_conditions`` (reparameterized.py:65-189)."""

    def __init__(self, swm: MultilayerShallowWater2D, qg: BaroclinicQG, helmholtz_lambdas, eigenvalues,
                 poisson_bc: str = "dst"):
        self.swm = swm
        self._qg = qg  # its plan holds the Helmholtz inversion of the projection
        self.helmholtz_lambdas = np.asarray(helmholtz_lambdas, dtype=np.float64)
        self.poisson_bc = poisson_bc
        m = swm.modal
        self._M = np.ascontiguousarray(m.Cm2l @ np.diag(np.asarray(eigenvalues, dtype=np.float64)) @ m.Cl2m)
        self._H = np.ascontiguousarray(swm.strat.H, dtype=np.float64)
        self._tws = None

    # -- delegates (reparameterized.py:95-123) ------------------------------
    @property
    def params(self):
        return self.swm.params

    @property
    def consts(self):
        return self.swm.consts

    @property
    def grid(self):
        return self.swm.grid

    @property
    def strat(self):
        return self.swm.strat

    @property
    def modal(self):
        return self.swm.modal

    # -- projection -----------------------------------------------------------
    def _project(self, h, u, v, h_deviation: bool = False, want_psi: bool = False):
        torch = _torch()
        qg = self._qg
        plan = qg._ensure_plan()
        ho, uo, vo = torch.empty_like(h), torch.empty_like(u), torch.empty_like(v)
        psi = torch.empty_like(h) if want_psi else None
        _lib.check(
            _lib.load().somaxhip_qg_project(
                plan, float(self.consts.f0), self._H.ctypes.data_as(_lib.c_double_p),
                self._M.ctypes.data_as(_lib.c_double_p), h.data_ptr(), u.data_ptr(), v.data_ptr(),
                int(h_deviation), ho.data_ptr(), uo.data_ptr(), vo.data_ptr(),
                psi.data_ptr() if want_psi else None, qg._ws.data_ptr(), self.swm._stream()),
            "somaxhip_qg_project",
        )
        return (ho, uo, vo, psi) if want_psi else (ho, uo, vo)

    def project(self, state) -> MultilayerSW2DState:
        """P = G (Q G)^-1 Q (reparameterized.py:142-178)."""
        h, u, v = self._project(*self.swm._fields(state))
        return MultilayerSW2DState(h=h, u=u, v=v)

    # -- SomaxModel -------------------------------------------------------------
    def vector_field(self, t: float, state, args=None):
        """The shallow-water tendencies (reparameterized.py:180-184)."""
        return self.swm.vector_field(t, state, args)

    def apply_boundary_conditions(self, state):
        """Shallow-water boundary condition, then the projection (:186-189)."""
        return self.project(self.swm.apply_boundary_conditions(state))

    def _step_inplace(self, h, u, v, dt: float, nsteps: int, solver: str = "tsit5"):
        torch = _torch()
        splan = self.swm._ensure_plan()
        qplan = self._qg._ensure_plan()
        if self._tws is None:
            nbytes = _lib.load().somaxhip_rqg_workspace_bytes(splan)
            self._tws = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=self.swm.device)
        code = _lib.SOLVER_TSIT5 if solver_name(solver) == "tsit5" else _lib.SOLVER_SSPRK3
        prm = self.swm._c_params()
        _lib.check(
            _lib.load().somaxhip_rqg_step(
                splan, qplan, code, float(self.consts.f0), self._H.ctypes.data_as(_lib.c_double_p),
                self._M.ctypes.data_as(_lib.c_double_p), h.data_ptr(), u.data_ptr(), v.data_ptr(), int(nsteps),
                float(dt), ctypes.byref(prm), self._qg._ws.data_ptr(), self._tws.data_ptr(), self.swm._stream()),
            "somaxhip_rqg_step",
        )

    def step(self, state, dt: float, nsteps: int = 1, solver="tsit5"):
        """``nsteps`` steps of the projected system; returns a new state."""
        h, u, v = (f.clone() for f in self.swm._fields(state))
        self._step_inplace(h, u, v, dt, nsteps, solver)
        return MultilayerSW2DState(h=h, u=u, v=v)

    def integrate(self, state0, t0: float, t1: float, dt: float, **kw) -> Solution:
        """As ``SomaxModel.integrate`` (model.py:53-88): the boundary condition
        (here including the projection) is applied to ``state0`` first;
        ``solver`` Tsit5 (default) or ``"ssprk3"``; ``saveat`` as for the QG
        models."""
        torch = _torch()
        solver = solver_name(kw.pop("solver", "tsit5"))
        saveat = kw.pop("saveat", None)
        if kw:
            raise TypeError(f"unexpected keyword arguments {sorted(kw)}")
        ts = save_times(t0, t1, dt, saveat)
        st = self.apply_boundary_conditions(state0)
        h, u, v = (f.clone() for f in self.swm._fields(st))
        t = float(t0)
        out = []
        for ts_k in ts:
            n = int(math.floor((ts_k - t) / dt + 1e-9))
            if n > 0:
                self._step_inplace(h, u, v, dt, n, solver)
                t += n * dt
            rem = ts_k - t
            if rem > 1e-12 * max(1.0, abs(ts_k)):
                self._step_inplace(h, u, v, rem, 1, solver)
                t = ts_k
            out.append((h.clone(), u.clone(), v.clone()))
        ys = MultilayerSW2DState(h=torch.stack([o[0] for o in out]), u=torch.stack([o[1] for o in out]),
                                 v=torch.stack([o[2] for o in out]))
        return Solution(ts=torch.tensor(ts, dtype=torch.float64), ys=ys)

    def diagnose(self, state) -> ReparamQGDiagnostics:
        """Shallow-water diagnostics plus psi and the ageostrophic velocity
        (reparameterized.py:191-222)."""
        d = self.swm.diagnose(state)
        h, u, v = self.swm._fields(state)
        _, ug, vg, psi = self._project(h, u, v, want_psi=True)
        return ReparamQGDiagnostics(energy=d.energy, total_energy=d.total_energy, enstrophy=d.enstrophy,
                                    total_enstrophy=d.total_enstrophy, potential_vorticity=d.potential_vorticity,
                                    relative_vorticity=d.relative_vorticity,
                                    kinetic_energy_field=d.kinetic_energy_field, psi=psi, u_ageostrophic=u - ug,
                                    v_ageostrophic=v - vg)

    @staticmethod
    def create(nx: int = 64, ny: int = 64, Lx: float = 4e6, Ly: float = 4e6, g: float = 9.81, f0: float = 9.375e-5,
               beta: float = 1.754e-11, n_layers: int = 3, H=(400.0, 1100.0, 2600.0),
               g_prime=(9.81, 0.025, 0.0125), stratification=None, lateral_viscosity: float = 0.0,
               bottom_drag: float = 0.0, wind_amplitude: float = 0.0, wind_profile: str = "doublegyre",
               bc: str = "wall", method: str = "upwind1", poisson_bc: str = "dst", *, precision: str = "fp32",
               helmholtz_eigenvalues: str = "consistent", device="cuda") -> "ReparameterizedQG":
        """``ReparameterizedQG.create`` (reparameterized.py:224-315): same
        arguments, defaults and ``ValueError``; ``precision`` and
        ``helmholtz_eigenvalues`` as for ``BaroclinicQG``."""
        if bc != "wall":
            raise ValueError(
                f"ReparameterizedQG requires wall BCs (got bc={bc!r}). "
                "The geostrophic projection uses Dirichlet Helmholtz "
                "inversion which is incompatible with periodic BCs."
            )
        if poisson_bc != "dst":
            raise NotImplementedError("the HIP path solves the projection's Helmholtz problem with DST-I only")
        if precision == "fp64":
            raise ValueError("precision 'fp64' is not available here: the projection runs on FP32 state "
                             "(precision must be 'fp32', 'mixed' or 'bf16x6')")
        swm = MultilayerShallowWater2D.create(
            nx=nx, ny=ny, Lx=Lx, Ly=Ly, g=g, f0=f0, beta=beta, n_layers=n_layers, H=H, g_prime=g_prime,
            stratification=stratification, lateral_viscosity=lateral_viscosity, bottom_drag=bottom_drag,
            wind_amplitude=wind_amplitude, wind_profile=wind_profile, bc=bc, method=method, device=device)
        strat = swm.strat
        qg = BaroclinicQG.create(nx=nx, ny=ny, Lx=Lx, Ly=Ly, f0=f0, beta=beta, n_layers=strat.nl,
                                 stratification=strat, precision=precision,
                                 helmholtz_eigenvalues=helmholtz_eigenvalues, device=device)
        eig = swm.modal.reference_eigenvalues if helmholtz_eigenvalues == "reference" else swm.modal.eigenvalues
        return ReparameterizedQG(swm, qg, f0 * f0 * np.asarray(eig), eig, poisson_bc=poisson_bc)
