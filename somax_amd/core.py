"""Model contract, state types and vertical modes (host side).

Mirrors ``somax/_src/core/types.py`` (State/Params/PhysConsts/Diagnostics),
``somax/_src/core/transforms.py`` (StratificationProfile, ModalTransform) and
``somax/_src/core/model.py`` (SomaxModel).  State leaves are ``torch`` tensors
on the GPU (float32, the reference's default dtype); torch only owns the
memory and the stream, every computation on them is a HIP kernel of
``libsomaxhip.so``.  Setup-time linear algebra (the vertical-mode
eigendecomposition) runs once in FP64 NumPy on the host, as in the reference
where it runs once in ``create()``.
"""

from __future__ import annotations

import abc
import dataclasses
from dataclasses import dataclass, field
from typing import Any

import numpy as np


# ---------------------------------------------------------------------------
# types (somax/_src/core/types.py:9-36)
# ---------------------------------------------------------------------------
class State:
    """Base class for model state vectors (types.py:9-15)."""


class Params:
    """Base class for differentiable model parameters (types.py:18-24)."""


class PhysConsts:
    """Base class for frozen physical constants (types.py:27-32)."""


class Diagnostics:
    """Base class for on-demand diagnostic quantities (types.py:35-40)."""


def replace(obj, **changes):
    """Functional update of a state/params dataclass (``eqx.tree_at`` analogue)."""
    return dataclasses.replace(obj, **changes)


# ---------------------------------------------------------------------------
# stratification and modes (somax/_src/core/transforms.py)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class StratificationProfile:
    """Layer thicknesses, reduced gravities and optional densities
    (transforms.py:11-148).  Arrays are FP64 NumPy (setup-time data)."""

    H: np.ndarray
    g_prime: np.ndarray
    rho: np.ndarray | None = None

    @property
    def nl(self) -> int:
        return int(np.asarray(self.H).shape[0])

    @property
    def total_depth(self) -> float:
        return float(np.sum(self.H))

    @staticmethod
    def from_N2_constant(N2: float, depth: float, n_layers: int, g: float = 9.81, rho0: float = 1025.0):
        """transforms.py:40-76."""
        H_val = depth / n_layers
        H = np.full(n_layers, H_val)
        g_prime = np.concatenate([[g], np.full(n_layers - 1, N2 * H_val)])
        drho = rho0 * N2 * H_val / g
        rho = rho0 + drho * np.arange(n_layers)
        return StratificationProfile(H=H, g_prime=g_prime, rho=rho)

    @staticmethod
    def from_N2_exponential(
        N2_surface: float, scale_depth: float, depth: float, n_layers: int, g: float = 9.81, rho0: float = 1025.0
    ):
        """transforms.py:78-115."""
        H_val = depth / n_layers
        H = np.full(n_layers, H_val)
        z_if = -np.arange(1, n_layers) * H_val
        g_prime = np.concatenate([[g], N2_surface * np.exp(z_if / scale_depth) * H_val])
        z_c = -(np.arange(n_layers) + 0.5) * H_val
        rho = rho0 + np.cumsum(rho0 * N2_surface * np.exp(z_c / scale_depth) * H_val / g)
        return StratificationProfile(H=H, g_prime=g_prime, rho=rho)

    @staticmethod
    def from_layers(H, g_prime, rho=None):
        """transforms.py:117-148, same ValueError messages."""
        if len(H) != len(g_prime):
            raise ValueError(f"H ({len(H)}) and g_prime ({len(g_prime)}) must have the same length")
        if rho is not None and len(rho) != len(H):
            raise ValueError(f"rho ({len(rho)}) must have the same length as H ({len(H)})")
        return StratificationProfile(
            H=np.asarray(H, dtype=np.float64),
            g_prime=np.asarray(g_prime, dtype=np.float64),
            rho=None if rho is None else np.asarray(rho, dtype=np.float64),
        )


def coupling_matrix(H, g_prime) -> np.ndarray:
    """Layer stretching matrix A of ``finitevolx.build_coupling_matrix``
    (called at transforms.py:192), MQGeometry form with the free-surface term
    ``1/(H0 g'0)`` on ``A[0, 0]``."""
    H = np.asarray(H, dtype=np.float64)
    g = np.asarray(g_prime, dtype=np.float64)
    nl = H.size
    A = np.zeros((nl, nl))
    if nl == 1:
        A[0, 0] = 1.0 / (H[0] * g[0])
        return A
    A[0, 0] = 1.0 / (H[0] * g[0]) + 1.0 / (H[0] * g[1])
    A[0, 1] = -1.0 / (H[0] * g[1])
    for k in range(1, nl - 1):
        A[k, k - 1] = -1.0 / (H[k] * g[k])
        A[k, k] = (1.0 / g[k] + 1.0 / g[k + 1]) / H[k]
        A[k, k + 1] = -1.0 / (H[k] * g[k + 1])
    A[-1, -2] = -1.0 / (H[-1] * g[-1])
    A[-1, -1] = 1.0 / (H[-1] * g[-1])
    return A


@dataclass(frozen=True)
class ModalTransform:
    """Layer <-> mode projections (transforms.py:151-224).

    ``eigenvalues`` are those of ``A`` that belong to the columns of ``Cm2l``
    (ascending).  ``reference_eigenvalues`` are what the reference stores at
    transforms.py:194 (``jnp.linalg.eigh`` of a non-symmetric ``A``, which
    JAX symmetrises first); for the stratifications used here they contain a
    negative value, which makes the reference's per-mode Helmholtz operator
    resonant, so models use ``eigenvalues`` unless asked otherwise.
    """

    Cl2m: np.ndarray
    Cm2l: np.ndarray
    eigenvalues: np.ndarray
    rossby_radii: np.ndarray
    reference_eigenvalues: np.ndarray = field(default=None)

    @staticmethod
    def from_physics(H, g_prime, f0: float) -> "ModalTransform":
        A = coupling_matrix(H, g_prime)
        w, R = np.linalg.eig(A)
        w = w.real
        R = R.real
        order = np.argsort(w)
        w = w[order]
        R = R[:, order]
        Cl2m = np.linalg.inv(R)
        with np.errstate(divide="ignore", invalid="ignore"):
            radii = 1.0 / (f0 * np.sqrt(w))
        ref = np.linalg.eigvalsh(0.5 * (A + A.T))
        return ModalTransform(Cl2m=Cl2m, Cm2l=R, eigenvalues=w, rossby_radii=radii, reference_eigenvalues=ref)

    @staticmethod
    def from_stratification(strat: StratificationProfile, f0: float) -> "ModalTransform":
        return ModalTransform.from_physics(strat.H, strat.g_prime, f0)

    def to_modal(self, x):
        """``einsum("lm,m...->l...", Cl2m, x)`` (transforms.py:218-220), host arrays."""
        return np.einsum("lm,m...->l...", self.Cl2m, np.asarray(x, dtype=np.float64))

    def to_layer(self, x):
        """``einsum("lm,m...->l...", Cm2l, x)`` (transforms.py:222-224), host arrays."""
        return np.einsum("lm,m...->l...", self.Cm2l, np.asarray(x, dtype=np.float64))


# ---------------------------------------------------------------------------
# grid (finitevolx.ArakawaCGrid2D.from_interior, baroclinic.py:278)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class ArakawaCGrid2D:
    nx: int
    ny: int
    Lx: float
    Ly: float

    @staticmethod
    def from_interior(nx: int, ny: int, Lx: float, Ly: float) -> "ArakawaCGrid2D":
        return ArakawaCGrid2D(nx=nx, ny=ny, Lx=Lx, Ly=Ly)

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


# ---------------------------------------------------------------------------
# model contract (somax/_src/core/model.py:12-95)
# ---------------------------------------------------------------------------
@dataclass
class Solution:
    """What ``integrate`` returns: the ``ts``/``ys`` fields of a diffrax
    Solution that the reference's callers read (``sol.ys.q``)."""

    ts: Any
    ys: Any
    stats: dict = field(default_factory=dict)


def save_times(t0: float, t1: float, dt: float, saveat) -> list[float]:
    """The save times of ``integrate`` (``t1`` alone when ``saveat`` is None),
    checked as diffrax checks them: a positive step, and save times that do
    not decrease and lie at or after ``t0``."""
    ts = [float(t1)] if saveat is None else [float(t) for t in saveat]
    if not (float(dt) > 0.0):
        raise ValueError(f"dt must be positive, got {dt!r}")
    if not ts:
        raise ValueError("saveat holds no save times")
    if any(t < float(t0) for t in ts):
        raise ValueError(f"save times must lie at or after t0 = {t0!r}, got {ts}")
    if any(b < a for a, b in zip(ts, ts[1:])):
        raise ValueError(f"save times must not decrease, got {ts}")
    return ts


def solver_name(solver) -> str:
    """``"tsit5"`` or ``"ssprk3"`` from a solver name or a solver object
    (``diffrax.Tsit5()`` -> ``"tsit5"``): the schemes the HIP path runs."""
    name = solver if isinstance(solver, str) else type(solver).__name__
    key = name.lower().replace("-", "").replace("_", "")
    if key == "tsit5":
        return "tsit5"
    if key in ("ssprk3", "ssprk33"):
        return "ssprk3"
    raise NotImplementedError(f"solver {name!r}: the HIP path implements Tsit5 and SSP-RK3")


class SomaxModel(abc.ABC):
    """Abstract model contract (model.py:12-95)."""

    @abc.abstractmethod
    def vector_field(self, t: float, state, args=None):
        """Right-hand side of the PDE (model.py:24-32)."""

    @abc.abstractmethod
    def apply_boundary_conditions(self, state):
        """Boundary enforcement (model.py:35-37)."""

    def build_terms(self):
        """``_rhs(t, state, args) = vector_field(t, apply_boundary_conditions(state), args)``
        (model.py:39-51), returned as a plain callable."""

        def _rhs(t, state, args=None):
            return self.vector_field(t, self.apply_boundary_conditions(state), args)

        return _rhs

    def diagnose(self, state):
        return {}
