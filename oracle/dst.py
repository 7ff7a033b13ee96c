"""DST-I Dirichlet Helmholtz solve over the full ghosted array (FP64).

Restates ``finitevolx.streamfunction_from_vorticity(q, dx, dy, bc="dst",
lambda_=lam)`` as called at ``somax/_src/models/qg/barotropic.py:119-121`` and
``finitevolx.pv_inversion(q_modal, dx, dy, lambda_=..., bc="dst")`` at
``somax/_src/models/qg/baroclinic.py:146-152``.  finitevolX v0.0.39 is an
un-vendored git dependency (pyproject.toml:32) and cannot be imported here, so
the convention is pinned from published outputs instead:

    phi = S_Ny [ (S_Ny f S_Nx) / (Lam_Ny[:, None] + Lam_Nx[None, :] - lam) ] S_Nx

with ``S_N[j, k] = sqrt(2/(N+1)) sin(pi (j+1)(k+1)/(N+1))`` (orthonormal,
symmetric, S @ S = I) and ``Lam_N[k] = (2 cos(pi (k+1)/(N+1)) - 2) / d**2``,
applied over the whole ``(Ny, Nx)`` array *including* the ghost ring.  This
reproduces the ten numbers printed by the executed tutorials
``content/tutorials/step09_laplace_2d.ipynb`` (cells 7, 11, 13),
``step10_poisson_2d.ipynb`` (cell 4) and ``step11_helmholtz_2d.ipynb`` (cell 6);
see ``tests/test_oracle_pins.py``.  It is the exact inverse of the 5-point
operator with homogeneous Dirichlet values at the virtual indices -1 and N.
"""

from __future__ import annotations

import numpy as np
import scipy.fft


def dst1_matrix(n: int) -> np.ndarray:
    """Orthonormal DST-I matrix ``S[j, k] = sqrt(2/(n+1)) sin(pi(j+1)(k+1)/(n+1))``."""
    j = np.arange(1, n + 1, dtype=np.float64)
    return np.sqrt(2.0 / (n + 1)) * np.sin(np.pi * np.outer(j, j) / (n + 1))


def dst1_eigenvalues(n: int, d: float) -> np.ndarray:
    """Eigenvalues of the 1-D 3-point second difference under DST-I, ``(2cos(pi k/(n+1)) - 2)/d^2``."""
    k = np.arange(1, n + 1, dtype=np.float64)
    return (2.0 * np.cos(np.pi * k / (n + 1)) - 2.0) / (d * d)


def dst2(f: np.ndarray, workers: int | None = None) -> np.ndarray:
    """Orthonormal 2-D DST-I over the last two axes (textbook ``scipy.fft.dstn``).

    ``workers`` threads split the independent 1-D transforms, which leaves
    every result bit unchanged."""
    return scipy.fft.dstn(f, type=1, axes=(-2, -1), norm="ortho", workers=workers)


def helmholtz_dst(f: np.ndarray, dx: float, dy: float, lam: float = 0.0, workers: int | None = None) -> np.ndarray:
    """Solve ``(nabla_h^2 - lam) phi = f`` by DST-I over the full array (ghosts included).

    ``f`` has shape ``(..., Ny, Nx)``; ``lam`` may be a scalar or broadcast
    against the leading axes (one value per mode, as ``pv_inversion`` takes a
    vector ``lambda_``).
    """
    f = np.asarray(f, dtype=np.float64)
    ny, nx = f.shape[-2:]
    lam = np.asarray(lam, dtype=np.float64)
    denom = (dst1_eigenvalues(ny, dy)[:, None] + dst1_eigenvalues(nx, dx)[None, :])
    denom = denom - lam[..., None, None] if lam.ndim else denom - lam
    return dst2(dst2(f, workers) / denom, workers)


def helmholtz_apply(phi: np.ndarray, dx: float, dy: float, lam: float = 0.0) -> np.ndarray:
    """Apply the operator that ``helmholtz_dst`` inverts: 5-point Laplacian with
    zero values at virtual indices -1 and N on every side, minus ``lam``.
    Used for the residual of iterative refinement and for brute-force pins."""
    phi = np.asarray(phi, dtype=np.float64)
    p = np.pad(phi, [(0, 0)] * (phi.ndim - 2) + [(1, 1), (1, 1)])
    lap = (p[..., 1:-1, 2:] - 2.0 * phi + p[..., 1:-1, :-2]) / (dx * dx) + (
        p[..., 2:, 1:-1] - 2.0 * phi + p[..., :-2, 1:-1]
    ) / (dy * dy)
    lam = np.asarray(lam, dtype=np.float64)
    if lam.ndim:
        lam = lam[..., None, None]
    return lap - lam * phi
