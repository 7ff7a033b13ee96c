"""Vertical stretching matrix and layer/mode transforms (FP64).

Restates ``ModalTransform.from_physics`` (``somax/_src/core/transforms.py:172-201``)
and the two finitevolX routines it delegates to, ``build_coupling_matrix(H, g')``
and ``decompose_vertical_modes(A, f0)``.  finitevolX is not importable, so the
MQGeometry formulation that the reference docstrings cite
(``somax/_src/models/qg/baroclinic.py:107``) is used:

* ``A`` is tridiagonal with ``A[0,0] = 1/(H0 g'0) + 1/(H0 g'1)`` (the ``g'0``
  free-surface term can be switched off to get the rigid-lid matrix of the old
  dev notebooks), ``A[k,k-1] = -1/(Hk g'k)``, ``A[k,k] = (1/g'k + 1/g'(k+1))/Hk``,
  ``A[k,k+1] = -1/(Hk g'(k+1))``, ``A[-1,-1] = 1/(H(-1) g'(-1))``.
* Modes are the eigen-pairs of ``A`` sorted by ascending eigenvalue;
  ``Cm2l`` holds the right eigenvectors as columns and ``Cl2m = Cm2l^-1``.

Pinned by closed forms (``Cl2m A Cm2l`` diagonal, ``Cl2m Cm2l = I``) and, for
the rigid-lid matrix, by the homogeneous-solution statistics printed in
``notebooks/dev/QG_Multilayer/mqg_demo_retangular.ipynb`` cell 18.  The
eigenvector normalisation is parity unpinned but does not change any
layer-space result (``Cm2l diag(.) Cl2m`` is invariant to column scaling).

The reference takes ``helmholtz_lambdas`` from ``jnp.linalg.eigh(A)``
(``transforms.py:194``) although ``A`` is not symmetric; JAX symmetrises its
input, so the reference values are ``eigvalsh((A + A.T)/2)``.
``reference_eigenvalues`` restates that; ``vertical_modes`` returns the
eigenvalues that belong to ``Cl2m``/``Cm2l``.
"""

from __future__ import annotations

import numpy as np


def coupling_matrix(H, g_prime, free_surface: bool = True) -> np.ndarray:
    """Layer stretching matrix ``A`` (units 1/m^2 * s^2 ... i.e. 1/(m * m/s^2))."""
    H = np.asarray(H, dtype=np.float64)
    g = np.asarray(g_prime, dtype=np.float64)
    nl = H.size
    A = np.zeros((nl, nl))
    top = 1.0 / (H[0] * g[0]) if free_surface else 0.0
    if nl == 1:
        A[0, 0] = top
        return A
    A[0, 0] = top + 1.0 / (H[0] * g[1])
    A[0, 1] = -1.0 / (H[0] * g[1])
    for k in range(1, nl - 1):
        A[k, k - 1] = -1.0 / (H[k] * g[k])
        A[k, k] = (1.0 / g[k] + 1.0 / g[k + 1]) / H[k]
        A[k, k + 1] = -1.0 / (H[k] * g[k + 1])
    A[-1, -2] = -1.0 / (H[-1] * g[-1])
    A[-1, -1] = 1.0 / (H[-1] * g[-1])
    return A


def vertical_modes(A: np.ndarray, f0: float):
    """Return ``(eigenvalues, Cl2m, Cm2l, rossby_radii)`` sorted by ascending eigenvalue."""
    w, R = np.linalg.eig(A)
    w = w.real
    R = R.real
    order = np.argsort(w)
    w = w[order]
    R = R[:, order]
    Cl2m = np.linalg.inv(R)
    with np.errstate(divide="ignore", invalid="ignore"):
        radii = 1.0 / (f0 * np.sqrt(w))
    return w, Cl2m, R, radii


def reference_eigenvalues(A: np.ndarray) -> np.ndarray:
    """What ``jnp.linalg.eigh(A)[0]`` returns for a non-symmetric ``A`` (input symmetrised)."""
    return np.linalg.eigvalsh(0.5 * (A + A.T))
