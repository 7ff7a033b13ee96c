"""Capacitance-matrix solve of the Helmholtz problem on a masked domain, FP64.

TEST INFRASTRUCTURE ONLY (see the package docstring).

The reference demonstrates it in notebooks/dev/QG_Multilayer/
helmholtz_decomp.ipynb (cells 19-33) and mqg_demo_octagonal.ipynb (cells 6,
12) through ``compute_capacitance_matrices`` and ``inverse_elliptic_dst_cmm``
of ``somax._src.operators.dst``, a module no longer in the tree.  The method
is the capacitance-matrix method of MQGeometry (Thiry et al., "MQGeometry-1.0",
GMD 2024), restated here from its published description:

* domain ``D = {mask == 1}``; irregular boundary ``B``: points with
  ``mask == 0`` and a point of ``D`` in their 3x3 neighbourhood
  (helmholtz_decomp.ipynb cells 23-25);
* ``G``: the rectangular DST-I Helmholtz solve over the full array (zero
  values beyond it), ``(lap - lam)^-1``;
* capacitance matrix ``C = (G restricted to B x B)^-1`` per plane, built from
  unit sources at the points of ``B``;
* solve: ``r0 = r * mask``; ``psi1 = G r0``; ``r0[B] = -C psi1[B]``;
  ``psi = (G r0) * mask``.  Then ``psi[B] = 0`` exactly, so ``psi`` solves
  ``(lap - lam) psi = r`` on ``D`` with ``psi = 0`` off ``D``.

Pinned by mathematics: the masked Dirichlet problem has one solution, which a
sparse direct solve over the points of ``D`` gives independently
(``masked_direct``); the notebook's own round trip (cells 29-33: a masked
field ``f`` is recovered from ``(lap - lam) f * mask``); and, for the octagon
mask of mqg_demo_octagonal.ipynb cell 6, the homogeneous-solution means that
cell 12 prints.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import dst


def octagon_center_mask(nx: int, ny: int) -> np.ndarray:
    """mqg_demo_octagonal.ipynb cell 6: ``(nx-1, ny-1)`` cell mask with the
    four corners cut along i + j < min(nx, ny) // 4."""
    mask = np.ones((nx - 1, ny - 1))
    k = min(nx // 4, ny // 4)
    for i in range(nx // 4):
        for j in range(ny // 4):
            if i + j < k:
                mask[i, j] = mask[i, -1 - j] = mask[-1 - i, j] = mask[-1 - i, -1 - j] = 0.0
    return mask


def irregular_boundary(mask: np.ndarray):
    """``(yids, xids)`` of the points outside ``mask`` that touch it in a 3x3
    window (helmholtz_decomp.ipynb cells 23-25), in row-major order."""
    m = np.asarray(mask) > 0.5
    p = np.pad(m, 1)
    near = np.zeros_like(m)
    for dy in (0, 1, 2):
        for dx in (0, 1, 2):
            near |= p[dy:dy + m.shape[0], dx:dx + m.shape[1]]
    return np.nonzero(~m & near)


def capacitance_matrices(shape, dx: float, dy: float, lambdas, yids, xids) -> np.ndarray:
    """``(P, M, M)``: per plane the inverse of ``G`` sampled at the boundary
    points, ``G[i, j] = (G e_j)[b_i]``."""
    lambdas = np.atleast_1d(np.asarray(lambdas, dtype=np.float64))
    M = len(yids)
    out = np.zeros((lambdas.size, M, M))
    for p, lam in enumerate(lambdas):
        g = np.empty((M, M))
        for j in range(M):
            e = np.zeros(shape)
            e[yids[j], xids[j]] = 1.0
            g[:, j] = dst.helmholtz_dst(e, dx, dy, lam)[yids, xids]
        out[p] = np.linalg.inv(g)
    return out


def solve(r, mask, dx: float, dy: float, lambdas, cap, yids, xids) -> np.ndarray:
    """The capacitance-matrix solve of ``(lap - lam_p) psi_p = r_p`` on the
    masked domain, for ``r`` of shape ``(P, Ny, Nx)``."""
    lambdas = np.atleast_1d(np.asarray(lambdas, dtype=np.float64))
    r0 = np.asarray(r, dtype=np.float64) * mask
    psi1 = dst.helmholtz_dst(r0, dx, dy, lambdas)
    if len(yids):
        for p in range(r0.shape[0]):
            r0[p, yids, xids] = -cap[p] @ psi1[p, yids, xids]
    return dst.helmholtz_dst(r0, dx, dy, lambdas) * mask


def masked_direct(r, mask, dx: float, dy: float, lam: float) -> np.ndarray:
    """Sparse direct solve of ``(lap - lam) psi = r`` at the points of ``D``
    with ``psi = 0`` elsewhere (5-point Laplacian), one plane."""
    m = np.asarray(mask) > 0.5
    ny, nx = m.shape
    idx = -np.ones(m.shape, dtype=np.int64)
    idx[m] = np.arange(m.sum())
    rows, cols, vals = [], [], []
    for j, i in zip(*np.nonzero(m)):
        k = idx[j, i]
        rows.append(k)
        cols.append(k)
        vals.append(-2.0 / dx**2 - 2.0 / dy**2 - lam)
        for dj, di, w in ((0, 1, dx), (0, -1, dx), (1, 0, dy), (-1, 0, dy)):
            jj, ii = j + dj, i + di
            if 0 <= jj < ny and 0 <= ii < nx and m[jj, ii]:
                rows.append(k)
                cols.append(idx[jj, ii])
                vals.append(1.0 / w**2)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(m.sum(), m.sum()))
    out = np.zeros(m.shape)
    out[m] = spla.spsolve(A, np.asarray(r, dtype=np.float64)[m])
    return out
