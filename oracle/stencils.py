"""C-grid stencil operators on ghosted ``(..., Ny, Nx)`` arrays (FP64).

Restates the finitevolX v0.0.39 operators that somax calls on the hot path
(``somax/_src/models/qg/baroclinic.py:158,175,184,188``;
``somax/_src/models/swm/multilayer.py:160-199``).  finitevolX cannot be
imported here, so each index convention below is an assumption stated in the
docstring (parity unpinned against finitevolX itself); each operator is pinned
by closed forms or invariants in ``tests/test_oracle_pins.py``.

Grid: T point ``(j, i)`` at ``(i dx, j dy)``; U ``(j, i)`` on the east face
``(i+1/2)``; V ``(j, i)`` on the north face ``(j+1/2)``; X ``(j, i)`` at the
north-east corner.  Row 0/col 0 and row -1/col -1 form the ghost ring.
Operators write the interior ``[1:-1, 1:-1]`` and leave the ghost ring 0.
"""

from __future__ import annotations

import numpy as np


def zero_boundaries(f: np.ndarray) -> np.ndarray:
    """``finitevolx.zero_boundaries``: ghost ring set to 0, interior kept."""
    g = np.array(f, dtype=np.float64, copy=True)
    g[..., 0, :] = 0.0
    g[..., -1, :] = 0.0
    g[..., :, 0] = 0.0
    g[..., :, -1] = 0.0
    return g


def _interior_like(f):
    return np.zeros_like(np.asarray(f, dtype=np.float64))


def laplacian(f: np.ndarray, dx: float, dy: float) -> np.ndarray:
    """``Difference2D.laplacian``: 5-point Laplacian on the interior, ghost ring 0."""
    f = np.asarray(f, dtype=np.float64)
    out = _interior_like(f)
    c = f[..., 1:-1, 1:-1]
    out[..., 1:-1, 1:-1] = (f[..., 1:-1, 2:] - 2.0 * c + f[..., 1:-1, :-2]) / (dx * dx) + (
        f[..., 2:, 1:-1] - 2.0 * c + f[..., :-2, 1:-1]
    ) / (dy * dy)
    return out


def biharmonic(f: np.ndarray, dx: float, dy: float) -> np.ndarray:
    """``laplacian(laplacian(f))``; the inner Laplacian's ghost ring is 0.

    Extension (no reference code): the hyperviscosity ``-nu4 del^4 q`` acts on
    q, mirroring the reference's harmonic ``nu del^2 q`` (baroclinic.py:188)."""
    return laplacian(laplacian(f, dx, dy), dx, dy)


def arakawa_jacobian(p: np.ndarray, q: np.ndarray, dx: float, dy: float) -> np.ndarray:
    """Arakawa (1966) energy- and enstrophy-conserving Jacobian ``J(p, q) = p_x q_y - p_y q_x``.

    Returns the interior ``(..., Ny-2, Nx-2)`` like ``finitevolx.arakawa_jacobian``
    (baroclinic.py:175).  It is the mean of the three second-order forms
    J++, J+x and Jx+ and reads the ghost ring of both inputs.
    """
    p = np.asarray(p, dtype=np.float64)
    q = np.asarray(q, dtype=np.float64)

    def s(a, dj, di):
        ny, nx = a.shape[-2:]
        return a[..., 1 + dj : ny - 1 + dj, 1 + di : nx - 1 + di]

    pE, pW, pN, pS = s(p, 0, 1), s(p, 0, -1), s(p, 1, 0), s(p, -1, 0)
    pNE, pNW, pSE, pSW = s(p, 1, 1), s(p, 1, -1), s(p, -1, 1), s(p, -1, -1)
    qE, qW, qN, qS = s(q, 0, 1), s(q, 0, -1), s(q, 1, 0), s(q, -1, 0)
    qNE, qNW, qSE, qSW = s(q, 1, 1), s(q, 1, -1), s(q, -1, 1), s(q, -1, -1)

    jpp = (pE - pW) * (qN - qS) - (pN - pS) * (qE - qW)
    jpx = pE * (qNE - qSE) - pW * (qNW - qSW) - pN * (qNE - qNW) + pS * (qSE - qSW)
    jxp = qN * (pNE - pNW) - qS * (pSE - pSW) - qE * (pNE - pSE) + qW * (pNW - pSW)
    return (jpp + jpx + jxp) / (12.0 * dx * dy)


# --------------------------------------------------------------------------
# WENO-Z flux-form PV advection (extension: no code in the current reference;
# the removed API used method="wenoz", num_pts=5, see
# notebooks/dev/QG_Multilayer/mqg_demo_retangular.ipynb cell 7).
# --------------------------------------------------------------------------

WENO_EPS = 1e-20


def weno5z_left(qm2, qm1, q0, qp1, qp2, eps=WENO_EPS):
    """WENO-Z (Borges et al. 2008, p = 1) value at face i+1/2 from cells i-2..i+2, upwind = left."""
    p0 = (2.0 * qm2 - 7.0 * qm1 + 11.0 * q0) / 6.0
    p1 = (-qm1 + 5.0 * q0 + 2.0 * qp1) / 6.0
    p2 = (2.0 * q0 + 5.0 * qp1 - qp2) / 6.0
    b0 = 13.0 / 12.0 * (qm2 - 2.0 * qm1 + q0) ** 2 + 0.25 * (qm2 - 4.0 * qm1 + 3.0 * q0) ** 2
    b1 = 13.0 / 12.0 * (qm1 - 2.0 * q0 + qp1) ** 2 + 0.25 * (qm1 - qp1) ** 2
    b2 = 13.0 / 12.0 * (q0 - 2.0 * qp1 + qp2) ** 2 + 0.25 * (3.0 * q0 - 4.0 * qp1 + qp2) ** 2
    tau = np.abs(b0 - b2)
    a0 = 0.1 * (1.0 + tau / (b0 + eps))
    a1 = 0.6 * (1.0 + tau / (b1 + eps))
    a2 = 0.3 * (1.0 + tau / (b2 + eps))
    return (a0 * p0 + a1 * p1 + a2 * p2) / (a0 + a1 + a2)


def weno3z_left(qm1, q0, qp1, eps=WENO_EPS):
    """3rd-order WENO-Z value at face i+1/2 from cells i-1..i+1, upwind = left."""
    p0 = (-qm1 + 3.0 * q0) / 2.0
    p1 = (q0 + qp1) / 2.0
    b0 = (q0 - qm1) ** 2
    b1 = (qp1 - q0) ** 2
    tau = np.abs(b0 - b1)
    a0 = (1.0 / 3.0) * (1.0 + tau / (b0 + eps))
    a1 = (2.0 / 3.0) * (1.0 + tau / (b1 + eps))
    return (a0 * p0 + a1 * p1) / (a0 + a1)


def _face_values_1d(qi: np.ndarray, vel: np.ndarray) -> np.ndarray:
    """Upwind WENO face values along the last axis.

    ``qi[..., c-1]`` holds interior cell c = 1..n; ``vel[..., f-1]`` is the
    velocity on face f+1/2 (between cells f and f+1) for f = 1..n-1.  Returns
    the reconstructed value on those n-1 faces.  The order drops near the
    walls: WENO5 where the five upwind-biased cells are all interior, else
    WENO3 where three are, else first-order upwind.
    """
    n = qi.shape[-1]
    # pad two cells each side so every stencil slice exists; padded cells are
    # only ever selected away by the masks below.
    p = np.concatenate([qi[..., :1], qi[..., :1], qi, qi[..., -1:], qi[..., -1:]], axis=-1)
    Q = lambda off: p[..., 2 + off : 2 + off + (n - 1)]  # noqa: E731 cell f+off, f = 1..n-1
    f = np.arange(1, n)
    l5 = weno5z_left(Q(-2), Q(-1), Q(0), Q(1), Q(2))
    l3 = weno3z_left(Q(-1), Q(0), Q(1))
    left = np.where((f - 2 >= 1) & (f + 2 <= n), l5, np.where((f - 1 >= 1) & (f + 1 <= n), l3, Q(0)))
    g = f + 1
    r5 = weno5z_left(Q(3), Q(2), Q(1), Q(0), Q(-1))
    r3 = weno3z_left(Q(2), Q(1), Q(0))
    right = np.where((g - 2 >= 1) & (g + 2 <= n), r5, np.where((g - 1 >= 1) & (g + 1 <= n), r3, Q(1)))
    return np.where(vel > 0.0, left, right)


def corner_streamfunction(psi: np.ndarray, psi_wall=0.0) -> np.ndarray:
    """psi at the (n+1) x (n+1) cell corners bounding the interior block.

    Corner ``(b, a)`` (b, a = 0..n) sits at ``(a+1/2, b+1/2)`` in T-index
    units.  Corners strictly inside the block take the 4-point T average;
    corners on the block's boundary lines take the wall value ``psi_wall``
    (a scalar, or one value per leading index), so the normal velocity through
    the walls is exactly 0 and the face velocities are exactly divergence free.
    """
    psi = np.asarray(psi, dtype=np.float64)
    c = 0.25 * (psi[..., :-1, :-1] + psi[..., :-1, 1:] + psi[..., 1:, :-1] + psi[..., 1:, 1:])
    w = np.asarray(psi_wall, dtype=np.float64)
    w = w[..., None] if w.ndim else w
    c[..., 0, :] = w
    c[..., -1, :] = w
    c[..., :, 0] = w
    c[..., :, -1] = w
    return c


def weno_velocities(psi: np.ndarray, dx: float, dy: float, psi_wall=0.0):
    """Face velocities from the corner streamfunction.

    ``u[..., j-1, f-1]`` on the x-face between interior cells (j, f), (j, f+1)
    is ``-dpsi/dy``; ``v[..., g-1, i-1]`` on the y-face between (g, i), (g+1, i)
    is ``dpsi/dx``."""
    C = corner_streamfunction(psi, psi_wall)
    n_y, n_x = C.shape[-2] - 1, C.shape[-1] - 1
    u = -(C[..., 1:, 1:n_x] - C[..., :-1, 1:n_x]) / dy
    v = (C[..., 1:n_y, 1:] - C[..., 1:n_y, :-1]) / dx
    return u, v


def weno_advection(psi: np.ndarray, q: np.ndarray, dx: float, dy: float, psi_wall=0.0) -> np.ndarray:
    """Flux-form tendency ``-div(u q)`` on the interior, ghost ring 0.

    Face velocities from ``weno_velocities``; face values of q from upwind
    WENO-Z (``_face_values_1d``).  Zero flux through the four walls, so the
    interior sum of the tendency is 0 (conservation of sum q).
    """
    q = np.asarray(q, dtype=np.float64)
    u, v = weno_velocities(psi, dx, dy, psi_wall)
    qi = q[..., 1:-1, 1:-1]
    n_y, n_x = qi.shape[-2:]
    qx = _face_values_1d(qi, u)
    qy = np.swapaxes(_face_values_1d(np.swapaxes(qi, -1, -2), np.swapaxes(v, -1, -2)), -1, -2)
    Fx = np.zeros(qi.shape[:-1] + (n_x + 1,))
    Fx[..., 1:-1] = u * qx
    Fy = np.zeros(qi.shape[:-2] + (n_y + 1, n_x))
    Fy[..., 1:-1, :] = v * qy
    out = np.zeros_like(q)
    out[..., 1:-1, 1:-1] = -(Fx[..., 1:] - Fx[..., :-1]) / dx - (Fy[..., 1:, :] - Fy[..., :-1, :]) / dy
    return out
