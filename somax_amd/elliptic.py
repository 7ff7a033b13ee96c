"""Helmholtz solve on a masked (non-rectangular) domain by the capacitance-
matrix method, on the HIP path.

Stands in for ``compute_capacitance_matrices`` / ``inverse_elliptic_dst_cmm``
of the former ``somax._src.operators.dst`` that
``notebooks/dev/QG_Multilayer/helmholtz_decomp.ipynb`` (cells 19-33) and
``mqg_demo_octagonal.ipynb`` (cell 12) call; that module is no longer in the
reference tree, so the arguments here are this library's own.

``MaskedHelmholtz.solve(rhs)`` returns ``psi`` with ``(lap - lam_p) psi_p =
rhs_p`` at the points where ``mask == 1`` and ``psi = 0`` elsewhere.  It runs
in ``libsomaxhip.so`` (``somaxhip_cmm_solve``): two rectangular DST-I
inversions of a QG plan (identity modal matrices, ghost ring kept) and an
FP64 capacitance correction at the points of the irregular boundary.  The
capacitance matrices are set-up tables, computed once in FP64 on the host.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .qg import _PREC


def _torch():
    import torch

    return torch


def irregular_boundary(mask):
    """``(yids, xids)``: the points with ``mask == 0`` that have a point of the
    mask in their 3x3 neighbourhood (helmholtz_decomp.ipynb cells 23-25),
    row-major."""
    m = np.asarray(mask) > 0.5
    p = np.pad(m, 1)
    near = np.zeros_like(m)
    for dy in range(3):
        for dx in range(3):
            near |= p[dy:dy + m.shape[0], dx:dx + m.shape[1]]
    return np.nonzero(~m & near)


def octagon_node_mask(nx: int, ny: int) -> np.ndarray:
    """``(ny, nx)`` node mask of the octagon of mqg_demo_octagonal.ipynb
    (cell 6): cells ``(nx-1, ny-1)`` with the corners i + j < min(nx, ny)//4
    cut, and a node inside when its four cells are (so the outer ring is
    outside)."""
    cells = np.ones((ny - 1, nx - 1))
    k = min(nx // 4, ny // 4)
    for i in range(ny // 4):
        for j in range(nx // 4):
            if i + j < k:
                cells[i, j] = cells[i, -1 - j] = cells[-1 - i, j] = cells[-1 - i, -1 - j] = 0.0
    c = np.pad(cells, 1)
    return ((c[1:, 1:] + c[:-1, 1:] + c[1:, :-1] + c[:-1, :-1]) > 3.5).astype(np.float64)


def _eig(n: int, d: float) -> np.ndarray:
    return (2.0 * np.cos(np.pi * np.arange(1, n + 1) / (n + 1)) - 2.0) / (d * d)


def _sine(n: int) -> np.ndarray:
    k = np.arange(1, n + 1, dtype=np.float64)
    return np.sqrt(2.0 / (n + 1)) * np.sin(np.pi * np.outer(k, k) / (n + 1))


def compute_capacitance_matrices(shape, dx: float, dy: float, lambdas, yids, xids, batch: int = 32) -> np.ndarray:
    """``(P, M, M)`` FP64: per plane the inverse of ``G[i, j] = (G e_j)[b_i]``,
    ``G = (lap - lam_p)^-1`` by DST-I over the full ``shape`` array.

    Set-up cost ``O(M Nx Ny Nx)`` per plane (about 5 s on 8 cores for the
    1020 boundary points of a 256^2 octagon): each unit source's transform is
    ``W * outer(Sy[:, y_j], Sx[:, x_j])``, multiplied by the x sine matrix in
    batches, and only the rows at the boundary points are summed over y."""
    Ny, Nx = shape
    lambdas = np.atleast_1d(np.asarray(lambdas, dtype=np.float64))
    yids, xids = np.asarray(yids), np.asarray(xids)
    M = yids.size
    Sy, Sx = _sine(Ny), _sine(Nx)
    D = _eig(Ny, dy)[:, None] + _eig(Nx, dx)[None, :]
    rows = Sy[yids, :]  # (M, Ny): y synthesis at the boundary rows
    out = np.empty((lambdas.size, M, M))
    for p, lam in enumerate(lambdas):
        w = 1.0 / (D - lam)
        g = np.empty((M, M))
        for lo in range(0, M, batch):
            hi = min(M, lo + batch)
            hat = Sy[:, yids[lo:hi]].T[:, :, None] * (w[None] * Sx[:, xids[lo:hi]].T[:, None, :])
            t = np.matmul(hat, Sx)[:, :, xids]  # (B, Ny, M): x synthesis at the boundary columns
            g[:, lo:hi] = np.einsum("ik,bki->ib", rows, t)
        out[p] = np.linalg.inv(g)
    return out


class MaskedHelmholtz:
    """Capacitance-matrix Helmholtz solver for ``P`` planes of ``(Ny, Nx)``
    sharing one mask, one Helmholtz parameter per plane."""

    def __init__(self, mask, dx: float, dy: float, lambdas, *, precision: str = "mixed", device="cuda"):
        torch = _torch()
        allowed = sorted(k for k in _PREC if k != "fp64")  # the masked solve runs on FP32 / BF16 plans
        if precision not in allowed:
            raise ValueError(f"precision must be one of {allowed}, got {precision!r}")
        mask = np.ascontiguousarray(np.asarray(mask, dtype=np.float32) > 0.5, dtype=np.float32)
        if mask.ndim != 2 or min(mask.shape) < 4:
            raise ValueError(f"mask must be 2-D with at least 4 points per side, got {mask.shape}")
        self.mask = mask
        self.shape = mask.shape
        self.dx, self.dy = float(dx), float(dy)
        self.lambdas = np.atleast_1d(np.asarray(lambdas, dtype=np.float64))
        self.P = self.lambdas.size
        self.precision = precision
        self.device = device
        self.yids, self.xids = irregular_boundary(mask)
        self.cap = np.ascontiguousarray(
            compute_capacitance_matrices(self.shape, self.dx, self.dy, self.lambdas, self.yids, self.xids))
        lib = _lib.load()
        _lib.init_device()
        Ny, Nx = self.shape
        keep = []

        def dp(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return a.ctypes.data_as(_lib.c_double_p)

        d = _lib.QGDesc()
        d.nx, d.ny, d.nl = Nx - 2, Ny - 2, self.P
        d.dx, d.dy = self.dx, self.dy
        eye = np.eye(self.P)
        d.cl2m, d.cm2l, d.lambdas = dp(eye), dp(eye), dp(self.lambdas)
        d.beta_y_row, d.wind_row, d.wind_scale = dp(np.zeros(Ny)), dp(np.zeros(Ny)), dp(np.zeros(self.P))
        d.zero_psi_ghost = 0
        d.advection = _lib.ADV_ARAKAWA
        d.precision = _PREC[precision]
        d.mass_correction = 0
        plan = ctypes.c_void_p()
        _lib.check(lib.somaxhip_qg_plan_create(ctypes.byref(d), ctypes.byref(plan)), "somaxhip_qg_plan_create")
        self._plan = plan
        bidx = np.ascontiguousarray(self.yids * Nx + self.xids, dtype=np.int32)
        cmm = ctypes.c_void_p()
        _lib.check(
            lib.somaxhip_cmm_create(plan, mask.ctypes.data_as(_lib.c_float_p), int(bidx.size),
                                    bidx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                    self.cap.ctypes.data_as(_lib.c_double_p), ctypes.byref(cmm)),
            "somaxhip_cmm_create",
        )
        self._cmm = cmm
        nbytes = int(lib.somaxhip_cmm_workspace_bytes(cmm))
        self._ws = torch.empty(nbytes, dtype=torch.uint8, device=device)

    @property
    def num_boundary_points(self) -> int:
        return int(self.yids.size)

    def __del__(self):
        lib = None
        try:
            lib = _lib.load()
        except Exception:
            return
        if getattr(self, "_cmm", None) is not None:
            lib.somaxhip_cmm_destroy(self._cmm)
        if getattr(self, "_plan", None) is not None:
            lib.somaxhip_qg_plan_destroy(self._plan)

    def solve(self, rhs):
        """``psi`` (float32 device tensor ``(P, Ny, Nx)``) for ``rhs`` of the
        same shape."""
        torch = _torch()
        shape = (self.P, *self.shape)
        if not isinstance(rhs, torch.Tensor) or rhs.device.type != "cuda":
            rhs = torch.as_tensor(np.asarray(rhs), dtype=torch.float32, device=self.device)
        if rhs.dtype != torch.float32:
            raise TypeError(f"rhs must be float32, got {rhs.dtype}")
        if tuple(rhs.shape) != shape:
            raise ValueError(f"rhs has shape {tuple(rhs.shape)}, expected {shape}")
        rhs = rhs.contiguous()
        psi = torch.empty_like(rhs)
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(_lib.load().somaxhip_cmm_solve(self._cmm, rhs.data_ptr(), psi.data_ptr(), self._ws.data_ptr(),
                                                  stream), "somaxhip_cmm_solve")
        return psi
