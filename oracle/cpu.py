"""ctypes fronts of the oracle's compiled restatements.

* ``oracle/cpu_step.cpp``: the C++/OpenMP Arakawa QG step with dense DST-I
  GEMMs (``CpuStep``), checked against the NumPy oracle.
* ``oracle/qgrhs.cpp``: the stencil terms of ``oracle.qg.rhs`` in FP64 (behind
  ``rhs(..., fast=True)``, the large-grid parity checks: NumPy's bits) and
  FP32 (the stencils of ``FftCpuStep``, the CPU baseline: the SSP-RK3 step
  with FFT-based DST-I, as the reference's JAX path computes it on a CPU).

TEST INFRASTRUCTURE ONLY (see the package docstring): loaded by tests and by
bench.py's cpu_baseline leg, never by the product.  ``build()`` compiles them
with g++ into ``oracle/_build/libqgcpu.so`` and ``oracle/_build/libqgrhs.so``
(x86-64-v3, OpenMP; the stencils without floating-point contraction, so the
FP64 instance keeps NumPy's rounding).
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SRC = HERE / "cpu_step.cpp"
LIB = HERE / "_build" / "libqgcpu.so"
RHS_SRC = HERE / "qgrhs.cpp"
RHS_LIB = HERE / "_build" / "libqgrhs.so"

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)


def build() -> Path:
    LIB.parent.mkdir(parents=True, exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-fopenmp", "-std=c++17", "-shared", "-fPIC", str(SRC), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed for {SRC.name}:\n{res.stdout}\n{res.stderr}")
    tmp.replace(LIB)
    tmp = RHS_LIB.with_suffix(".so.tmp")
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fopenmp", "-std=c++17", "-shared", "-fPIC",
           "-Wall", str(RHS_SRC), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed for {RHS_SRC.name}:\n{res.stdout}\n{res.stderr}")
    tmp.replace(RHS_LIB)
    return LIB


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} is missing: run oracle.cpu.build()")
        lib = ctypes.CDLL(str(LIB))
        lib.qgcpu_create.restype = ctypes.c_void_p
        lib.qgcpu_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_int]
        lib.qgcpu_step.restype = None
        lib.qgcpu_step.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        lib.qgcpu_destroy.restype = None
        lib.qgcpu_destroy.argtypes = [ctypes.c_void_p]
        lib.qgcpu_max_threads.restype = ctypes.c_int
        lib.qgcpu_set_threads.argtypes = [ctypes.c_int]
        _lib = lib
    return _lib


class CpuStep:
    """The Arakawa QG step of an oracle ``QGSetup`` (no WENO, no mass
    correction) in FP32 or FP64 on ``threads`` host threads."""

    def __init__(self, setup, fp64: bool = False, threads: int | None = None):
        if setup.advection != "arakawa" or setup.mass_correction:
            raise ValueError("the C++ step restates the Arakawa path without mass correction")
        lib = load()
        if threads:
            lib.qgcpu_set_threads(int(threads))
        self.threads = int(lib.qgcpu_max_threads())
        self.fp64 = fp64
        self.dtype = np.float64 if fp64 else np.float32
        keep = [np.ascontiguousarray(a, dtype=np.float64) for a in
                (setup.Cl2m, setup.Cm2l, setup.lambdas, setup.beta_y[:, 0], setup.wind_forcing[:, 0],
                 setup.wind_scale)]
        self._keep = keep
        p = [k.ctypes.data_as(_dp) for k in keep]
        self._h = lib.qgcpu_create(setup.nx, setup.ny, setup.nl, setup.dx, setup.dy, *p, int(setup.zero_psi_ghost),
                                   setup.nu, setup.nu4, setup.kappa, setup.tau0, int(fp64))
        self.shape = (setup.nl, setup.Ny, setup.Nx)

    def step(self, q: np.ndarray, dt: float, nsteps: int = 1) -> np.ndarray:
        out = np.ascontiguousarray(np.asarray(q, dtype=self.dtype).reshape(self.shape)).copy()
        load().qgcpu_step(self._h, out.ctypes.data, int(nsteps), float(dt))
        return out

    def __del__(self):
        if getattr(self, "_h", None):
            load().qgcpu_destroy(self._h)
            self._h = None


_rhs_lib = None
_fp = ctypes.POINTER(ctypes.c_float)


def _load_rhs():
    global _rhs_lib
    if _rhs_lib is None:
        if not RHS_LIB.exists():
            raise RuntimeError(f"{RHS_LIB} is missing: run oracle.cpu.build()")
        lib = ctypes.CDLL(str(RHS_LIB))
        for name, p in (("qgrhs_stencil_f64", _dp), ("qgrhs_stencil_f32", _fp)):
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                           p, p, p, p, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, p]
        lib.qgrhs_set_threads.argtypes = [ctypes.c_int]
        lib.qgrhs_set_threads.restype = None
        lib.qgrhs_max_threads.restype = ctypes.c_int
        _rhs_lib = lib
    return _rhs_lib


def qg_stencil_rhs(s, psi, q, qt, wall, dtype=np.float64) -> np.ndarray:
    """Stencil terms of ``oracle.qg.rhs`` for setup ``s``: advection of ``qt``
    by ``psi``, wind, bottom drag, harmonic and biharmonic viscosity, in
    ``dtype`` (float64: NumPy's bits; float32: the CPU baseline)."""
    lib = _load_rhs()
    dtype = np.dtype(dtype)
    fn, p = (lib.qgrhs_stencil_f64, _dp) if dtype == np.float64 else (lib.qgrhs_stencil_f32, _fp)
    shape = (s.nl, s.Ny, s.Nx)
    arrs = [np.ascontiguousarray(np.broadcast_to(a, shape), dtype=dtype) for a in (psi, q, qt)]
    wind = np.ascontiguousarray(np.broadcast_to(s.wind_forcing, shape[1:]), dtype=dtype)
    coef = np.ascontiguousarray(s.tau0 * np.asarray(s.wind_scale, dtype=np.float64))
    wall = np.ascontiguousarray(np.broadcast_to(np.asarray(wall, dtype=np.float64), (s.nl,)))
    out = np.empty(shape, dtype=dtype)
    fn(s.nl, s.Ny, s.Nx, s.dx, s.dy, 1 if s.advection == "weno" else 0, arrs[0].ctypes.data_as(p),
       arrs[1].ctypes.data_as(p), arrs[2].ctypes.data_as(p), wind.ctypes.data_as(p), coef.ctypes.data_as(_dp),
       float(s.kappa), float(s.nu), float(s.nu4), wall.ctypes.data_as(_dp), out.ctypes.data_as(p))
    return out


def host_threads() -> int:
    """Host threads this process may use (its CPU affinity)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


class FftCpuStep:
    """The SSP-RK3 step of an oracle ``QGSetup`` in FP32 on the host, computed
    as the reference's JAX path computes it on a CPU: the DST-I by FFT
    (pocketfft through ``scipy.fft``, single precision), the layer/mode
    projections as small matrix products, the stencil terms of
    ``oracle/qgrhs.cpp`` (FP32, OpenMP), FP64 domain means for the mass
    constraint.  Arakawa or WENO-Z, with or without the mass constraint, on
    ``threads`` host threads (transforms and stencils)."""

    def __init__(self, setup, threads: int | None = None):
        import scipy.fft  # noqa: F401

        from . import dst

        s = setup
        self.s = s
        self.threads = int(threads or host_threads())
        _load_rhs().qgrhs_set_threads(self.threads)
        f32 = np.float32
        self.cl2m = np.asarray(s.Cl2m, dtype=f32)
        self.cm2l = np.asarray(s.Cm2l, dtype=f32)
        den = dst.dst1_eigenvalues(s.Ny, s.dy)[:, None] + dst.dst1_eigenvalues(s.Nx, s.dx)[None, :]
        self.den = (den[None] - np.asarray(s.lambdas)[:, None, None]).astype(f32)
        self.beta_y = np.broadcast_to(s.beta_y, (s.Ny, s.Nx)).astype(f32)
        self.mass = bool(s.mass_correction)
        if self.mass:
            self.homsol = np.asarray(s.homsol, dtype=f32)
            self.hmean = np.asarray(s.homsol_mean, dtype=np.float64)
        self.shape = (s.nl, s.Ny, s.Nx)

    def _mix(self, M, f):
        return (M @ f.reshape(self.s.nl, -1)).reshape(self.shape)

    def rhs(self, q: np.ndarray) -> np.ndarray:
        import scipy.fft

        s, w = self.s, self.threads
        qm = self._mix(self.cl2m, q)
        f = scipy.fft.dstn(qm, type=1, axes=(-2, -1), norm="ortho", workers=w)
        pm = scipy.fft.dstn(f / self.den, type=1, axes=(-2, -1), norm="ortho", workers=w)
        wall = np.zeros(s.nl)
        if self.mass:
            alpha = -pm.mean(axis=(-2, -1), dtype=np.float64) / self.hmean
            pm = pm + alpha.astype(np.float32)[:, None, None] * self.homsol
            wall = np.asarray(s.Cm2l) @ alpha
        psi = self._mix(self.cm2l, pm)
        if s.zero_psi_ghost:
            psi[:, 0, :] = psi[:, -1, :] = 0.0
            psi[:, :, 0] = psi[:, :, -1] = 0.0
        return qg_stencil_rhs(s, psi, q, q + self.beta_y, wall, dtype=np.float32)

    def _bc(self, q):
        g = q.copy()
        g[:, 0, :] = g[:, -1, :] = 0.0
        g[:, :, 0] = g[:, :, -1] = 0.0
        return g

    def step(self, q: np.ndarray, dt: float, nsteps: int = 1) -> np.ndarray:
        h = np.float32(dt)
        q = np.ascontiguousarray(np.asarray(q, dtype=np.float32).reshape(self.shape))
        for _ in range(nsteps):
            q1 = q + h * self.rhs(self._bc(q))
            q2 = np.float32(0.75) * q + np.float32(0.25) * (q1 + h * self.rhs(self._bc(q1)))
            q = q / np.float32(3.0) + np.float32(2.0 / 3.0) * (q2 + h * self.rhs(self._bc(q2)))
        return q
