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
        lib.fftcpu_prepare.restype = None
        lib.fftcpu_prepare.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int] + [_fp] * 6
        lib.fftcpu_divide.restype = None
        lib.fftcpu_divide.argtypes = [ctypes.c_long, _fp, _fp]
        lib.fftcpu_finish.restype = None
        lib.fftcpu_finish.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp, _fp, _dp, _fp, _dp, ctypes.c_int,
                                      _fp, _dp]
        lib.fftcpu_combine.restype = None
        lib.fftcpu_combine.argtypes = [ctypes.c_long, _fp, ctypes.c_float, _fp, ctypes.c_float, _fp, ctypes.c_float,
                                       _fp]
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
        self._lib = _load_rhs()
        self._lib.qgrhs_set_threads(self.threads)
        f32 = np.float32
        self.cl2m = np.ascontiguousarray(s.Cl2m, dtype=f32)
        self.cm2l = np.ascontiguousarray(s.Cm2l, dtype=f32)
        self.cm2l64 = np.ascontiguousarray(s.Cm2l, dtype=np.float64)
        den = dst.dst1_eigenvalues(s.Ny, s.dy)[:, None] + dst.dst1_eigenvalues(s.Nx, s.dx)[None, :]
        self.den = np.ascontiguousarray((den[None] - np.asarray(s.lambdas)[:, None, None]).astype(f32))
        self.beta_row = np.ascontiguousarray(np.broadcast_to(s.beta_y, (s.Ny, s.Nx))[:, 0], dtype=f32)
        self.mass = bool(s.mass_correction)
        if self.mass:
            self.homsol = np.ascontiguousarray(s.homsol, dtype=f32)
            self.hmean = np.ascontiguousarray(s.homsol_mean, dtype=np.float64)
        self.shape = (s.nl, s.Ny, s.Nx)

    def rhs(self, q: np.ndarray) -> np.ndarray:
        """The RHS of the boundary-conditioned state BC(q): elementwise passes
        in C++ (OpenMP), sine transforms by pocketfft."""
        import scipy.fft

        s, w, lib = self.s, self.threads, self._lib
        p = lambda a: a.ctypes.data_as(_fp)  # noqa: E731
        qbc, qt, qm, psi = (np.empty(self.shape, dtype=np.float32) for _ in range(4))
        lib.fftcpu_prepare(s.nl, s.Ny, s.Nx, p(q), p(self.beta_row), p(self.cl2m), p(qbc), p(qt), p(qm))
        f = scipy.fft.dstn(qm, type=1, axes=(-2, -1), norm="ortho", workers=w, overwrite_x=True)
        lib.fftcpu_divide(f.size, p(f), p(self.den))
        pm = np.ascontiguousarray(scipy.fft.dstn(f, type=1, axes=(-2, -1), norm="ortho", workers=w,
                                                 overwrite_x=True), dtype=np.float32)
        wall = np.zeros(s.nl)
        lib.fftcpu_finish(s.nl, s.Ny, s.Nx, p(pm), p(self.homsol) if self.mass else None,
                          self.hmean.ctypes.data_as(_dp) if self.mass else None, p(self.cm2l),
                          self.cm2l64.ctypes.data_as(_dp), int(s.zero_psi_ghost), p(psi), wall.ctypes.data_as(_dp))
        return qg_stencil_rhs(s, psi, qbc, qt, wall, dtype=np.float32)

    def step(self, q: np.ndarray, dt: float, nsteps: int = 1) -> np.ndarray:
        """``nsteps`` SSP-RK3 steps (Shu-Osher form) from ``q``."""
        h = float(dt)
        q = np.array(np.asarray(q, dtype=np.float32).reshape(self.shape), order="C")
        q1, q2 = np.empty_like(q), np.empty_like(q)
        n = q.size
        p = lambda a: a.ctypes.data_as(_fp)  # noqa: E731
        comb = self._lib.fftcpu_combine
        for _ in range(nsteps):
            comb(n, p(q1), 0.0, p(q), 1.0, p(q), h, p(self.rhs(q)))
            comb(n, p(q2), 0.75, p(q), 0.25, p(q1), h, p(self.rhs(q1)))
            comb(n, p(q), 1.0 / 3.0, p(q), 2.0 / 3.0, p(q2), h, p(self.rhs(q2)))
        return q
