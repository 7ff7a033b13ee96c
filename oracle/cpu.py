"""ctypes front of oracle/cpu_step.cpp, the C++/OpenMP restatement of the
QG SSP-RK3 step used as the multi-core CPU baseline.

TEST INFRASTRUCTURE ONLY (see the package docstring): loaded by tests and by
bench.py's cpu_baseline leg, never by the product.  ``build()`` compiles it
with g++ into ``oracle/_build/libqgcpu.so`` (x86-64-v3, OpenMP).
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SRC = HERE / "cpu_step.cpp"
LIB = HERE / "_build" / "libqgcpu.so"

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
