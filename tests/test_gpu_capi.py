"""The C ABI from a plain-C host (tests/capi/qg_step_driver.c, built with gcc
against include/somaxhip.h and libsomaxhip.so, device memory from the HIP
runtime): the same QG steps as the Python path, bitwise."""

from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent


def _build(tmp: Path) -> Path:
    exe = tmp / "qg_step_driver"
    lib = ROOT / "somax_amd" / "lib"
    cmd = ["gcc", "-O2", "-std=c11", "-D__HIP_PLATFORM_AMD__", f"-I{ROOT / 'include'}", "-I/opt/rocm/include",
           str(ROOT / "tests" / "capi" / "qg_step_driver.c"), "-o", str(exe), f"-L{lib}", "-lsomaxhip",
           "-L/opt/rocm/lib", "-lamdhip64", f"-Wl,-rpath,{lib}", "-Wl,-rpath,/opt/rocm/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return exe


@pytest.mark.parametrize("precision", ["fp32", "mixed", "fp64"])
def test_c_host_matches_python_path(gpu, tmp_path, precision):
    import torch

    from somax_amd import BaroclinicQGState, _lib
    from somax_amd import testcases as tc

    print(f"[capi] C driver, 3x128x128, {precision}", flush=True)
    exe = _build(tmp_path)
    model = tc.config_baroclinic(128, precision=precision)
    dtype = np.float64 if precision == "fp64" else np.float32
    q = tc.random_pv(model, seed=3).astype(dtype)
    nsteps, dt = 5, 1200.0
    p = model.params
    with open(tmp_path / "in.bin", "wb") as f:
        prec = {"fp32": _lib.PREC_FP32, "mixed": _lib.PREC_MIXED, "fp64": _lib.PREC_FP64}[precision]
        f.write(np.array([model.grid.nx, model.grid.ny, 3, nsteps, prec], dtype="<i4").tobytes())
        f.write(np.array([model.grid.dx, model.grid.dy, dt, p.lateral_viscosity, p.biharmonic_viscosity,
                          p.bottom_drag, p.wind_amplitude], dtype="<f8").tobytes())
        for a in (model._Cl2m, model._Cm2l, model._lambdas, model._beta_row, model._wind_row, model._wind_scale):
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        f.write(q.astype("<f8" if precision == "fp64" else "<f4").tobytes())
    res = subprocess.run([str(exe), str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True,
                         text=True, timeout=120)
    print(res.stdout.strip(), res.stderr.strip(), flush=True, file=sys.stdout)
    assert res.returncode == 0
    got = np.fromfile(tmp_path / "out.bin", dtype="<f8" if precision == "fp64" else "<f4").reshape(q.shape)
    want = model.step(BaroclinicQGState(q=torch.from_numpy(q).cuda()), dt, nsteps=nsteps, solver="ssprk3")
    assert np.array_equal(got, want.q.cpu().numpy())
