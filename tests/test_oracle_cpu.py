"""The C++/OpenMP restatement of the QG step (oracle/cpu_step.cpp, the
multi-core CPU baseline) against the NumPy oracle: FP64 to round-off, FP32
to its precision, with viscosity, hyperviscosity, drag, beta and wind on."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import cpu, qg as oqg


@pytest.fixture(scope="module")
def built():
    if not cpu.LIB.exists():
        cpu.build()
    return True


@pytest.mark.parametrize("n,ny", [(32, 32), (40, 24)])
def test_cpu_step_matches_oracle(built, n, ny):
    s = oqg.baroclinic_setup(nx=n, ny=ny, Lx=4e6, Ly=4e6 * ny / n, lateral_viscosity=50.0, bottom_drag=1e-7,
                             wind_amplitude=1e-11, biharmonic_viscosity=1e11)
    q = np.random.default_rng(n).standard_normal((3, ny + 2, n + 2)) * 1e-6
    ref = oqg.ssprk3_step(s, oqg.ssprk3_step(s, q, 1200.0), 1200.0)
    got64 = cpu.CpuStep(s, fp64=True).step(q, 1200.0, 2)
    got32 = cpu.CpuStep(s, fp64=False).step(q, 1200.0, 2)
    rel = lambda a: float(np.linalg.norm(a - ref) / np.linalg.norm(ref))  # noqa: E731
    assert rel(got64) < 1e-13 and rel(got32) < 1e-6


def test_cpu_step_rejects_extensions(built):
    with pytest.raises(ValueError):
        cpu.CpuStep(oqg.baroclinic_setup(nx=16, ny=16, advection="weno"))
