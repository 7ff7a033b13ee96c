"""Writes tests/golden/longrun_swm128_jet.json: the FP64 oracle's long-run
statistics of the 2-layer shallow-water baroclinic jet (the 512^2 benchmark
jet's physics at 128 x 128 cells): channel of side 2048 km, periodic in x
with walls in y, f-plane f0 = 1e-4 /s, H = 500/3500 m, g' = 9.81/0.02 m/s^2,
balanced 0.5 m/s upper-layer jet of width 200 km, seeded white interface
noise of 1 m (seed 11), energy-enstrophy-conserving vorticity flux, no
viscosity, drag or wind; SSP-RK3 at CFL 0.5 for 30 model days, sampled every
500 steps: the kinetic energy of every sample (the thickness-weighted
0.5 |u|^2 at T points, summed over layers and averaged over the interior)
and the time mean of the samples' zonal KE spectra
(tests/test_gpu_longrun.py compares both time means).

Run from the repository root: ``python tests/golden/make_longrun_swm128_jet.py``
(about 20 minutes on one core).
"""

from __future__ import annotations

import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import swm as oswm  # noqa: E402
from somax_amd import testcases as tc  # noqa: E402

N, DAYS, EVERY, SEED = 128, 30, 500, 11
ARGS = dict(nx=N, ny=N, Lx=tc.JET_L, Ly=tc.JET_L, f0=tc.JET_F0, beta=0.0, H=tc.JET_H, g_prime=tc.JET_GP,
            bc="channel", vorticity_flux="energy_enstrophy")


def stats(h, u, v):
    """(kinetic energy, zonal KE spectrum) of one state: u and v averaged to
    T points, KE = sum over layers of the interior mean of h 0.5 |u|^2; the
    spectrum is 0.5 (|u_m|^2 + |v_m|^2) of the zonal Fourier coefficients of
    the interior rows, averaged over rows, layers weighted by H / sum(H)."""
    h, u, v = (np.asarray(a, dtype=np.float64) for a in (h, u, v))
    uT = 0.5 * (u[:, 1:-1, 1:-1] + u[:, 1:-1, :-2])
    vT = 0.5 * (v[:, 1:-1, 1:-1] + v[:, :-2, 1:-1])
    ke = float(np.sum(np.mean(h[:, 1:-1, 1:-1] * 0.5 * (uT**2 + vT**2), axis=(-2, -1))))
    n = uT.shape[-1]
    uh = np.fft.rfft(uT, axis=-1) / n
    vh = np.fft.rfft(vT, axis=-1) / n
    w = np.asarray(tc.JET_H, dtype=np.float64) / np.sum(tc.JET_H)
    spec = np.einsum("l,lm->m", w, 0.5 * np.mean(np.abs(uh) ** 2 + np.abs(vh) ** 2, axis=1))
    return ke, spec


def main() -> int:
    s = oswm.setup(**ARGS)
    h, u, v = tc.baroclinic_jet_fields(N, seed=SEED)
    # the GPU starts from the float32 fields
    h, u, v = (a.astype(np.float32).astype(np.float64) for a in (h, u, v))
    dt = tc.cfl_dt(tc.JET_H, tc.JET_GP, s.dx, s.dy)
    nsamples = int(DAYS * 86400.0 / dt) // EVERY
    ke, spec = [], []
    t0 = time.time()
    for k in range(nsamples):
        for _ in range(EVERY):
            h, u, v = oswm.ssprk3_step(s, h, u, v, dt)
        e, sp = stats(h, u, v)
        ke.append(e)
        spec.append(sp)
        if (k + 1) % 10 == 0:
            print(f"step {(k + 1) * EVERY}: KE {e:.6e} ({time.time() - t0:.0f} s)", flush=True)
    out = {"workload": "2-layer SWM baroclinic jet, 128x128 channel, energy-enstrophy flux, SSP-RK3 CFL 0.5",
           "n": N, "days": DAYS, "every": EVERY, "nsamples": nsamples, "dt": dt, "seed": SEED, "args": ARGS,
           "kinetic_energy": ke, "spectrum_mean": np.mean(np.asarray(spec), axis=0).tolist()}
    path = Path(__file__).resolve().parent / "longrun_swm128_jet.json"
    path.write_text(json.dumps(out))
    print(f"wrote {path} ({path.stat().st_size} bytes)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
