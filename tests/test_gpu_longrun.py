"""Long-run statistics of the mixed-precision HIP step against the FP64
oracle: the time-mean kinetic energy and the time-mean isotropic kinetic-
energy spectrum must agree within 1% (the flow is chaotic, so fields are
compared through their statistics).

3-layer QG double gyre at 64^2 (the 128^2 benchmark configuration scaled to
the grid: biharmonic viscosity by (128/n)^4), seeded k^-3 initial PV, 6000
SSP-RK3 steps of 1200 s (about 83 days), sampled every 30 steps.  The
spectrum is the layer-thickness-weighted 1/2 |k|^2 |psi_hat|^2 of the
orthonormal DST-I of the interior streamfunction, binned by |k| in steps of
the fundamental wavenumber; bins holding under 1e-4 of the energy are left
out of the comparison.  The 128^2 WENO-Z double gyre and the shallow-water
baroclinic jet (30 model days) are compared with golden statistics of the
FP64 oracle under tests/golden/.
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.fft

from oracle import qg as oqg

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


def _spectrum(psi, H, dx, dy):
    """Isotropic KE spectrum of psi (nl, Ny, Nx) with a ghost ring."""
    inner = np.asarray(psi, dtype=np.float64)[:, 1:-1, 1:-1]
    nl, ny, nx = inner.shape
    hat = scipy.fft.dstn(inner, type=1, axes=(-2, -1), norm="ortho")
    ky = np.pi * np.arange(1, ny + 1) / ((ny + 1) * dy)
    kx = np.pi * np.arange(1, nx + 1) / ((nx + 1) * dx)
    k2 = ky[:, None] ** 2 + kx[None, :] ** 2
    w = np.asarray(H, dtype=np.float64) / np.sum(H)
    e = 0.5 * np.einsum("l,lij->ij", w, hat**2) * k2
    dk = min(kx[0], ky[0])
    bins = np.minimum((np.sqrt(k2) / dk).astype(int), min(nx, ny))
    return np.bincount(bins.ravel(), weights=e.ravel(), minlength=min(nx, ny) + 1)


def test_long_run_statistics_match_oracle(gpu):
    import torch

    from somax_amd import BaroclinicQG, BaroclinicQGState
    from somax_amd import testcases as tc

    n, nsteps, every, dt = 64, 6000, 30, 1200.0
    H = (400.0, 1100.0, 2600.0)
    args = dict(nx=n, ny=n, Lx=tc.BASIN, Ly=tc.BASIN, f0=tc.F0, beta=tc.BETA, H=H, g_prime=(9.81, 0.025, 0.0125),
                bottom_drag=tc.DRAG, wind_amplitude=tc.wind_amplitude(),
                biharmonic_viscosity=tc.NU4_128 * (128.0 / n) ** 4)
    model = BaroclinicQG.create(n_layers=3, precision="mixed", **args)
    setup = oqg.baroclinic_setup(**args)
    q = tc.random_pv(model, 17).astype(np.float32)
    qd = torch.from_numpy(q).cuda()
    ref = q.astype(np.float64)
    dx = dy = tc.BASIN / n
    ke_g, ke_o, sp_g, sp_o = [], [], [], []
    _log(f"[longrun] QG {n}^2 mixed, {nsteps} steps of {dt:.0f} s, sample every {every}")
    for k in range(nsteps // every):
        qd = model.step(BaroclinicQGState(q=qd), dt, nsteps=every).q
        for _ in range(every):
            ref = oqg.ssprk3_step(setup, ref, dt)
        dg = model.diagnose(BaroclinicQGState(q=qd))
        do = oqg.diagnose(setup, ref)
        ke_g.append(float(dg.total_kinetic_energy))
        ke_o.append(float(do["total_kinetic_energy"]))
        sp_g.append(_spectrum(dg.psi.cpu().numpy(), H, dx, dy))
        sp_o.append(_spectrum(do["psi"], H, dx, dy))
        if (k + 1) % 40 == 0:
            _log(f"  step {(k + 1) * every}: KE gpu {ke_g[-1]:.6e} oracle {ke_o[-1]:.6e}")
    assert np.all(np.isfinite(ke_g))
    mean_g, mean_o = np.mean(ke_g), np.mean(ke_o)
    r_ke = abs(mean_g - mean_o) / mean_o
    sg, so = np.mean(sp_g, axis=0), np.mean(sp_o, axis=0)
    keep = so > 1e-4 * so.sum()
    r_sp = float(np.max(np.abs(sg[keep] - so[keep]) / so[keep]))
    final = float(np.linalg.norm(qd.cpu().numpy() - ref) / np.linalg.norm(ref))
    _log(f"  time-mean KE rel diff {r_ke:.3e}; spectrum max rel diff over {int(keep.sum())} bins {r_sp:.3e}; "
         f"final-state rel-L2 {final:.3e}")
    assert r_ke < 0.01 and r_sp < 0.01


def test_long_run_statistics_qg128_weno_match_oracle(gpu):
    """The multilayer double gyre at 128^2 with WENO-Z PV flux, 10,000 mixed
    SSP-RK3 steps of 1200 s (139 days) sampled every 50 steps, against the
    FP64 oracle's statistics of the same run held in
    tests/golden/longrun_qg128_weno.json (written by
    tests/golden/make_longrun_qg128_weno.py): time-mean total kinetic energy
    and time-mean isotropic KE spectrum within 1%."""
    import json
    from pathlib import Path

    import torch

    from somax_amd import BaroclinicQG, BaroclinicQGState
    from somax_amd import testcases as tc

    fx = json.loads((Path(__file__).resolve().parent / "golden" / "longrun_qg128_weno.json").read_text())
    n, nsteps, every, dt = fx["n"], fx["nsteps"], fx["every"], fx["dt"]
    args = dict(fx["args"])
    H = tuple(args["H"])
    model = BaroclinicQG.create(n_layers=3, precision="mixed", advection=fx["advection"], **args)
    psi0 = tc.random_streamfunction(n, n, args["Lx"], args["Ly"], 3, fx["seed"], 0.05)
    q = tc.pv_from_streamfunction(psi0, model.grid.dx, model.grid.dy, model._lambdas, model._Cl2m, model._Cm2l)
    qd = torch.from_numpy(q.astype(np.float32)).cuda()
    dx, dy = model.grid.dx, model.grid.dy
    ke, sp = [], []
    _log(f"[longrun] QG {n}^2 WENO-Z mixed, {nsteps} steps of {dt:.0f} s, sample every {every}")
    for k in range(nsteps // every):
        model._step_inplace(qd, dt, every)
        d = model.diagnose(BaroclinicQGState(q=qd))
        ke.append(float(d.total_kinetic_energy))
        sp.append(_spectrum(d.psi.cpu().numpy(), H, dx, dy))
        if (k + 1) % 40 == 0:
            _log(f"  step {(k + 1) * every}: KE gpu {ke[-1]:.6e} oracle {fx['total_kinetic_energy'][k]:.6e}")
    assert np.all(np.isfinite(ke))
    mean_g, mean_o = np.mean(ke), np.mean(fx["total_kinetic_energy"])
    r_ke = abs(mean_g - mean_o) / mean_o
    sg, so = np.mean(sp, axis=0), np.asarray(fx["spectrum_mean"])
    keep = so > 1e-4 * so.sum()
    r_sp = float(np.max(np.abs(sg[keep] - so[keep]) / so[keep]))
    _log(f"  time-mean KE rel diff {r_ke:.3e}; spectrum max rel diff over {int(keep.sum())} bins {r_sp:.3e}")
    assert r_ke < 0.01 and r_sp < 0.01


def test_long_run_statistics_swm128_jet_match_oracle(gpu):
    """The 2-layer shallow-water baroclinic jet (the 512^2 benchmark jet's
    physics at 128^2, energy-enstrophy-conserving flux) for 30 model days of
    SSP-RK3 at CFL 0.5, sampled every 500 steps, against the FP64 oracle's
    statistics of the same run held in tests/golden/longrun_swm128_jet.json
    (written by tests/golden/make_longrun_swm128_jet.py, whose ``stats`` both
    sides use): time-mean kinetic energy and time-mean zonal KE spectrum
    within 1% (bins holding under 1e-4 of the energy left out)."""
    import importlib.util
    import json
    from pathlib import Path

    import torch

    from somax_amd import MultilayerShallowWater2D, MultilayerSW2DState
    from somax_amd import testcases as tc

    golden = Path(__file__).resolve().parent / "golden"
    spec = importlib.util.spec_from_file_location("make_longrun_swm128_jet", golden / "make_longrun_swm128_jet.py")
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    fx = json.loads((golden / "longrun_swm128_jet.json").read_text())
    n, every, nsamples, dt = fx["n"], fx["every"], fx["nsamples"], fx["dt"]
    args = {k: (tuple(v) if isinstance(v, list) else v) for k, v in fx["args"].items()}
    model = MultilayerShallowWater2D.create(n_layers=2, **args)
    h, u, v = (torch.from_numpy(a.astype(np.float32)).cuda() for a in tc.baroclinic_jet_fields(n, seed=fx["seed"]))
    st = MultilayerSW2DState(h=h, u=u, v=v)
    ke, sp = [], []
    _log(f"[longrun] SWM jet {n}^2, {nsamples * every} steps of {dt:.1f} s, sample every {every}")
    for k in range(nsamples):
        st = model.step(st, dt, every)
        e, s = gen.stats(st.h.cpu().numpy(), st.u.cpu().numpy(), st.v.cpu().numpy())
        ke.append(e)
        sp.append(s)
        if (k + 1) % 9 == 0:
            _log(f"  step {(k + 1) * every}: KE gpu {e:.6e} oracle {fx['kinetic_energy'][k]:.6e}")
    assert np.all(np.isfinite(ke))
    mean_g, mean_o = np.mean(ke), np.mean(fx["kinetic_energy"])
    r_ke = abs(mean_g - mean_o) / mean_o
    sg, so = np.mean(sp, axis=0), np.asarray(fx["spectrum_mean"])
    keep = so > 1e-4 * so.sum()
    r_sp = float(np.max(np.abs(sg[keep] - so[keep]) / so[keep]))
    _log(f"  time-mean KE rel diff {r_ke:.3e}; spectrum max rel diff over {int(keep.sum())} bins {r_sp:.3e}")
    assert r_ke < 0.01 and r_sp < 0.01
