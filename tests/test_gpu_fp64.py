"""GPU parity of the FP64 QG form (``precision="fp64"``) against the FP64 oracle.

The FP64 form runs every term in double precision on float64 state, as the
reference runs its CPU path with x64 enabled (the 64^2 barotropic double
gyre, 1000 SSP-RK3 steps of 4000 s).  GPU and oracle receive the same FP64
inputs, so the remaining differences are rounding: the oracle's sine
transforms are FFTs (scipy.fft.dstn), the GPU's are dense FP64 GEMMs, and
sums run in other orders.  Tolerances (relative L2 over the whole array):

* PV inversion and one step: 1e-12;
* RHS (vector_field): 1e-10 (Arakawa / WENO differences of products);
* 100 steps: 1e-10; the 1000-step barotropic run: 1e-9.

Each bar is at least four orders of magnitude below the north-star bars that
the FP32 / BF16 forms meet (1e-6 after 1 step, 1e-4 after 100).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import qg as oqg

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _log(msg):
    print(msg, flush=True)


def _pair(n, nl=3, advection="arakawa", ny=None, **kw):
    """An FP64 GPU model and the oracle setup with identical parameters."""
    from somax_amd import BaroclinicQG
    from somax_amd import testcases as tc

    ny = n if ny is None else ny
    H = (400.0, 1100.0, 2600.0)[:nl]
    gp = (9.81, 0.025, 0.0125)[:nl]
    args = dict(nx=n, ny=ny, Lx=tc.BASIN, Ly=tc.BASIN * ny / n, f0=tc.F0, beta=tc.BETA, H=H, g_prime=gp,
                bottom_drag=tc.DRAG, wind_amplitude=tc.wind_amplitude(),
                biharmonic_viscosity=kw.pop("biharmonic_viscosity", tc.NU4_128 * (128.0 / n) ** 4),
                lateral_viscosity=kw.pop("lateral_viscosity", 0.0))
    model = BaroclinicQG.create(n_layers=nl, precision="fp64", advection=advection, **args, **kw)
    setup = oqg.baroclinic_setup(advection=advection, **args, **kw)
    return model, setup


def _state(model, seed, ghost_noise=False):
    import torch

    from somax_amd import testcases as tc

    q = tc.random_pv(model, seed)
    if ghost_noise:
        rng = np.random.default_rng(seed + 1000)
        g = rng.standard_normal(q.shape) * np.abs(q).max()
        g[:, 1:-1, 1:-1] = 0.0
        q = q + g
    return q, torch.from_numpy(q).cuda()


@pytest.mark.parametrize("n,ny,nl,mass", [(64, 64, 3, False), (64, 64, 3, True), (96, 64, 2, False),
                                          (64, 128, 1, False), (300, 260, 2, True)])
def test_fp64_inversion(gpu, n, ny, nl, mass):
    """Non-square grids catch a transposed sine matrix or eigenvalue table;
    transforms up to 256 x 256 run on FP64 MFMA, the 302 x 262 case on the
    LDS-tiled FMA kernel."""
    _log(f"[fp64 inversion] nx={n} ny={ny} nl={nl} mass={mass}")
    model, setup = _pair(n, nl=nl, ny=ny, mass_correction=mass)
    q, qd = _state(model, seed=3)
    psi = model._invert_pv(qd).cpu().numpy()
    assert psi.dtype == np.float64
    r = rel(psi, oqg.invert_pv(setup, q))
    _log(f"  rel-L2 psi = {r:.3e}")
    assert r < 1e-12


@pytest.mark.parametrize("nl,advection,mass", [(5, "weno", True), (8, "arakawa", True), (8, "weno", False)])
def test_fp64_deep_stratification(gpu, nl, advection, mass):
    """5- and 8-layer stacks (the layer mixes and the stage kernel's layer
    loops at their longest): inversion and one SSP-RK3 step against the
    oracle, to FP64 round-off."""
    from somax_amd import BaroclinicQG
    from somax_amd import testcases as tc

    _log(f"[fp64 layers] nl={nl} advection={advection} mass={mass}")
    w = np.linspace(300.0, 700.0, nl)
    H = tuple(float(x) for x in w * (4000.0 / w.sum()))
    gp = (9.81,) + tuple(0.03 / (1.0 + k) for k in range(nl - 1))
    args = dict(nx=64, ny=48, Lx=tc.BASIN, Ly=tc.BASIN * 0.75, f0=tc.F0, beta=tc.BETA, H=H, g_prime=gp,
                bottom_drag=tc.DRAG, wind_amplitude=tc.wind_amplitude(), biharmonic_viscosity=tc.NU4_128 * 16.0,
                lateral_viscosity=0.0)
    model = BaroclinicQG.create(n_layers=nl, precision="fp64", advection=advection, mass_correction=mass, **args)
    setup = oqg.baroclinic_setup(advection=advection, mass_correction=mass, **args)
    q, qd = _state(model, seed=11 + nl)
    r = rel(model._invert_pv(qd).cpu().numpy(), oqg.invert_pv(setup, q))
    q1 = qd.clone()
    model._step_inplace(q1, 4000.0, 1)
    r1 = rel(q1.cpu().numpy(), oqg.ssprk3_step(setup, q, 4000.0))
    _log(f"  rel-L2 psi {r:.3e}, one step q {r1:.3e}")
    assert r < 1e-12 and r1 < 1e-12


@pytest.mark.parametrize("advection,mass,nu", [("arakawa", False, 0.0), ("arakawa", False, 50.0),
                                               ("weno", False, 0.0), ("weno", True, 0.0), ("arakawa", True, 50.0)])
def test_fp64_rhs(gpu, advection, mass, nu):
    """build_terms' RHS (BC applied first) and vector_field on a state whose
    ghost ring is not zero (BC not applied), against the oracle's RHS."""
    from somax_amd import BaroclinicQGState

    _log(f"[fp64 rhs] advection={advection} mass={mass} nu={nu}")
    model, setup = _pair(64, advection=advection, mass_correction=mass, lateral_viscosity=nu)
    q, qd = _state(model, seed=5, ghost_noise=True)
    dq = model.build_terms()(0.0, BaroclinicQGState(q=qd)).q.cpu().numpy()
    ref = oqg.rhs(setup, oqg.apply_bc(q))
    r = rel(dq, ref)
    dq_raw = model.vector_field(0.0, BaroclinicQGState(q=qd)).q.cpu().numpy()
    r_raw = rel(dq_raw, oqg.rhs(setup, q))
    _log(f"  rel-L2 rhs {r:.3e}, without BC {r_raw:.3e}")
    assert r < 1e-10 and r_raw < 1e-10


@pytest.mark.parametrize("advection,mass", [("arakawa", False), ("weno", True)])
def test_fp64_ssprk3(gpu, advection, mass):
    """One and 100 SSP-RK3 steps of 1200 s at 64^2, three layers."""
    import torch

    _log(f"[fp64 ssprk3] advection={advection} mass={mass}")
    model, setup = _pair(64, advection=advection, mass_correction=mass)
    q, qd = _state(model, seed=7)
    q1 = qd.clone()
    model._step_inplace(q1, 1200.0, 1)
    ref1 = oqg.ssprk3_step(setup, q, 1200.0)
    r1 = rel(q1.cpu().numpy(), ref1)
    q100 = qd.clone()
    model._step_inplace(q100, 1200.0, 100)
    torch.cuda.synchronize()
    ref100 = oqg.run(setup, q, 1200.0, 100)
    r100 = rel(q100.cpu().numpy(), ref100)
    rp = rel(model._invert_pv(q100).cpu().numpy(), oqg.invert_pv(setup, ref100))
    _log(f"  1 step {r1:.3e}; 100 steps q {r100:.3e} psi {rp:.3e}")
    assert r1 < 1e-12 and r100 < 1e-10 and rp < 1e-10


def test_fp64_barotropic_thousand_steps(gpu):
    """The reference's CPU configuration in FP64: the 1-layer double gyre at
    64^2 (DST-I Poisson inversion, psi ghost ring kept, biharmonic
    2e11 m^4/s, drag 3.6e-8 /s, wind 0.08 N/m^2 over 4000 m), seeded k^-3
    initial PV, 1000 SSP-RK3 steps of 4000 s."""
    import torch

    from somax_amd import testcases as tc

    _log("[fp64 barotropic] 1000 steps n=64")
    model = tc.config_barotropic(64, precision="fp64")
    setup = oqg.barotropic_setup(nx=64, ny=64, Lx=tc.BASIN, Ly=tc.BASIN, f0=tc.F0, beta=tc.BETA,
                                 bottom_drag=tc.DRAG, wind_amplitude=tc.wind_amplitude() / 4000.0,
                                 biharmonic_viscosity=tc.NU4_128)
    q = tc.random_pv(model, seed=8)[0]
    qd = torch.from_numpy(q).cuda()[None]
    model._step_inplace(qd, 4000.0, 1000)
    torch.cuda.synchronize()
    ref = oqg.run(setup, q[None], 4000.0, 1000, fast=True)
    r = rel(qd.cpu().numpy(), ref)
    rp = rel(model._invert_pv(qd[0]).cpu().numpy(), oqg.invert_pv(setup, ref)[0])
    _log(f"  1000 steps: rel-L2 q {r:.3e} psi {rp:.3e}")
    assert r < 1e-9 and rp < 1e-9


def test_fp64_tsit5_and_integrate(gpu):
    """Tsit5 (integrate's default solver): 1 and 5 steps against the oracle's
    explicit Tsit5, and integrate() over 3 steps."""
    import torch

    from somax_amd import BaroclinicQGState

    _log("[fp64 tsit5] n=64 weno")
    model, setup = _pair(64, advection="weno")
    q, qd = _state(model, seed=9)
    q5 = qd.clone()
    model._step_tsit5_inplace(q5, 1200.0, 5)
    torch.cuda.synchronize()
    ref = q.copy()
    for k in range(5):
        ref = oqg.tsit5_step(setup, ref, 1200.0)
        if k == 0:
            ref1 = ref.copy()
    r5 = rel(q5.cpu().numpy(), ref)
    q1 = qd.clone()
    model._step_tsit5_inplace(q1, 1200.0, 1)
    r1 = rel(q1.cpu().numpy(), ref1)
    sol = model.integrate(BaroclinicQGState(q=qd), 0.0, 3600.0, 1200.0)
    ref3 = oqg.apply_bc(q)
    for _ in range(3):
        ref3 = oqg.tsit5_step(setup, ref3, 1200.0)
    r3 = rel(sol.ys.q[-1].cpu().numpy(), ref3)
    _log(f"  1 step {r1:.3e}, 5 steps {r5:.3e}, integrate 3 steps {r3:.3e}")
    assert r1 < 1e-12 and r5 < 1e-11 and r3 < 1e-11


@pytest.mark.parametrize("nl", [1, 3])
def test_fp64_diagnose(gpu, nl):
    from somax_amd import BaroclinicQGState

    _log(f"[fp64 diagnose] nl={nl}")
    model, setup = _pair(64, nl=nl)
    q, qd = _state(model, seed=13)
    d = model.diagnose(BaroclinicQGState(q=qd))
    ref = oqg.diagnose(setup, q)
    for name, got, want in [("psi", d.psi, ref["psi"]), ("u", d.u, ref["u"]), ("v", d.v, ref["v"]),
                            ("zeta", d.relative_vorticity, ref["relative_vorticity"])]:
        r = rel(got.cpu().numpy(), want)
        _log(f"  {name} {r:.3e}")
        assert r < 1e-12, name
    ke = d.kinetic_energy.cpu().numpy()
    ens = d.enstrophy.cpu().numpy()
    assert np.allclose(ke, ref["kinetic_energy"], rtol=1e-12, atol=0)
    assert np.allclose(ens, ref["enstrophy"], rtol=1e-12, atol=0)


def test_fp64_plans_refuse_the_other_form(gpu):
    """An FP64 plan refuses the FP32 calls and an FP32 plan the FP64 calls;
    ensembles, the masked solve's plans and the reparameterized model stay
    FP32 / BF16."""
    import ctypes

    import torch

    from somax_amd import QGEnsemble, ReparameterizedQG, _lib
    from somax_amd import testcases as tc

    model64 = tc.config_barotropic(64, precision="fp64")
    model32 = tc.config_barotropic(64, precision="fp32")
    lib = _lib.load()
    q64 = torch.zeros((1, 66, 66), dtype=torch.float64, device="cuda")
    q32 = torch.zeros((1, 66, 66), dtype=torch.float32, device="cuda")
    p64, p32 = model64._ensure_plan(), model32._ensure_plan()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.somaxhip_qg_invert(p64, q32.data_ptr(), q32.data_ptr(), 1, model64._ws.data_ptr(), st)
    assert rc != 0 and "FP64" in _lib.last_error()
    rc = lib.somaxhip_qg_invert_f64(p32, q64.data_ptr(), q64.data_ptr(), 1, model32._ws.data_ptr(), st)
    assert rc != 0 and "SOMAXHIP_PREC_FP64" in _lib.last_error()
    with pytest.raises(TypeError, match="float64"):
        model64._invert_pv(q32)
    with pytest.raises(_lib.SomaxHipError, match="FP64"):
        QGEnsemble([model64, model64])
    with pytest.raises(ValueError, match="fp64"):
        ReparameterizedQG.create(nx=32, ny=32, precision="fp64")


@pytest.mark.parametrize("advection", ["arakawa", "weno"])
def test_fp64_barotropic_step_and_psi_out(gpu, advection):
    """1 and 50 SSP-RK3 steps of the barotropic 64^2 model against the
    oracle, and psi_out = the streamfunction of the last stage's input state
    (q2 = 3/4 q0 + 1/4 (q1 + dt f(q1)) of the last step)."""
    import ctypes

    import torch

    from somax_amd import _lib
    from somax_amd import testcases as tc

    _log(f"[fp64 barotropic step] advection={advection}")
    model = tc.config_barotropic(64, precision="fp64", advection=advection)
    setup = oqg.barotropic_setup(nx=64, ny=64, Lx=tc.BASIN, Ly=tc.BASIN, f0=tc.F0, beta=tc.BETA,
                                 bottom_drag=tc.DRAG, wind_amplitude=tc.wind_amplitude() / 4000.0,
                                 biharmonic_viscosity=tc.NU4_128, advection=advection)
    q0 = tc.random_pv(model, seed=21)
    qd = torch.from_numpy(q0).cuda()
    psi = torch.zeros_like(qd)
    plan = model._ensure_plan()
    prm = model._c_params()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(_lib.load().somaxhip_qg_step_f64(plan, qd.data_ptr(), psi.data_ptr(), 1, 4000.0, ctypes.byref(prm),
                                                model._ws.data_ptr(), st), "somaxhip_qg_step_f64")
    torch.cuda.synchronize()
    dt = 4000.0
    q1 = q0 + dt * oqg.rhs(setup, oqg.apply_bc(q0))
    q2 = 0.75 * q0 + 0.25 * (q1 + dt * oqg.rhs(setup, oqg.apply_bc(q1)))
    r_psi = rel(psi.cpu().numpy(), oqg.invert_pv(setup, oqg.apply_bc(q2)))
    r1 = rel(qd.cpu().numpy(), oqg.ssprk3_step(setup, q0, dt))
    q50 = torch.from_numpy(q0).cuda()
    model._step_inplace(q50, dt, 50)
    torch.cuda.synchronize()
    r50 = rel(q50.cpu().numpy(), oqg.run(setup, q0, dt, 50))
    _log(f"  1 step {r1:.3e}, psi_out {r_psi:.3e}, 50 steps {r50:.3e}")
    assert r1 < 1e-12 and r_psi < 1e-12 and r50 < 1e-11
