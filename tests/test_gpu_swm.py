"""GPU parity of the multilayer shallow-water path against the FP64 oracle.

The GPU receives float32 fields and the oracle the same values widened to
FP64.  Tolerances (relative L2 over the whole array, per field):

* apply_boundary_conditions: exact (copies and zeros).
* RHS: dh 1e-5, du and dv 1e-5 (FP32 products of O(1) factors; the pressure
  gradient is formed from exact FP32 differences of h).
* One SSP-RK3 step: h, u, v 1e-6; 100 steps: 1e-4.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import swm as oswm

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _log(msg):
    print(msg, flush=True)


def _jet(n, bc="channel", seed=0, **kw):
    """The baroclinic jet at n^2 (viscosity and drag as given) and its oracle setup."""
    from somax_amd import testcases as tc

    model, (h, u, v) = tc.baroclinic_jet_swm(n, seed=seed, **kw)
    if bc != "channel":
        model = type(model).create(nx=n, ny=n, Lx=tc.JET_L, Ly=tc.JET_L, f0=tc.JET_F0, beta=0.0, n_layers=2,
                                   H=(500.0, 3500.0), g_prime=(9.81, 0.02), bc=bc, **kw)
    s = oswm.setup(nx=n, ny=n, Lx=tc.JET_L, Ly=tc.JET_L, f0=tc.JET_F0, beta=0.0, H=(500.0, 3500.0),
                   g_prime=(9.81, 0.02), bc=bc, **kw)
    rng = np.random.default_rng(seed + 7)
    v = 0.05 * rng.standard_normal(v.shape)  # cross-jet flow so every term is active
    f32 = [np.asarray(a, dtype=np.float32) for a in (h, u, v)]
    return model, s, f32


def _state(f32):
    import torch

    from somax_amd import MultilayerSW2DState

    h, u, v = (torch.from_numpy(a.copy()).cuda() for a in f32)
    return MultilayerSW2DState(h=h, u=u, v=v)


KW = dict(lateral_viscosity=50.0, bottom_drag=1e-7, wind_amplitude=2e-5)


@pytest.mark.parametrize("bc", ["periodic", "wall", "channel"])
def test_apply_bc_matches_oracle(gpu, bc):
    _log(f"[swm] apply_bc {bc}")
    model, s, f32 = _jet(64, bc, **KW)
    st = model.apply_boundary_conditions(_state(f32))
    ref = oswm.apply_bc(s, *[a.astype(np.float64) for a in f32])
    for name, a, b in zip("huv", (st.h, st.u, st.v), ref):
        assert np.array_equal(a.cpu().numpy().astype(np.float64), b), name


@pytest.mark.parametrize("bc", ["periodic", "wall", "channel"])
@pytest.mark.parametrize("apply_bc", [False, True])
def test_rhs_matches_oracle(gpu, bc, apply_bc):
    _log(f"[swm] rhs {bc} apply_bc={apply_bc}")
    model, s, f32 = _jet(128, bc, **KW)
    st = _state(f32)
    d = model.build_terms()(0.0, st) if apply_bc else model.vector_field(0.0, st)
    x = [a.astype(np.float64) for a in f32]
    ref = oswm.rhs(s, *(oswm.apply_bc(s, *x) if apply_bc else x))
    rs = [rel(a.cpu().numpy(), b) for a, b in zip((d.h, d.u, d.v), ref)]
    _log(f"  rel-L2 dh {rs[0]:.3e} du {rs[1]:.3e} dv {rs[2]:.3e}")
    assert max(rs) < 1e-5


@pytest.mark.parametrize("n,bc", [(128, "channel"), (128, "wall"), (128, "periodic"), (512, "channel")])
def test_one_step_matches_oracle(gpu, n, bc):
    from somax_amd import testcases as tc

    _log(f"[swm] one step n={n} {bc}")
    model, s, f32 = _jet(n, bc, **KW)
    dt = tc.jet_dt(model)
    out = model.step(_state(f32), dt, 1)
    ref = oswm.ssprk3_step(s, *[a.astype(np.float64) for a in f32], dt)
    rs = [rel(a.cpu().numpy(), b) for a, b in zip((out.h, out.u, out.v), ref)]
    _log(f"  dt={dt:.2f}s rel-L2 h {rs[0]:.3e} u {rs[1]:.3e} v {rs[2]:.3e}")
    assert max(rs) < 1e-6


@pytest.mark.parametrize("bc", ["channel", "wall", "periodic"])
def test_tsit5_step_matches_oracle(gpu, bc):
    """Tsit5 (integrate's default solver, core/model.py:75): one step to 1e-6."""
    from somax_amd import testcases as tc

    _log(f"[swm] Tsit5 one step n=128 {bc}")
    model, s, f32 = _jet(128, bc, **KW)
    dt = tc.jet_dt(model)
    out = model.step(_state(f32), dt, 1, solver="tsit5")
    ref = oswm.tsit5_step(s, *[a.astype(np.float64) for a in f32], dt)
    rs = [rel(a.cpu().numpy(), b) for a, b in zip((out.h, out.u, out.v), ref)]
    _log(f"  rel-L2 h {rs[0]:.3e} u {rs[1]:.3e} v {rs[2]:.3e}")
    assert max(rs) < 1e-6


def test_hundred_steps_match_oracle(gpu):
    from somax_amd import testcases as tc

    _log("[swm] 100 steps n=128 channel")
    model, s, f32 = _jet(128, "channel", **KW)
    dt = tc.jet_dt(model)
    out = model.step(_state(f32), dt, 100)
    ref = oswm.run(s, *[a.astype(np.float64) for a in f32], dt, 100)
    rs = [rel(a.cpu().numpy(), b) for a, b in zip((out.h, out.u, out.v), ref)]
    _log(f"  rel-L2 h {rs[0]:.3e} u {rs[1]:.3e} v {rs[2]:.3e}")
    assert max(rs) < 1e-4


def test_rest_state_and_forcing_locality(gpu):
    """tests/models/test_swm_multilayer.py:94-127 on the HIP path."""
    import torch

    from somax_amd import MultilayerShallowWater2D, MultilayerSW2DState

    m = MultilayerShallowWater2D.create(nx=16, ny=16, n_layers=2, H=(500.0, 500.0), g_prime=(9.81, 0.02))
    shape = (2, m.grid.Ny, m.grid.Nx)
    h = torch.ones(shape, device="cuda") * torch.tensor([500.0, 500.0], device="cuda")[:, None, None]
    rest = m.apply_boundary_conditions(MultilayerSW2DState(h=h, u=torch.zeros(shape, device="cuda"),
                                                           v=torch.zeros(shape, device="cuda")))
    d = m.vector_field(0.0, rest)
    assert max(float(d.h.abs().max()), float(d.u.abs().max()), float(d.v.abs().max())) < 1e-10
    mw = MultilayerShallowWater2D.create(nx=16, ny=16, n_layers=2, H=(500.0, 500.0), g_prime=(9.81, 0.02),
                                         wind_amplitude=1e-5)
    d = mw.vector_field(0.0, rest)
    assert float(d.u[0].abs().max()) > 0 and float(d.u[-1].abs().max()) < 1e-15
    md = MultilayerShallowWater2D.create(nx=16, ny=16, n_layers=2, H=(500.0, 500.0), g_prime=(9.81, 0.02),
                                         bottom_drag=1e-4)
    uni = MultilayerSW2DState(h=rest.h, u=torch.ones(shape, device="cuda"), v=rest.v)
    d = md.vector_field(0.0, md.apply_boundary_conditions(uni))
    assert torch.isfinite(d.u).all()
    assert float((d.u[-1, 1:-1, 1:-1] + 1e-4).abs().max()) < 1e-9  # drag on the bottom layer only


@pytest.mark.parametrize("bc", ["periodic", "channel", "wall"])
def test_mass_conserved_over_steps(gpu, bc):
    from somax_amd import testcases as tc

    _log(f"[swm] mass over 50 steps, {bc}")
    model, s, f32 = _jet(128, bc, **KW)
    st = model.apply_boundary_conditions(_state(f32))
    m0 = st.h[:, 1:-1, 1:-1].double().sum(dim=(-2, -1)).cpu().numpy()
    out = model.step(st, tc.jet_dt(model), 50)
    m1 = out.h[:, 1:-1, 1:-1].double().sum(dim=(-2, -1)).cpu().numpy()
    r = np.abs(m1 - m0) / m0
    _log(f"  relative mass change per layer {r}")
    assert r.max() < 1e-6


def test_reference_instability_case_integrates(gpu):
    """tests/models/test_swm_multilayer.py:254-266."""
    import torch

    from somax_amd import testcases as tc

    model, state0 = tc.baroclinic_instability_swm(nx=16, ny=16)
    sol = model.integrate(state0, t0=0.0, t1=1000.0, dt=10.0)
    assert torch.isfinite(sol.ys.h).all() and sol.ys.h.shape[0] == 1


@pytest.mark.parametrize("bc", ["channel", "wall"])
def test_diagnose_matches_oracle(gpu, bc):
    """diagnose (multilayer.py:225-255): energy and enstrophy per layer to 1e-6
    relative (FP64 sums), PV / vorticity / KE fields to 1e-5 relative L2."""
    _log(f"[swm] diagnose {bc}")
    model, s, f32 = _jet(128, bc, **KW)
    d = model.diagnose(_state(f32))
    ref = oswm.diagnose(s, *[a.astype(np.float64) for a in f32])
    rs = {k: rel(getattr(d, a).cpu().numpy(), ref[k]) for k, a in
          (("energy", "energy"), ("enstrophy", "enstrophy"), ("potential_vorticity", "potential_vorticity"),
           ("relative_vorticity", "relative_vorticity"), ("kinetic_energy_field", "kinetic_energy_field"))}
    _log("  " + " ".join(f"{k} {v:.2e}" for k, v in rs.items()))
    assert rs["energy"] < 1e-6 and rs["enstrophy"] < 1e-6
    assert rs["potential_vorticity"] < 1e-5 and rs["relative_vorticity"] < 1e-5 and rs["kinetic_energy_field"] < 1e-5
    assert np.all(d.energy.cpu().numpy() >= 0)


# ---------------------------------------------------------------------------
# single-layer NonlinearShallowWater2D (nonlinear_2d.py)
# ---------------------------------------------------------------------------
def _nonlinear(n, bc, seed=4):
    from somax_amd import NonlinearShallowWater2D

    kw = dict(nx=n, ny=n, H0=100.0, lateral_viscosity=20.0, bottom_drag=1e-6, wind_amplitude=2e-6, bc=bc)
    model = NonlinearShallowWater2D.create(**kw)
    s = oswm.nonlinear_setup(**kw)
    rng = np.random.default_rng(seed)
    shape = (s.Ny, s.Nx)
    h = 100.0 + 0.5 * rng.standard_normal(shape)
    u = 0.05 * rng.standard_normal(shape)
    v = 0.05 * rng.standard_normal(shape)
    return model, s, [np.asarray(a, dtype=np.float32) for a in (h, u, v)]


def _nl_state(f32):
    import torch

    from somax_amd import NonlinearSW2DState

    h, u, v = (torch.from_numpy(a.copy()).cuda() for a in f32)
    return NonlinearSW2DState(h=h, u=u, v=v)


@pytest.mark.parametrize("bc", ["periodic", "wall"])
def test_nonlinear_rhs_and_step_match_oracle(gpu, bc):
    _log(f"[swm1] single layer n=64 {bc}")
    model, s, f32 = _nonlinear(64, bc)
    x = [a.astype(np.float64)[None] for a in f32]
    d = model.build_terms()(0.0, _nl_state(f32))
    ref = oswm.rhs(s, *oswm.apply_bc(s, *x))
    rs = [rel(a.cpu().numpy(), b[0]) for a, b in zip((d.h, d.u, d.v), ref)]
    _log(f"  rhs rel-L2 dh {rs[0]:.3e} du {rs[1]:.3e} dv {rs[2]:.3e}")
    assert d.h.shape == f32[0].shape and max(rs) < 1e-5
    dt = 20.0
    for solver, one in (("ssprk3", oswm.ssprk3_step), ("tsit5", oswm.tsit5_step)):
        out = model.step(_nl_state(f32), dt, 1, solver=solver)
        ref = one(s, *x, dt)
        rs = [rel(a.cpu().numpy(), b[0]) for a, b in zip((out.h, out.u, out.v), ref)]
        _log(f"  one {solver} step rel-L2 h {rs[0]:.3e} u {rs[1]:.3e} v {rs[2]:.3e}")
        assert max(rs) < 1e-6


@pytest.mark.timeout(300)
def test_nonlinear_1024_step_matches_oracle(gpu):
    """The single-layer model at the size the benchmark times it (1024^2,
    walls): one SSP-RK3 step against the oracle."""
    _log("[swm1] single layer n=1024 wall")
    model, s, f32 = _nonlinear(1024, "wall", seed=9)
    x = [a.astype(np.float64)[None] for a in f32]
    dt = 5.0
    out = model.step(_nl_state(f32), dt, 1)
    ref = oswm.ssprk3_step(s, *x, dt)
    rs = [rel(a.cpu().numpy(), b[0]) for a, b in zip((out.h, out.u, out.v), ref)]
    _log(f"  one step rel-L2 h {rs[0]:.3e} u {rs[1]:.3e} v {rs[2]:.3e}")
    assert max(rs) < 1e-6


def test_nonlinear_reference_behaviours(gpu):
    """tests/models/test_swm_2d.py:137-159."""
    import torch

    from somax_amd import NonlinearShallowWater2D, NonlinearSW2DDiagnostics, NonlinearSW2DState

    model = NonlinearShallowWater2D.create(nx=16, ny=16, H0=100.0, lateral_viscosity=100.0)
    g = model.grid
    x = torch.arange(g.Nx, dtype=torch.float32, device="cuda") * g.dx
    y = torch.arange(g.Ny, dtype=torch.float32, device="cuda") * g.dy
    Y, X = torch.meshgrid(y, x, indexing="ij")
    h0 = 100.0 + 0.1 * torch.exp(-((X - g.Lx / 2) ** 2 + (Y - g.Ly / 2) ** 2) / (2 * (g.Lx / 5) ** 2))
    state0 = NonlinearSW2DState(h=h0, u=torch.zeros_like(h0), v=torch.zeros_like(h0))
    sol = model.integrate(state0, t0=0.0, t1=10.0, dt=0.1)
    assert bool(torch.isfinite(sol.ys.h).all()) and sol.ys.h.shape == (1, g.Ny, g.Nx)
    rest = NonlinearSW2DState(h=torch.full_like(h0, 100.0), u=torch.zeros_like(h0), v=torch.zeros_like(h0))
    d = model.diagnose(rest)
    assert isinstance(d, NonlinearSW2DDiagnostics)
    assert d.potential_vorticity.shape == h0.shape and d.relative_vorticity.shape == h0.shape
    assert d.energy.ndim == 0 and float(d.energy) > 0


# ---------------------------------------------------------------------------
# energy-enstrophy-conserving (triad) vorticity flux (extension)
# ---------------------------------------------------------------------------
EEC = dict(KW, vorticity_flux="energy_enstrophy")


@pytest.mark.parametrize("bc", ["periodic", "wall", "channel"])
@pytest.mark.parametrize("apply_bc", [False, True])
def test_energy_enstrophy_rhs_matches_oracle(gpu, bc, apply_bc):
    _log(f"[swm eec] rhs {bc} apply_bc={apply_bc}")
    model, s, f32 = _jet(128, bc, **EEC)
    st = _state(f32)
    d = model.build_terms()(0.0, st) if apply_bc else model.vector_field(0.0, st)
    x = [a.astype(np.float64) for a in f32]
    ref = oswm.rhs(s, *(oswm.apply_bc(s, *x) if apply_bc else x))
    rs = [rel(a.cpu().numpy(), b) for a, b in zip((d.h, d.u, d.v), ref)]
    _log(f"  rel-L2 dh {rs[0]:.3e} du {rs[1]:.3e} dv {rs[2]:.3e}")
    assert max(rs) < 1e-5
    # the triad form differs from the average form on this flow
    _, s_avg, _ = _jet(128, bc, **KW)
    avg = oswm.rhs(s_avg, *(oswm.apply_bc(s, *x) if apply_bc else x))
    assert rel(ref[1], avg[1]) > 1e-5


def test_energy_enstrophy_rhs_nonsquare_periodic(gpu):
    """A 100 x 60 periodic grid (partial tiles on both axes, images of row and
    column 1 past the last X row and column)."""
    import torch

    from somax_amd import MultilayerShallowWater2D, MultilayerSW2DState

    _log("[swm eec] rhs 100 x 60 periodic, 3 layers")
    args = dict(nx=100, ny=60, Lx=1e6, Ly=6e5, f0=1e-4, beta=1e-11, n_layers=3, H=(500.0, 1000.0, 2500.0),
                g_prime=(9.81, 0.02, 0.01), bc="periodic", lateral_viscosity=20.0)
    model = MultilayerShallowWater2D.create(vorticity_flux="energy_enstrophy", **args)
    s = oswm.setup(vorticity_flux="energy_enstrophy", **{k: v for k, v in args.items() if k != "n_layers"})
    rng = np.random.default_rng(12)
    shape = s.shape
    h = (np.asarray(s.H)[:, None, None] + rng.standard_normal(shape)).astype(np.float32)
    u = (0.2 * rng.standard_normal(shape)).astype(np.float32)
    v = (0.2 * rng.standard_normal(shape)).astype(np.float32)
    st = MultilayerSW2DState(*(torch.from_numpy(a).cuda() for a in (h, u, v)))
    d = model.build_terms()(0.0, st)
    ref = oswm.rhs(s, *oswm.apply_bc(s, *(a.astype(np.float64) for a in (h, u, v))))
    rs = [rel(a.cpu().numpy(), b) for a, b in zip((d.h, d.u, d.v), ref)]
    _log(f"  rel-L2 dh {rs[0]:.3e} du {rs[1]:.3e} dv {rs[2]:.3e}")
    assert max(rs) < 1e-5


@pytest.mark.parametrize("nl,flux", [(6, "average"), (8, "energy_enstrophy")])
def test_deep_stack_rhs_and_step_match_oracle(gpu, nl, flux):
    """6 and 8 layers (the most a plan takes) on a 96 x 64 walled basin: each
    layer's workgroup sums the pressure terms of every layer above it, so the
    deep stacks exercise the longest of those chains."""
    import torch

    from somax_amd import MultilayerShallowWater2D, MultilayerSW2DState

    _log(f"[swm] deep stack nl={nl} flux={flux}")
    w = np.linspace(300.0, 900.0, nl)
    H = tuple(float(x) for x in w * (4000.0 / w.sum()))
    gp = (9.81,) + tuple(0.02 / (1.0 + k) for k in range(nl - 1))
    args = dict(nx=96, ny=64, Lx=9.6e5, Ly=6.4e5, f0=1e-4, beta=1e-11, H=H, g_prime=gp, bc="wall",
                lateral_viscosity=20.0, bottom_drag=1e-7, wind_amplitude=2e-5, vorticity_flux=flux)
    model = MultilayerShallowWater2D.create(n_layers=nl, **args)
    s = oswm.setup(**args)
    rng = np.random.default_rng(40 + nl)
    shape = s.shape
    h = (np.asarray(s.H)[:, None, None] + rng.standard_normal(shape)).astype(np.float32)
    u = (0.2 * rng.standard_normal(shape)).astype(np.float32)
    v = (0.2 * rng.standard_normal(shape)).astype(np.float32)
    f64 = [a.astype(np.float64) for a in (h, u, v)]
    st = MultilayerSW2DState(*(torch.from_numpy(a.copy()).cuda() for a in (h, u, v)))
    d = model.build_terms()(0.0, st)
    ref = oswm.rhs(s, *oswm.apply_bc(s, *f64))
    rs = [rel(a.cpu().numpy(), b) for a, b in zip((d.h, d.u, d.v), ref)]
    dt = 0.3 * 1e4 / np.sqrt(9.81 * 4000.0)
    out = model.step(MultilayerSW2DState(*(torch.from_numpy(a.copy()).cuda() for a in (h, u, v))), dt, 1)
    ref1 = oswm.ssprk3_step(s, *f64, dt)
    r1 = [rel(a.cpu().numpy(), b) for a, b in zip((out.h, out.u, out.v), ref1)]
    _log(f"  rhs rel-L2 dh {rs[0]:.3e} du {rs[1]:.3e} dv {rs[2]:.3e}; "
         f"one step h {r1[0]:.3e} u {r1[1]:.3e} v {r1[2]:.3e}")
    assert max(rs) < 1e-5 and max(r1) < 1e-6


@pytest.mark.parametrize("nx,ny,bc,flux", [(2, 2, "wall", "average"), (3, 2, "periodic", "energy_enstrophy"),
                                            (5, 4, "channel", "average"), (2, 7, "wall", "energy_enstrophy")])
def test_tiny_grid_rhs_and_step_match_oracle(gpu, nx, ny, bc, flux):
    """The smallest grids (down to 2 x 2 cells): one partial tile per
    layer, periodic images and walls a cell apart."""
    import torch

    from somax_amd import MultilayerShallowWater2D, MultilayerSW2DState

    _log(f"[swm] tiny {nx}x{ny} bc={bc} flux={flux}")
    args = dict(nx=nx, ny=ny, Lx=2e4 * nx, Ly=2e4 * ny, f0=1e-4, beta=1e-11, H=(500.0, 3500.0),
                g_prime=(9.81, 0.02), bc=bc, lateral_viscosity=20.0, bottom_drag=1e-7, wind_amplitude=2e-5,
                vorticity_flux=flux)
    model = MultilayerShallowWater2D.create(n_layers=2, **args)
    s = oswm.setup(**args)
    rng = np.random.default_rng(nx * 10 + ny)
    shape = s.shape
    h = (np.asarray(s.H)[:, None, None] + rng.standard_normal(shape)).astype(np.float32)
    u = (0.2 * rng.standard_normal(shape)).astype(np.float32)
    v = (0.2 * rng.standard_normal(shape)).astype(np.float32)
    f64 = [a.astype(np.float64) for a in (h, u, v)]
    st = MultilayerSW2DState(*(torch.from_numpy(a.copy()).cuda() for a in (h, u, v)))
    d = model.build_terms()(0.0, st)
    ref = oswm.rhs(s, *oswm.apply_bc(s, *f64))
    rs = [rel(a.cpu().numpy(), b) for a, b in zip((d.h, d.u, d.v), ref)]
    dt = 0.3 * 2e4 / np.sqrt(9.81 * 4000.0)
    out = model.step(MultilayerSW2DState(*(torch.from_numpy(a.copy()).cuda() for a in (h, u, v))), dt, 1)
    ref1 = oswm.ssprk3_step(s, *f64, dt)
    r1 = [rel(a.cpu().numpy(), b) for a, b in zip((out.h, out.u, out.v), ref1)]
    _log(f"  rhs rel-L2 dh {rs[0]:.3e} du {rs[1]:.3e} dv {rs[2]:.3e}; "
         f"one step h {r1[0]:.3e} u {r1[1]:.3e} v {r1[2]:.3e}")
    assert max(rs) < 1e-5 and max(r1) < 1e-6


@pytest.mark.timeout(300)
def test_jet_thirty_days_conserves_invariants(gpu):
    """The benchmark jet as stated: 2 layers at 512^2 in the channel,
    energy-enstrophy-conserving flux, no dissipation or forcing, SSP-RK3 at
    CFL 0.5 for 30 model days (91,414 steps).  The FP64 oracle cannot run this
    long at this size, so the run is checked by what the discretisation
    conserves: each layer's mass exactly (flux-form continuity; FP32 sums,
    measured 2e-9), and total energy and potential enstrophy up to the time
    stepping (the triad flux conserves both before it; measured -1.5e-6 and
    +3.3e-6).  The 128^2 version is compared with the oracle's statistics in
    tests/test_gpu_longrun.py."""
    import torch

    from somax_amd import MultilayerShallowWater2D, MultilayerSW2DState
    from somax_amd import testcases as tc

    n = 512
    _log(f"[swm] jet {n}^2, 30 days")
    model = MultilayerShallowWater2D.create(nx=n, ny=n, Lx=tc.JET_L, Ly=tc.JET_L, f0=tc.JET_F0, beta=0.0, n_layers=2,
                                            H=tc.JET_H, g_prime=tc.JET_GP, bc="channel",
                                            vorticity_flux="energy_enstrophy")
    h, u, v = (torch.from_numpy(a.astype(np.float32)).cuda() for a in tc.baroclinic_jet_fields(n, seed=3))
    st = MultilayerSW2DState(h=h, u=u, v=v)
    dt = tc.jet_dt(model)
    nsteps = int(30 * 86400 / dt)

    def invariants(state):
        d = model.diagnose(state)
        mass = state.h[:, 1:-1, 1:-1].double().sum(dim=(-2, -1)).cpu().numpy()
        return mass, float(d.total_energy), float(d.total_enstrophy)

    m0, e0, z0 = invariants(st)
    done = 0
    while done < nsteps:
        k = min(5000, nsteps - done)
        st = model.step(st, dt, k)
        done += k
        _log(f"  {done} steps")
    m1, e1, z1 = invariants(st)
    dm = np.abs(m1 - m0) / m0
    de, dz = (e1 - e0) / e0, (z1 - z0) / z0
    _log(f"  mass rel change {dm}, energy {de:.3e}, enstrophy {dz:.3e}")
    assert all(bool(torch.isfinite(f).all()) for f in (st.h, st.u, st.v))
    assert dm.max() < 1e-7 and abs(de) < 1e-4 and abs(dz) < 1e-3


@pytest.mark.parametrize("n,bc", [(128, "periodic"), (128, "wall"), (512, "channel")])
def test_energy_enstrophy_one_step_matches_oracle(gpu, n, bc):
    from somax_amd import testcases as tc

    _log(f"[swm eec] one step n={n} {bc}")
    model, s, f32 = _jet(n, bc, **EEC)
    dt = tc.jet_dt(model)
    out = model.step(_state(f32), dt, 1)
    ref = oswm.ssprk3_step(s, *[a.astype(np.float64) for a in f32], dt)
    rs = [rel(a.cpu().numpy(), b) for a, b in zip((out.h, out.u, out.v), ref)]
    _log(f"  dt={dt:.2f}s rel-L2 h {rs[0]:.3e} u {rs[1]:.3e} v {rs[2]:.3e}")
    assert max(rs) < 1e-6


def test_energy_enstrophy_hundred_steps_match_oracle(gpu):
    from somax_amd import testcases as tc

    _log("[swm eec] 100 steps n=128 channel")
    model, s, f32 = _jet(128, "channel", **EEC)
    dt = tc.jet_dt(model)
    out = model.step(_state(f32), dt, 100)
    ref = oswm.run(s, *[a.astype(np.float64) for a in f32], dt, 100)
    rs = [rel(a.cpu().numpy(), b) for a, b in zip((out.h, out.u, out.v), ref)]
    _log(f"  rel-L2 h {rs[0]:.3e} u {rs[1]:.3e} v {rs[2]:.3e}")
    assert max(rs) < 1e-4
