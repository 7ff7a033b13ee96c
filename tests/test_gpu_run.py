"""The runner's chunk protocol on the HIP path (somax_amd/run.py) against the
NumPy restatement of the reference's reductions (oracle/run.py) and the
reference's own runner tests (tests/test_cli_run.py).

Bars: min, max and the non-finite counts exact; the NaN-skipping mean to 1e-12
relative (an FP64 sum in a different order from NumPy's pairwise one).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import run as orun

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


def _same(got, want):
    assert got["nan"] == want["nan"]
    for k in ("min", "max"):
        assert (math.isnan(got[k]) and math.isnan(want[k])) or got[k] == want[k], (k, got, want)
    g, w = got["mean"], want["mean"]
    assert (math.isnan(g) and math.isnan(w)) or (math.isinf(w) and g == w) or g == pytest.approx(w, rel=1e-12, abs=1e-30)


def _planted(rng, n, k):
    x = rng.standard_normal(n).astype(np.float32) * 3.0 + 0.5
    if n and k:
        x[rng.choice(n, size=min(k, n), replace=False)] = rng.choice([np.nan, np.inf, -np.inf], size=min(k, n))
    return x


@pytest.mark.parametrize("n", [0, 1, 3, 5, 4096, 4097, (1 << 20) + 3, 4_500_001])
@pytest.mark.parametrize("offset", [0, 1])
def test_field_stats_matches_numpy(gpu, n, offset):
    """Vector and scalar paths (offset 1 misaligns the view), the tail, the
    block cap (4.5M values > 1024 blocks x 4096)."""
    import torch

    from somax_amd.run import _summary, field_stats

    _log(f"[run] field_stats n={n} offset={offset}")
    rng = np.random.default_rng(n + offset)
    x = _planted(rng, n + offset, 7)
    dev = torch.from_numpy(x).cuda()[offset:]
    rows = field_stats([dev])
    _same(_summary(rows[0], n), orun.field_summary(x[offset:]))


def test_many_fields_and_all_nonfinite(gpu):
    """70 fields (two launches of at most 64), all-NaN, all-Inf and mixed."""
    import torch

    from somax_amd.run import _summary, field_stats

    rng = np.random.default_rng(5)
    xs = [_planted(rng, int(rng.integers(1, 20000)), int(rng.integers(0, 4))) for _ in range(67)]
    xs += [np.full(33, np.nan, np.float32), np.full(9, -np.inf, np.float32), np.array([np.inf, np.nan], np.float32)]
    rows = field_stats([torch.from_numpy(x).cuda() for x in xs])
    assert rows.shape == (70, 5)
    for r, x in zip(rows, xs):
        _same(_summary(r, x.size), orun.field_summary(x))


def test_state_diagnostics_and_assert_finite(gpu):
    """tests/test_cli_run.py:52-110 on device states."""
    import torch

    from somax_amd import IntegrationDivergedError, MultilayerSW2DState, assert_finite_state, state_diagnostics

    ones = torch.ones((3, 4, 4), device="cuda")
    z = torch.zeros((3, 4, 4), device="cuda")
    clean = MultilayerSW2DState(h=ones, u=z, v=z)
    assert_finite_state(clean, mode="run")
    d = state_diagnostics(clean)
    assert d["h"] == {"min": 1.0, "mean": 1.0, "max": 1.0, "nan": 0}
    h = ones.clone()
    h[1, 0, 0] = float("nan")
    with pytest.raises(IntegrationDivergedError) as e:
        assert_finite_state(MultilayerSW2DState(h=h, u=z, v=z), mode="run")
    assert "h" in str(e.value) and "1/" in str(e.value) and "Mid-integration" in str(e.value)
    with pytest.raises(IntegrationDivergedError, match=r"non-finite"):
        assert_finite_state(MultilayerSW2DState(h=torch.full((2, 3, 3), float("inf"), device="cuda"),
                                                u=z[:2, :3, :3], v=z[:2, :3, :3]), mode="run")
    h = ones.clone()
    h[0] = float("nan")
    with pytest.raises(IntegrationDivergedError) as e:
        assert_finite_state(MultilayerSW2DState(h=h, u=z, v=z), mode="run")
    assert "already non-finite at the first saved step" in str(e.value) and "CFL" in str(e.value)
    assert "h (16/16 at t0)" in str(e.value)
    with pytest.raises(IntegrationDivergedError, match=r"spinup integration"):
        assert_finite_state(MultilayerSW2DState(h=torch.full((2, 3, 3), float("nan"), device="cuda"),
                                                u=z[:2, :3, :3], v=z[:2, :3, :3]), mode="spinup")
    d = state_diagnostics(MultilayerSW2DState(h=h, u=z, v=z))
    _same(d["h"], orun.field_summary(h.cpu().numpy()))


def _swm(n=32, **kw):
    from somax_amd import MultilayerShallowWater2D

    return MultilayerShallowWater2D.create(nx=n, ny=n, Lx=1e6, Ly=1e6, f0=1e-4, beta=1.6e-11, n_layers=2,
                                           H=(1000.0, 4000.0), g_prime=(9.81, 0.02), bc="wall", **kw)


def _swm_state(model, seed=0, amp=1.0):
    import torch

    from somax_amd import MultilayerSW2DState

    rng = np.random.default_rng(seed)
    shape = (model.strat.nl, model.grid.Ny, model.grid.Nx)
    a = rng.standard_normal(shape)
    for _ in range(4):
        a = 0.25 * (np.roll(a, 1, -1) + np.roll(a, -1, -1) + np.roll(a, 1, -2) + np.roll(a, -1, -2))
    h = np.asarray(model.strat.H)[:, None, None] + amp * a / np.abs(a).max()
    t = lambda x: torch.from_numpy(np.asarray(x, np.float32)).cuda()  # noqa: E731
    return MultilayerSW2DState(h=t(h), u=t(np.zeros(shape)), v=t(np.zeros(shape)))


def test_chunked_matches_saveat_and_logs(gpu):
    """Chunks reproduce one integrate call with the same save times; the log
    has chunk 0..n plus the closing line and the physical scalars
    (tests/test_cli_run.py:221-272)."""
    import torch

    from somax_amd import chunked_integrate

    _log("[run] chunked SWM 32^2, 2 layers, 4 diagnostic chunks per save")
    model = _swm(lateral_viscosity=50.0, wind_amplitude=1e-5)
    s0 = _swm_state(model)
    dt = 30.0
    lines = []
    res = chunked_integrate(model, s0, [0.0, 300.0, 900.0], dt, diagnostics_per_save=4, log=lines.append)
    ref = model.integrate(s0, 0.0, 900.0, dt, saveat=[300.0, 900.0])
    assert res.ys.h.shape[0] == 3 and torch.equal(res.ys.h[0], s0.h)
    # Each integrate call applies the boundary condition to its initial state
    # (core/model.py:75), so the chunked run resets the entries the wall
    # condition sets (ghost ring, u at the east wall face, v at the north one)
    # at every chunk, while one long call lets them drift under the RHS; only
    # the other entries are compared.  They differ by the float32 rounding of
    # the total thickness (ulp(4000 m) = 2.4e-4 m) at each chunk boundary,
    # which the barotropic gravity wave turns into ~1e-5 m/s of velocity.
    sel = {"h": (slice(1, -1), slice(1, -1)), "u": (slice(1, -1), slice(1, -2)), "v": (slice(1, -2), slice(1, -1))}
    bars = {"h": 1e-6, "u": 1e-3, "v": 1e-3}
    for f in ("h", "u", "v"):
        a, b = getattr(res.ys, f)[1:][(..., *sel[f])], getattr(ref.ys, f)[(..., *sel[f])]
        err, scale = float((a - b).abs().max()), float(b.abs().max())
        _log(f"  {f}: max |chunked - saveat| = {err:.3e} (max |{f}| {scale:.3e})")
        assert err <= bars[f] * scale, f
    assert sum(1 for ln in lines if ln.startswith("chunk ")) == 9
    assert lines[0].startswith("chunk 0/8") and "initial state" in lines[0]
    assert "total_energy=" in lines[1] and "total_enstrophy=" in lines[1]
    assert lines[-1].startswith("all 8 chunks completed")
    assert not any("!!" in ln for ln in lines)


def test_supercritical_dt_raises(gpu):
    """tests/test_cli_run.py:188-217: a time step far above the gravity-wave
    CFL limit (c = sqrt(g H) = 221 m/s, dx = 31 km, dt = 1200 s: CFL 8.5)
    ends the run with IntegrationDivergedError at the chunk where the state
    turns non-finite; the energy-growth warning comes first."""
    from somax_amd import IntegrationDivergedError, chunked_integrate

    _log("[run] supercritical SWM 32^2 dt=1200")
    model = _swm()
    s0 = _swm_state(model, seed=1)
    lines = []
    with pytest.raises(IntegrationDivergedError) as e:
        chunked_integrate(model, s0, np.linspace(0.0, 1200.0 * 400, 41), 1200.0, log=lines.append, mode="spinup")
    msg = str(e.value)
    assert "somax-sim spinup integration produced non-finite values during chunk" in msg
    assert lines[-1].startswith("ABORT at chunk") and "NaN=" in lines[-2]


def test_energy_growth_warning_and_abort(gpu):
    """The protocol itself, on a model whose state grows 20x per chunk: a
    warning on every chunk after the first, then the abort at the chunk
    where float32 overflows (20^30 > 3.4e38)."""
    from dataclasses import dataclass

    import torch

    from somax_amd import IntegrationDivergedError, MultilayerSW2DState, Solution, chunked_integrate

    @dataclass(frozen=True)
    class Energy:
        total_energy: object

    class Grow:
        def integrate(self, s, t0, t1, dt):
            return Solution(ts=torch.tensor([t1]), ys=MultilayerSW2DState(h=(20 * s.h)[None], u=s.u[None],
                                                                           v=s.v[None]))

        def diagnose(self, s):
            return Energy(total_energy=(s.h.double() ** 2).sum())

    ones = torch.ones((1, 8, 8), device="cuda")
    lines = []
    with pytest.raises(IntegrationDivergedError, match=r"during chunk 30/40"):
        chunked_integrate(Grow(), MultilayerSW2DState(h=ones, u=0 * ones, v=0 * ones), np.arange(41.0), 1.0,
                          log=lines.append)
    grew = [ln for ln in lines if "!! energy grew 400.0x" in ln]
    assert len(grew) == 28 and lines[-1] == "ABORT at chunk 30/40: non-finite state"


def test_qg_chunked_and_ensemble_flags(gpu):
    """QG chunk lines carry the QG scalars; the ensemble's per-member flags
    find the one member with planted NaN."""
    import torch

    from somax_amd import QGEnsemble, chunked_integrate, doublegyre_baroclinic_qg

    _log("[run] QG 64^2 chunks and ensemble member flags")
    model, state0 = doublegyre_baroclinic_qg(nx=64, ny=64)
    lines = []
    res = chunked_integrate(model, state0, [0.0, 7200.0], 1200.0, diagnostics_per_save=2, log=lines.append,
                            solver="ssprk3")
    assert res.ys.q.shape[0] == 2 and bool(torch.isfinite(res.ys.q).all())
    assert "total_kinetic_energy=" in lines[1] and "q[1/s]=" in lines[1]
    ens = QGEnsemble([model, doublegyre_baroclinic_qg(nx=32, ny=32)[0], model], concurrent=False, graph=False)
    packed = ens.pack([state0.q, torch.zeros((3, 34, 34)), state0.q])
    packed[ens.offsets[2] + 100] = float("nan")
    flags = ens.member_diagnostics(packed)
    assert [f["nan"] for f in flags] == [0, 0, 1]
    _same(flags[0], orun.field_summary(state0.q.cpu().numpy()))


def test_restart_through_zarr_store(gpu, tmp_path):
    """tests/test_cli_run.py:275-306: a spin-up's final state written to a
    zarr store and read back continues the run as one uninterrupted
    integration: the interior bitwise.  (The restart's integrate call zeroes
    the ghost ring of its initial state, as the reference's does,
    model.py:75; within one call the ghost values evolve with the stage
    combination, and every RHS evaluation re-applies the boundary condition,
    so only the ghost ring can differ.)"""
    import torch

    from somax_amd import doublegyre_baroclinic_qg, io

    _log("[run] QG 64^2 restart through a zarr store")
    model, q0 = doublegyre_baroclinic_qg(nx=64, ny=64)
    dt = 1200.0
    spin = model.integrate(q0, 0.0, 12 * dt, dt, solver="ssprk3")
    final = type(q0)(q=spin.ys.q[-1])
    io.save_dataset(io.state_to_dataset(final, time=12 * dt), tmp_path / "final_state.zarr")
    restart = io.dataset_to_state(io.load_dataset(tmp_path / "final_state.zarr"))
    assert type(restart) is type(q0) and restart.q.device.type == "cuda"
    assert torch.equal(restart.q, final.q)
    cont = model.integrate(restart, 12 * dt, 24 * dt, dt, solver="ssprk3")
    whole = model.integrate(q0, 0.0, 24 * dt, dt, solver="ssprk3")
    a, b = cont.ys.q[-1], whole.ys.q[-1]
    _log(f"  ghost ring max |restart - whole| {float((a - b).abs().max()):.3e}")
    assert torch.equal(a[..., 1:-1, 1:-1], b[..., 1:-1, 1:-1])
