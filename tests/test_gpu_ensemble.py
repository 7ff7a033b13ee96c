"""GPU tests of QG ensembles (members of uneven grid size and layer count).

Each member of an ensemble runs the arithmetic of the single-model step, so
the bar against ``model.step`` is agreement to FP32 round-off (the kernels are
deterministic and identical per plane; 1e-7 relative leaves room only for a
different launch order of identical work).  Against the FP64 oracle the
one-step bar is the single model's, 1e-6.  Graph replays are checked one by
one against eager steps, with unrelated eager work and synchronisations in
between.
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


# (nx, ny, nl, kind): kind "bc" = BaroclinicQG, "bt" = BarotropicQG
MEMBERS = [(64, 64, 1, "bc"), (64, 64, 3, "bc"), (64, 64, 1, "bt"), (96, 64, 2, "bc"), (128, 128, 3, "bc"),
           (64, 64, 2, "bc"), (96, 64, 2, "bc")]


def _specs(seed=5):
    """(kind, create kwargs) per member; tau0 drawn uniformly in [0.04, 0.12]."""
    from somax_amd import testcases as tc

    rng = np.random.default_rng(seed)
    out = []
    for nx, ny, nl, kind in MEMBERS:
        tau0 = float(rng.uniform(0.04, 0.12))
        if kind == "bt":
            out.append(("bt", dict(nx=nx, ny=ny, Lx=tc.BASIN, Ly=tc.BASIN, f0=tc.F0, beta=tc.BETA,
                                   bottom_drag=tc.DRAG, wind_amplitude=tc.wind_amplitude(tau0) / 4000.0,
                                   biharmonic_viscosity=tc.NU4_128)))
        else:
            out.append(("bc", tc.ensemble_member_args(nx, nl, tau0, ny=ny)))
    return out


def _models(precision, seed=5):
    from somax_amd import BaroclinicQG, BarotropicQG

    return [BarotropicQG.create(**a, precision=precision) if k == "bt" else
            BaroclinicQG.create(**a, precision=precision) for k, a in _specs(seed)]


def _oracle(i, seed=5):
    kind, a = _specs(seed)[i]
    if kind == "bt":
        return oqg.barotropic_setup(**a)
    a = {k: v for k, v in a.items() if k != "n_layers"}
    return oqg.baroclinic_setup(**a)


def _states(models, seed=21):
    from somax_amd import testcases as tc

    return [tc.random_pv(m, seed + i).astype(np.float32).reshape(m._nl, m.grid.Ny, m.grid.Nx)
            for i, m in enumerate(models)]


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
@pytest.mark.parametrize("concurrent,graph,merged", [(False, False, False), (True, False, False),
                                                     (True, True, False), (False, False, True),
                                                     (True, True, True)])
def test_ensemble_matches_single_models(gpu, precision, concurrent, graph, merged):
    import torch

    from somax_amd import QGEnsemble

    _log(f"[ensemble] precision={precision} concurrent={concurrent} graph={graph} merged={merged}")
    models = _models(precision)
    qs = _states(models)
    ens = QGEnsemble(models, concurrent=concurrent, graph=graph, merged=merged)
    assert len(ens) == len(MEMBERS)
    packed = ens.pack(qs)
    dt, nsteps = 1200.0, 3
    ens.step(packed, dt, nsteps)
    torch.cuda.synchronize()
    worst = 0.0
    for i, m in enumerate(models):
        q = torch.from_numpy(qs[i]).cuda()
        m._step_inplace(q, dt, nsteps)
        r = rel(ens.member(packed, i).cpu().numpy(), q.cpu().numpy())
        worst = max(worst, r)
    _log(f"  worst member rel-L2 vs single model = {worst:.3e}")
    assert worst < 1e-7


def _deep_member_args(nx, ny, nl, tau0):
    """A double-gyre member with an nl-layer stack of 4000 m (reduced
    gravities falling with depth), for layer counts past the 1-3 of MEMBERS."""
    from somax_amd import testcases as tc

    w = np.linspace(300.0, 700.0, nl)
    H = tuple(float(x) for x in w * (4000.0 / w.sum()))
    gp = (9.81,) + tuple(0.03 / (1.0 + k) for k in range(nl - 1))
    return dict(nx=nx, ny=ny, Lx=tc.BASIN, Ly=tc.BASIN * ny / nx, f0=tc.F0, beta=tc.BETA, n_layers=nl, H=H,
                g_prime=gp, bottom_drag=tc.DRAG, wind_amplitude=tc.wind_amplitude(tau0, tc.BASIN * ny / nx),
                biharmonic_viscosity=tc.NU4_128 * (128.0 / nx) ** 4)


@pytest.mark.parametrize("graph,merged", [(True, False), (False, True)])
def test_ensemble_deep_members_match_single_models(gpu, graph, merged):
    """Members of 4, 5 and 8 layers beside shallow ones in one class (the
    kernels' 4- and 8-slot instances over a mixed plane table), plus an
    8-layer member on a folded 256^2 grid; WENO-Z and the mass constraint."""
    import torch

    from somax_amd import BaroclinicQG, QGEnsemble
    from somax_amd import testcases as tc

    _log(f"[ensemble] deep members graph={graph} merged={merged}")
    shapes = [(64, 64, 5), (64, 64, 8), (64, 64, 1), (64, 64, 4), (96, 64, 8), (256, 256, 8)]
    rng = np.random.default_rng(17)
    models = [BaroclinicQG.create(**_deep_member_args(nx, ny, nl, float(rng.uniform(0.04, 0.12))),
                                  precision="mixed", advection="weno", mass_correction=True) for nx, ny, nl in shapes]
    qs = [tc.random_pv(m, 90 + i, zero_mean=True).astype(np.float32).reshape(m._nl, m.grid.Ny, m.grid.Nx)
          for i, m in enumerate(models)]
    ens = QGEnsemble(models, concurrent=True, graph=graph, merged=merged)
    packed = ens.pack(qs)
    dt, nsteps = 1200.0, 2
    ens.step(packed, dt, nsteps)
    torch.cuda.synchronize()
    worst = 0.0
    for i, m in enumerate(models):
        q = torch.from_numpy(qs[i]).cuda()
        m._step_inplace(q, dt, nsteps)
        worst = max(worst, rel(ens.member(packed, i).cpu().numpy(), q.cpu().numpy()))
    _log(f"  worst member rel-L2 vs single model = {worst:.3e}")
    assert worst < 1e-7


def test_ensemble_graph_replays(gpu):
    import torch

    from somax_amd import QGEnsemble

    _log("[ensemble] graph replays vs eager, mixed")
    models = _models("mixed")
    qs = _states(models)
    eager = QGEnsemble(models, concurrent=False, graph=False)
    graph = QGEnsemble(models, concurrent=True, graph=True)
    pe, pg = eager.pack(qs), graph.pack(qs)
    scratch = torch.randn(1 << 20, device="cuda")
    for k in range(4):
        graph.step(pg, 1200.0, 1)
        scratch.mul_(1.0001).add_(0.5)  # eager work between replays
        torch.cuda.synchronize()
        eager.step(pe, 1200.0, 1)
        torch.cuda.synchronize()
        r = rel(pg.cpu().numpy(), pe.cpu().numpy())
        _log(f"  replay {k}: rel-L2 = {r:.3e}")
        assert r < 1e-7
    # two replays per call
    graph.step(pg, 1200.0, 2)
    eager.step(pe, 1200.0, 2)
    torch.cuda.synchronize()
    assert rel(pg.cpu().numpy(), pe.cpu().numpy()) < 1e-7


def test_ensemble_one_step_vs_oracle(gpu):
    import torch

    from somax_amd import QGEnsemble

    _log("[ensemble] one step vs FP64 oracle, mixed")
    models = _models("mixed")
    qs = _states(models)
    ens = QGEnsemble(models)
    packed = ens.pack(qs)
    dt = 1200.0
    ens.step(packed, dt, 1)
    torch.cuda.synchronize()
    for i in (0, 1, 2, 4):
        setup = _oracle(i)
        q0 = qs[i].astype(np.float64)
        if setup.nl == 1 and MEMBERS[i][3] == "bt":
            q0 = q0[0]
        ref = oqg.ssprk3_step(setup, q0, dt)
        r = rel(ens.member(packed, i).cpu().numpy().reshape(ref.shape), ref)
        _log(f"  member {i} {MEMBERS[i]}: rel-L2 q = {r:.3e}")
        assert r < 1e-6


def test_ensemble_set_params(gpu):
    import dataclasses

    import torch

    from somax_amd import QGEnsemble

    _log("[ensemble] set_params")
    models = _models("fp32")
    qs = _states(models)
    ens = QGEnsemble(models, concurrent=True, graph=True)
    packed = ens.pack(qs)
    ens.step(packed, 1200.0, 1)  # captures a graph with the first parameters
    new = [dataclasses.replace(m.params, wind_amplitude=2.0 * m.params.wind_amplitude, biharmonic_viscosity=0.0)
           for m in models]
    ens.set_params(new)
    packed = ens.pack(qs)
    ens.step(packed, 1200.0, 2)
    torch.cuda.synchronize()
    for i, m in enumerate(models):
        assert m.params.biharmonic_viscosity == 0.0
        q = torch.from_numpy(qs[i]).cuda()
        m._step_inplace(q, 1200.0, 2)
        assert rel(ens.member(packed, i).cpu().numpy(), q.cpu().numpy()) < 1e-7


def test_merged_groups_beyond_eight_classes(gpu):
    """Ten grid classes merge into two launch groups (at most eight classes
    per merged launch); every member still equals its single model bit for
    bit after two steps."""
    import torch

    from somax_amd import QGEnsemble
    from somax_amd import testcases as tc

    _log("[ensemble] 10 classes, merged launch groups, mixed")
    models = [tc.ensemble_member(n, 1 + i % 3, 0.04 + 0.008 * i, precision="mixed")
              for i, n in enumerate(range(32, 112, 8))]
    ens = QGEnsemble(models, concurrent=True, graph=True, merged=True)
    assert ens.num_classes == 10
    qs = _states(models)
    packed = ens.pack(qs)
    ens.step(packed, 1200.0, 2)
    torch.cuda.synchronize()
    for i, m in enumerate(models):
        q = torch.from_numpy(qs[i]).cuda()
        m._step_inplace(q, 1200.0, 2)
        assert torch.equal(ens.member(packed, i), q.view(ens.shapes[i])), i


def test_ensemble_mixed_precision_classes(gpu):
    import torch

    from somax_amd import QGEnsemble
    from somax_amd import testcases as tc

    _log("[ensemble] fp32 and mixed members of one grid form two classes")
    models = [tc.ensemble_member(64, 3, 0.08, precision="fp32"), tc.ensemble_member(64, 3, 0.08, precision="mixed")]
    ens = QGEnsemble(models)
    assert ens.num_classes == 2
    qs = _states(models)
    packed = ens.pack(qs)
    ens.step(packed, 1200.0, 1)
    torch.cuda.synchronize()
    for i, m in enumerate(models):
        q = torch.from_numpy(qs[i]).cuda()
        m._step_inplace(q, 1200.0, 1)
        assert rel(ens.member(packed, i).cpu().numpy(), q.cpu().numpy()) < 1e-7


@pytest.mark.parametrize("rows", [8])
def test_stage_tile_rows_leave_results_unchanged(gpu, rows):
    """The autotuner's stage-tile knob: 8-row tiles give the 16-row results
    bit for bit (same arithmetic, other launch shape); the knob is refused
    once the ensemble has stepped."""
    import torch

    from somax_amd import QGEnsemble

    _log(f"[ensemble] stage tile rows {rows} vs 16, mixed")
    models = _models("mixed")
    qs = _states(models)
    ref = QGEnsemble(models, concurrent=False, graph=False)
    alt = QGEnsemble(models, concurrent=False, graph=True)
    for c in range(alt.num_classes):
        alt.set_stage_rows(c, rows)
    assert all(info["stage_rows"] == rows for info in alt.class_info())
    pr, pa = ref.pack(qs), alt.pack(qs)
    ref.step(pr, 1200.0, 3)
    alt.step(pa, 1200.0, 3)
    torch.cuda.synchronize()
    assert torch.equal(pr, pa)
    with pytest.raises(RuntimeError, match="fixed once the ensemble has stepped"):
        alt.set_stage_rows(0, 16)


def test_autotune_picks_a_valid_configuration(gpu):
    """QGEnsemble.autotune: every candidate timed, the chosen ensemble the
    fastest of them, and its step still matches the FP64 oracle."""
    import torch

    from somax_amd import QGEnsemble

    _log("[ensemble] autotune, mixed")
    models = _models("mixed")
    qs = _states(models)
    ens, report = QGEnsemble.autotune(models, qs, 1200.0, steps=2, log=_log)
    chosen = report["chosen"]
    timed = [c["ms_per_step"] for c in report["candidates"]]
    assert len(report["candidates"]) >= 6 and chosen["ms_per_step"] <= min(timed) + 1e-9
    assert len(chosen["classes"]) == ens.num_classes
    packed = ens.pack(qs)
    ens.step(packed, 1200.0, 1)
    torch.cuda.synchronize()
    for i in (0, 4):
        setup = _oracle(i)
        ref = oqg.ssprk3_step(setup, qs[i].astype(np.float64), 1200.0)
        assert rel(ens.member(packed, i).cpu().numpy().reshape(ref.shape), ref) < 1e-6


@pytest.mark.parametrize("merged", [False, True])
def test_uneven_256_member_ensemble(gpu, merged):
    """The 256-member ensemble: 96 members at 64^2, 96 at 128^2, 48 at 256^2
    and 16 at 512^2 (the 256^2 and 512^2 classes with folded transforms),
    layer counts cycling 1/2/3, tau0 ~ U[0.04, 0.12] N/m^2, WENO-Z PV flux
    with the FP64 mass constraint, mixed inversion, in one captured graph:
    concurrent class streams, or one launch group whose every kernel phase
    is a single launch over the four classes.  After one step every member
    equals its single model bit for bit, and one member per (grid, layer
    count) meets the one-step bar (1e-6) against the FP64 oracle."""
    import torch

    from somax_amd import QGEnsemble
    from somax_amd import testcases as tc

    _log(f"[ensemble] 256 uneven members, WENO-Z + mass constraint, mixed, graph, merged={merged}")
    models, qs = tc.config_ensemble(seed=3, precision="mixed")
    assert len(models) == 256
    ens = QGEnsemble(models, concurrent=True, graph=True, merged=merged)
    info = ens.class_info()
    assert [(c["nx"], c["members"], c["fold"]) for c in info] == [(512, 16, 1), (256, 48, 1), (128, 96, 0),
                                                                  (64, 96, 0)]
    packed = ens.pack(qs)
    dt = 1200.0
    ens.step(packed, dt, 1)
    torch.cuda.synchronize()
    assert torch.isfinite(packed).all()
    mismatched = []
    for i, m in enumerate(models):
        q = torch.from_numpy(qs[i]).cuda()
        m._step_inplace(q, dt, 1)
        if not torch.equal(ens.member(packed, i), q.view(ens.shapes[i])):
            mismatched.append(i)
    _log(f"  members differing from their single model: {mismatched}")
    assert not mismatched
    seen = set()
    for i, m in enumerate(models):
        key = (m.grid.nx, m._nl)
        if key in seen:
            continue
        seen.add(key)
        a = {k: v for k, v in tc.ensemble_member_args(m.grid.nx, m._nl, 0.0).items() if k != "n_layers"}
        a["wind_amplitude"] = m.params.wind_amplitude  # the member's own tau0
        setup = oqg.baroclinic_setup(advection="weno", mass_correction=True, **a)
        ref = oqg.ssprk3_step(setup, qs[i].astype(np.float64), dt, fast=True)
        r = rel(ens.member(packed, i).cpu().numpy(), ref)
        _log(f"  member {i} ({m.grid.nx}^2, {m._nl} layers): rel-L2 q vs oracle = {r:.3e}")
        assert r < 1e-6
    assert len(seen) == 12
    if merged:
        # one launch per kernel phase per RK stage for all four classes
        prof = ens.profile(packed, dt, 1)
        launches = {k: v[0] for k, v in prof.items()}
        _log(f"  launches per step: {launches}")
        for k in ("to_modal", "residual", "mass_constraint", "to_layer", "qg_stage"):
            assert launches[k] == 3, k
        assert launches["dst_gemm_bf16x3"] == 12 and launches["dst_gemm_bf16x6"] == 6
        assert launches["dst_gemm_bf16x1"] == 6


@pytest.mark.parametrize("merged", [False, True])
def test_ensemble_graph_recaptures(gpu, merged):
    """A concurrent, graph-captured ensemble recaptures its step whenever the
    parameters switch the stage kernel (biharmonic on / off, which also
    regroups merged launch groups), dt changes or the state buffer moves; five
    captures in a row on one ensemble, each followed by steps checked bit for
    bit against the single models."""
    import dataclasses

    import torch

    from somax_amd import QGEnsemble

    _log(f"[ensemble] five graph captures on one concurrent ensemble, merged={merged}")
    models = _models("mixed")
    qs = _states(models)
    ens = QGEnsemble(models, concurrent=True, graph=True, merged=merged)
    base = [m.params for m in models]
    no_bih = [dataclasses.replace(p, biharmonic_viscosity=0.0) for p in base]
    packed = ens.pack(qs)
    singles = [torch.from_numpy(q).cuda() for q in qs]
    plan = [("first capture", None, 1200.0, False), ("biharmonic off", no_bih, 1200.0, False),
            ("dt changed", None, 900.0, False), ("state moved", None, 900.0, True),
            ("biharmonic on", base, 900.0, False)]
    for k, (what, params, dt, move) in enumerate(plan):
        if params is not None:
            ens.set_params(params)
        if move:
            packed = packed.clone()
        ens.step(packed, dt, 2)
        for m, q in zip(models, singles):
            m._step_inplace(q, dt, 2)
        torch.cuda.synchronize()
        same = all(torch.equal(ens.member(packed, i), q.view(ens.shapes[i])) for i, q in enumerate(singles))
        _log(f"  capture {k + 1} ({what}): members equal their single models: {same}")
        assert same
