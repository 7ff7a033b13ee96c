#!/usr/bin/env python3
"""Throughput of the MI355X QG time step in layer-cell-steps/s.

Workload (default): the 3-layer QG double gyre at 1024 x 1024 interior cells
(3 x 1026 x 1026 float32 state), H = 400/1100/2600 m, g' = 9.81/0.025/0.0125,
WENO-Z PV flux, biharmonic viscosity 2e11*(1/8)^4 m^4/s, bottom drag
3.6e-8 /s, tau0 = 0.08 N/m^2, dt = 1200 s, the homogeneous-solution mass
constraint (FP64 domain means), seeded k^-3 random initial streamfunction
with zero layer means, advanced by the SSP-RK3 step with the mixed-precision
MFMA inversion (BF16 hi/lo 3-product GEMMs + one refinement pass).  The
plain all-FP32 step on the same state is timed in the same run as the
comparator, and the 256-member uneven ensemble (same physics) beside them.

A layer-cell-step is one interior cell of one layer advanced by one full
SSP-RK3 step.  Inputs are resident in HBM before the timed region; the timed
region is exactly --steps steps bracketed by torch.cuda.synchronize() and HIP
events on the stream the kernels run on.

Prints one JSON line.  ``roofline`` is taken from a second, per-launch
profiled run (HIP events around every launch, somaxhip_qg_step_profile);
``cpu_baseline`` times the FP32 host restatement of the same step with
FFT-based DST-I (oracle/cpu.py:FftCpuStep) on all host threads and on one,
for a bounded sample of the same workload.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

PEAK = {"hbm_GBs": 8000.0, "bf16_TFLOPs": 2500.0, "fp32_TFLOPs": 157.3}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def build_model(n: int, precision: str, advection: str, mass: bool):
    from somax_amd import testcases as tc

    return tc.config_baroclinic(n, precision=precision, advection=advection, mass_correction=mass)


def time_steps(model, q0, dt: float, steps: int, warmup: int):
    import torch

    q = q0.clone()
    if warmup:
        model._step_inplace(q, dt, warmup)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(st)
    model._step_inplace(q, dt, steps)
    e1.record(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1)
    if not torch.isfinite(q).all():
        raise RuntimeError("non-finite state after the timed steps")
    return ms / steps, wall


def kernel_profile(model, q0, dt: float, steps: int):
    import torch

    from somax_amd import _lib

    q = q0.clone()
    plan = model._ensure_plan()
    stats = (_lib.KernelStat * 16)()
    nst = ctypes.c_int(0)
    prm = model._c_params()
    torch.cuda.synchronize()
    _lib.check(
        _lib.load().somaxhip_qg_step_profile(plan, q.data_ptr(), steps, dt, ctypes.byref(prm), model._ws.data_ptr(),
                                            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream), stats, 16,
                                            ctypes.byref(nst)),
        "somaxhip_qg_step_profile",
    )
    out = {}
    for k in range(nst.value):
        s = stats[k]
        out[s.name.decode()] = dict(launches=s.launches, total_ms=s.total_ms, flops=s.flops, bytes=s.bytes)
    return out


def measured_traffic(name: str) -> dict:
    """HBM bytes per launch of a workload's kernels, measured with rocprofv3
    FETCH_SIZE / WRITE_SIZE passes and scaled by a copy of known size
    (profiles/hbm_traffic_*.json); empty if absent."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", name)
    try:
        with open(path) as f:
            return json.load(f)["bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        return {}


def fit_to_step(prof: dict, ms_per_step: float, steps: int) -> float:
    """Factor (at most 1) that scales the HIP-event launch durations of a
    profiled run so that their sum per step does not exceed the timed step:
    events around every single launch add their own overhead, so the raw
    per-launch times over-count; the scaled times share the measured step in
    the raw proportions."""
    per_step = sum(v["total_ms"] for v in prof.values()) / steps
    return min(1.0, ms_per_step / per_step) if per_step > 0 else 1.0


def roofline_entry(name: str, st: dict, scale: float = 1.0) -> dict:
    """Roofline of kernel ``name`` from its profile entry; ``scale`` is the
    ``fit_to_step`` factor applied to its mean launch duration (the raw
    event-timed mean is reported beside it)."""
    raw_ms = st["total_ms"] / st["launches"]
    avg_ms = raw_ms * scale
    if name.startswith("dst_gemm"):
        peak = PEAK["fp32_TFLOPs"] if name.endswith("fp32") else PEAK["bf16_TFLOPs"]
        per_launch = st["flops"] / st["launches"]
        achieved = per_launch / (avg_ms * 1e-3) / 1e12
        issued = achieved * (3.0 if name.endswith("bf16x3") else 1.0)
        return {"kernel": name, "bound": "mfma", "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), "issued_frac": round(issued / peak, 4),
                "avg_launch_us": round(avg_ms * 1e3, 2), "event_launch_us": round(raw_ms * 1e3, 2),
                "alg_flops_per_launch": per_launch, "traffic": None}
    per_launch = st["bytes"] / st["launches"]
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    return {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK["hbm_GBs"], "unit": "GB/s",
            "frac": round(achieved / PEAK["hbm_GBs"], 4), "avg_launch_us": round(avg_ms * 1e3, 2),
            "event_launch_us": round(raw_ms * 1e3, 2), "alg_bytes_per_launch": per_launch, "traffic": None}


def cpu_baseline(n: int, q64, dt: float, steps: int, advection: str, mass: bool) -> dict:
    """The same workload on the host: the FP32 restatement of the step with
    FFT-based DST-I (oracle/cpu.py:FftCpuStep: pocketfft sine transforms,
    OpenMP stencils), timed for ``steps`` steps on every host thread this
    process may use, and for one step on one thread."""
    from oracle import cpu as ocpu
    from oracle import qg as oqg
    from somax_amd import testcases as tc

    setup = oqg.baroclinic_setup(nx=n, ny=n, Lx=tc.BASIN, Ly=tc.BASIN, f0=tc.F0, beta=tc.BETA, bottom_drag=tc.DRAG,
                                 wind_amplitude=tc.wind_amplitude(), biharmonic_viscosity=tc.NU4_128 * (128 / n) ** 4,
                                 advection=advection, mass_correction=mass)
    nl = q64.shape[0]
    if not (ocpu.LIB.exists() and ocpu.RHS_LIB.exists()):
        ocpu.build()
    q32 = q64.astype("float32")
    out = {}
    for label, threads, k in (("all", ocpu.host_threads(), steps), ("one", 1, 1)):
        cs = ocpu.FftCpuStep(setup, threads=threads)
        cs.step(q32, dt, 1)  # warm-up
        t0 = time.perf_counter()
        cs.step(q32, dt, k)
        el = time.perf_counter() - t0
        out[label] = (round(nl * n * n * k / el, 1), threads, k, el)
    v, th, k, el = out["all"]
    v1, _, k1, el1 = out["one"]
    return {"value": v, "unit": "layer-cell-steps/s", "cores": th, "kind": "port",
            "sample": f"{k} SSP-RK3 steps of the same {nl}x{n}x{n} workload ({advection}"
                      f"{', mass constraint' if mass else ''}) by the FP32 host restatement of the step with "
                      f"FFT-based DST-I (oracle/cpu.py:FftCpuStep: pocketfft sine transforms, OpenMP stencils) on "
                      f"{th} host threads, {el:.1f} s",
            "one_thread": {"value": v1, "steps": k1, "sample": f"the same step on one thread, {el1:.1f} s"}}


def bench_ensemble(dt: float, steps: int, warmup: int) -> dict:
    """The 256-member uneven ensemble (96 x 64^2, 96 x 128^2, 48 x 256^2,
    16 x 512^2, layers cycling 1/2/3, tau0 ~ U[0.04, 0.12], WENO-Z PV flux and
    the mass constraint) advanced as one hipGraph-captured ensemble step with
    the mixed inversion, against the same members stepped one after another
    with their own plain FP32 plans."""
    import numpy as np
    import torch

    from somax_amd import QGEnsemble
    from somax_amd import testcases as tc

    log("[bench] ensemble: building 256 members")
    models, qs = tc.config_ensemble(seed=7, precision="mixed")
    cells = sum(m._nl * m.grid.nx * m.grid.ny for m in models)
    out = {"members": len(models), "layer_cells": cells,
           "counts": {f"{n}x{n}": c for n, c in tc.ENSEMBLE_COUNTS},
           "physics": "WENO-Z PV flux, FP64 mass constraint, zero-mean seeded k^-3 initial streamfunctions"}

    def timed(fn, k):
        torch.cuda.synchronize()
        st = torch.cuda.current_stream()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        fn(k)
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / k

    for label, conc, graph, merged in (("graph", True, True, False), ("merged_graph", True, True, True),
                                        ("eager", False, False, False)):
        ens = QGEnsemble(models, concurrent=conc, graph=graph, merged=merged)
        packed = ens.pack(qs)
        ens.step(packed, dt, max(warmup, 1))
        ms = timed(lambda k: ens.step(packed, dt, k), steps)
        if not torch.isfinite(packed).all():
            raise RuntimeError("non-finite ensemble state")
        out[f"mixed_{label}"] = {"ms_per_step": round(ms, 4), "value": round(cells / (ms * 1e-3), 1)}
        log(f"[bench] ensemble mixed {label}: {ms:.3f} ms/step")
        if label in ("graph", "merged_graph"):
            prof = ens.profile(packed, dt, 1)
            key = "kernel_profile_eager" if label == "graph" else "kernel_profile_merged_eager"
            out[key] = {k: {"launches": v[0], "total_us": round(1e3 * v[1], 1)} for k, v in prof.items()}
        del ens, packed

    log("[bench] ensemble: autotune (transform layout, member packing, stage tile rows)")
    ens, report = QGEnsemble.autotune(models, qs, dt, steps=3)
    packed = ens.pack(qs)
    ms = ens.time_step(packed, dt, steps)
    out["mixed_autotuned"] = {"ms_per_step": round(ms, 4), "value": round(cells / (ms * 1e-3), 1),
                              "chosen": {k: v for k, v in report["chosen"].items() if k != "ms_per_step"},
                              "candidates_timed": len(report["candidates"])}
    log(f"[bench] ensemble mixed autotuned: {ms:.3f} ms/step")
    del ens, packed

    log("[bench] ensemble: per-member plain FP32 steps")
    plain = [tc.ensemble_member(m.grid.nx, m._nl, 0.0, precision="fp32", advection="weno", mass_correction=True)
             for m in models]
    for p, m in zip(plain, models):
        p.params = m.params
    qd = [torch.from_numpy(q.astype(np.float32)).cuda() for q in qs]
    for p, q in zip(plain, qd):
        p._step_inplace(q, dt, 1)

    def per_member(k):
        for _ in range(k):
            for p, q in zip(plain, qd):
                p._step_inplace(q, dt, 1)

    ms = timed(per_member, steps)
    out["plain_per_member"] = {"ms_per_step": round(ms, 4), "value": round(cells / (ms * 1e-3), 1)}
    best = min(out["mixed_graph"]["ms_per_step"], out["mixed_autotuned"]["ms_per_step"])
    out["speedup_vs_plain_per_member"] = round(ms / best, 3)
    log(f"[bench] per-member plain: {ms:.3f} ms/step")
    return out


def bench_swm(steps: int, warmup: int) -> dict:
    """The 2-layer shallow-water baroclinic jet: 512 x 512 channel (periodic in
    x, walls in y), f-plane, H = 500/3500 m, g' = 9.81/0.02, balanced 0.5 m/s
    jet plus 1 m interface noise, energy-enstrophy-conserving (triad) PV flux,
    SSP-RK3 at CFL 0.5; one fused stage kernel per RK stage.  The reference's
    average-form flux is timed beside it."""
    import numpy as np
    import torch

    from somax_amd import _lib
    from somax_amd import testcases as tc

    model, (h, u, v) = tc.baroclinic_jet_swm(512, seed=3, lateral_viscosity=50.0, vorticity_flux="energy_enstrophy")
    avg_model, _ = tc.baroclinic_jet_swm(512, seed=3, lateral_viscosity=50.0)
    dt = tc.jet_dt(model)
    fields = [torch.from_numpy(a.astype(np.float32)).cuda() for a in (h, u, v)]
    model._step_inplace(*fields, dt, max(warmup, 1))
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    model._step_inplace(*fields, dt, steps)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    if not all(torch.isfinite(f).all() for f in fields):
        raise RuntimeError("non-finite shallow-water state")
    stats = (_lib.KernelStat * 4)()
    nst = ctypes.c_int(0)
    prm = model._c_params()
    _lib.check(
        _lib.load().somaxhip_swm_step_profile(model._plan, *(f.data_ptr() for f in fields), 2, dt, ctypes.byref(prm),
                                             model._ws.data_ptr(), ctypes.c_void_p(st.cuda_stream), stats, 4,
                                             ctypes.byref(nst)),
        "somaxhip_swm_step_profile",
    )
    k = stats[0]
    cells = model._nl * model.grid.nx * model.grid.ny
    st_prof = {"swm_stage": dict(launches=k.launches, total_ms=k.total_ms, flops=0.0, bytes=k.bytes)}
    roof = roofline_entry("swm_stage", st_prof["swm_stage"], fit_to_step(st_prof, ms, 2))
    roof["traffic"] = measured_traffic("hbm_traffic_swm512_een.json").get("swm_stage")
    avg_fields = [torch.from_numpy(a.astype(np.float32)).cuda() for a in (h, u, v)]
    avg_ms = _event_ms(lambda: avg_model._step_inplace(*avg_fields, dt, 1), steps)
    return {"workload": "swm2x512x512_channel_jet_energy_enstrophy_ssprk3", "dt": round(dt, 3),
            "ms_per_step": round(ms, 4), "value": round(cells / (ms * 1e-3), 1), "unit": "layer-cell-steps/s",
            "roofline": roof,
            "average_flux": {"ms_per_step": round(avg_ms, 4), "value": round(cells / (avg_ms * 1e-3), 1)}}


def bench_cmm(solves: int, cpu_solves: int) -> dict:
    """The masked-domain Helmholtz solve of mqg_demo_octagonal.ipynb: the
    octagon of 256 x 256 nodes, three rigid-lid modes (H = 400/1100/2600 m),
    mixed-precision inversions; one capacitance-matrix solve = two rectangular
    DST-I inversions + the FP64 boundary correction.  CPU beside it: the FP64
    NumPy oracle of the same solve."""
    import time

    import numpy as np
    import torch

    from somax_amd.core import coupling_matrix
    from somax_amd.elliptic import MaskedHelmholtz, octagon_node_mask

    N, L, f0 = 256, 5120e3, 9.375e-5
    A = coupling_matrix([400.0, 1100.0, 2600.0], [9.81, 0.025, 0.0125])
    A[0, 0] -= 1.0 / (400.0 * 9.81)  # rigid lid
    lams = f0**2 * np.sort(np.linalg.eigvals(A).real)
    mask = octagon_node_mask(N, N)
    d = L / (N - 1)
    t = time.perf_counter()
    solver = MaskedHelmholtz(mask, d, d, lams, precision="mixed")
    setup_s = time.perf_counter() - t
    rhs64 = np.random.default_rng(11).standard_normal((3, N, N)) * 1e-10
    rhs = torch.from_numpy(rhs64.astype(np.float32)).cuda()
    psi = solver.solve(rhs)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(solves):
        solver.solve(rhs)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / solves
    pts = int(mask.sum()) * 3
    cpu, err = None, None
    if cpu_solves > 0:  # the CPU leg: the oracle timed on the same solve, its result checks the GPU's
        from oracle import cmm as ocmm

        from threadpoolctl import threadpool_limits

        r = rhs.double().cpu().numpy()
        with threadpool_limits(1):
            t = time.perf_counter()
            for _ in range(cpu_solves):
                want = ocmm.solve(r, mask, d, d, lams, solver.cap, solver.yids, solver.xids)
            cs = (time.perf_counter() - t) / cpu_solves
        got = psi.double().cpu().numpy()
        err = float(f"{np.linalg.norm(got - want) / np.linalg.norm(want):.3e}")
        cpu = {"value": round(pts / cs, 1), "unit": "masked-point-solves/s", "cores": 1, "kind": "port",
               "sample": f"{cpu_solves} solves of the same 3x256x256 octagon, FP64 NumPy oracle (oracle/cmm.py)"}
    return {"workload": "cmm_octagon3x256x256_mixed", "boundary_points": solver.num_boundary_points,
            "ms_per_solve": round(ms, 4), "value": round(pts / (ms * 1e-3), 1), "unit": "masked-point-solves/s",
            "rel_l2_vs_oracle": err, "setup_s": round(setup_s, 2), "cpu_baseline": cpu}


def _event_ms(fn, reps: int) -> float:
    """Mean milliseconds of ``fn()`` over ``reps`` calls, HIP events on the
    current stream (the stream the library launches on)."""
    import torch

    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def bench_rows(mixed, q0, dt: float) -> dict:
    """The remaining paths on their own: the Tsit5 step of the 1024^2 model
    (integrate's default solver, 6 RHS per step), the ReparameterizedQG step
    (3 layers at 512^2, Tsit5: a projection per stage), the single-layer
    nonlinear shallow-water step (1024^2), the FP64 barotropic 64^2 step
    beside the FP64 oracle on one thread, and the between-chunk field
    statistics of the 1024^2 state (one read of 12.6 MB: an HBM-bound
    reduction, reported against the HBM peak)."""
    import ctypes

    import numpy as np
    import torch

    from somax_amd import MultilayerSW2DState, NonlinearShallowWater2D, NonlinearSW2DState, _lib
    from somax_amd import testcases as tc
    from somax_amd.run import _workspace

    out = {}
    nl, Ny, Nx = q0.shape
    cells = nl * (Ny - 2) * (Nx - 2)
    q = q0.clone()
    ms = _event_ms(lambda: mixed._step_tsit5_inplace(q, dt, 1), 10)
    out["qg_tsit5"] = {"workload": f"qg{nl}x{Ny - 2}x{Nx - 2}_tsit5_mixed", "ms_per_step": round(ms, 4),
                       "value": round(cells / (ms * 1e-3), 1), "unit": "layer-cell-steps/s"}

    rq, st = tc.doublegyre_reparameterized_qg(nx=512, ny=512, precision="mixed")
    h, u, v = (f.clone() for f in rq.swm._fields(rq.apply_boundary_conditions(st)))
    ms = _event_ms(lambda: rq._step_inplace(h, u, v, 20.0, 1, "tsit5"), 10)  # dx 7.8 km, c 200 m/s
    c = 3 * 512 * 512
    out["rqg_tsit5"] = {"workload": "reparameterized_qg3x512x512_tsit5_mixed", "ms_per_step": round(ms, 4),
                        "value": round(c / (ms * 1e-3), 1), "unit": "layer-cell-steps/s"}

    sw = NonlinearShallowWater2D.create(nx=1024, ny=1024, bc="wall", lateral_viscosity=10.0)
    shape = (sw.grid.Ny, sw.grid.Nx)
    rng = np.random.default_rng(5)
    h0 = torch.from_numpy((100.0 + 0.1 * rng.standard_normal(shape)).astype(np.float32)).cuda()
    z = torch.zeros(shape, device="cuda")
    s0 = NonlinearSW2DState(h=h0, u=z, v=z.clone())
    ml = sw._sync()
    fields = [f.clone() for f in ml._fields(MultilayerSW2DState(h=h0[None], u=z[None], v=z[None]))]
    ms = _event_ms(lambda: ml._step_inplace(*fields, 5.0, 1), 20)
    out["nonlinear_swm_ssprk3"] = {"workload": "nonlinear_swm1024x1024_wall_ssprk3", "ms_per_step": round(ms, 4),
                                   "value": round(1024 * 1024 / (ms * 1e-3), 1), "unit": "cell-steps/s"}
    del s0

    # the reference's own CPU configuration in FP64: the 1-layer 64^2 double
    # gyre, SSP-RK3 at 4000 s, as precision="fp64" on the GPU and as the FP64
    # NumPy oracle on one host thread (the cpu_baseline leg)
    bt = tc.config_barotropic(64, precision="fp64")
    qb0 = tc.random_pv(bt, seed=8)
    qb = torch.from_numpy(qb0).cuda()
    ms = _event_ms(lambda: bt._step_inplace(qb, 4000.0, 100), 3) / 100
    row = {"workload": "barotropic_qg64x64_ssprk3_fp64", "ms_per_step": round(ms, 4),
           "value": round(64 * 64 / (ms * 1e-3), 1), "unit": "layer-cell-steps/s", "dtype": "f64"}
    from oracle import qg as oqg
    from threadpoolctl import threadpool_limits

    setup = oqg.barotropic_setup(nx=64, ny=64, Lx=tc.BASIN, Ly=tc.BASIN, f0=tc.F0, beta=tc.BETA,
                                 bottom_drag=tc.DRAG, wind_amplitude=tc.wind_amplitude() / 4000.0,
                                 biharmonic_viscosity=tc.NU4_128)
    with threadpool_limits(1):
        t = time.perf_counter()
        oqg.run(setup, qb0, 4000.0, 50)
        cs = (time.perf_counter() - t) / 50
    row["cpu_baseline"] = {"value": round(64 * 64 / cs, 1), "unit": "layer-cell-steps/s", "cores": 1, "kind": "port",
                           "sample": "50 steps of the same model, FP64 NumPy oracle (oracle/qg.py: scipy.fft "
                                     f"sine transforms), {1e3 * cs:.2f} ms per step"}
    out["barotropic_fp64"] = row

    ws = _workspace(q0.device)
    res = torch.empty(5, dtype=torch.float64, device="cuda")
    ptrs = (ctypes.c_void_p * 1)(q0.data_ptr())
    counts = (ctypes.c_size_t * 1)(q0.numel())
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    lib = _lib.load()
    ms = _event_ms(lambda: _lib.check(lib.somaxhip_field_stats(ptrs, counts, 1, res.data_ptr(), ws.data_ptr(),
                                                                stream), "somaxhip_field_stats"), 50)
    nbytes = 4.0 * q0.numel()
    gbs = nbytes / (ms * 1e-3) / 1e9
    out["field_stats"] = {"workload": f"field statistics of the {nl}x{Ny}x{Nx} state", "us_per_call": round(1e3 * ms, 2),
                          "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK["hbm_GBs"],
                                       "unit": "GB/s", "frac": round(gbs / PEAK["hbm_GBs"], 4),
                                       "alg_bytes_per_call": nbytes}}
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=50,
                    help="untimed steps before each timed region (the GPU's clocks settle within a few dozen)")
    ap.add_argument("--n", type=int, default=1024, help="interior cells per side")
    ap.add_argument("--dt", type=float, default=1200.0)
    ap.add_argument("--advection", default="weno", choices=["arakawa", "weno"])
    ap.add_argument("--no-mass", action="store_true", help="leave out the homogeneous-solution mass constraint")
    ap.add_argument("--profile-steps", type=int, default=2)
    ap.add_argument("--cpu-steps", type=int, default=12, help="host steps timed for cpu_baseline (0 = skip)")
    ap.add_argument("--ensemble-steps", type=int, default=10, help="steps timed for the 256-member ensemble (0 = skip)")
    ap.add_argument("--swm-steps", type=int, default=50, help="steps timed for the shallow-water jet (0 = skip)")
    ap.add_argument("--cmm-solves", type=int, default=20, help="masked-domain solves timed (0 = skip)")
    ap.add_argument("--rows", type=int, default=1, help="time the Tsit5 / ReparameterizedQG / nonlinear SWM / "
                    "field-statistics paths (0 = skip)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path "
              f"(--gpus {args.gpus}, WORLD_SIZE {world}); members and rows of a single model are not sharded.",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from somax_amd import testcases as tc

    n, dt = args.n, args.dt
    mass = not args.no_mass
    log(f"[bench] building models n={n} advection={args.advection} mass_constraint={mass}")
    mixed = build_model(n, "mixed", args.advection, mass)
    plain = build_model(n, "fp32", args.advection, mass)
    q64 = tc.random_pv(mixed, seed=2024, zero_mean=mass)
    q0 = torch.from_numpy(q64.astype(np.float32)).cuda()
    nl = q0.shape[0]
    cells = nl * n * n

    log("[bench] timing mixed")
    ms_mixed, wall_mixed = time_steps(mixed, q0, dt, args.steps, args.warmup)
    log(f"[bench] mixed {ms_mixed:.3f} ms/step")
    log("[bench] timing plain fp32")
    ms_plain, _ = time_steps(plain, q0, dt, args.steps, args.warmup)
    log(f"[bench] plain {ms_plain:.3f} ms/step")

    log("[bench] timing bf16x6 (three-piece operands, six products, no refinement)")
    x6 = build_model(n, "bf16x6", args.advection, mass)
    ms_x6, _ = time_steps(x6, q0, dt, args.steps, args.warmup)
    log(f"[bench] bf16x6 {ms_x6:.3f} ms/step")
    del x6

    log("[bench] per-launch profile")
    prof_mixed = kernel_profile(mixed, q0, dt, args.profile_steps)
    prof_plain = kernel_profile(plain, q0, dt, args.profile_steps)
    dom = max(prof_mixed, key=lambda k: prof_mixed[k]["total_ms"])
    fit_mixed = fit_to_step(prof_mixed, ms_mixed, args.profile_steps)
    fit_plain = fit_to_step(prof_plain, ms_plain, args.profile_steps)
    roof = roofline_entry(dom, prof_mixed[dom], fit_mixed)
    roof_stencil = roofline_entry("qg_stage", prof_mixed["qg_stage"], fit_mixed)
    dom_p = max(prof_plain, key=lambda k: prof_plain[k]["total_ms"])
    roof_plain = roofline_entry(dom_p, prof_plain[dom_p], fit_plain)
    workload = f"qg{nl}x{n}x{n}_{args.advection}{'_mass' if mass else ''}_ssprk3"
    if n == 1024:
        tr = measured_traffic(f"hbm_traffic_{workload}_mixed.json")
        roof["traffic"] = tr.get(dom)
        roof_stencil["traffic"] = tr.get("qg_stage")

    ens = bench_ensemble(dt, args.ensemble_steps, 2) if args.ensemble_steps > 0 else None
    swm = None
    if args.swm_steps > 0:
        log("[bench] shallow-water jet 2x512x512")
        swm = bench_swm(args.swm_steps, 2)
        log(f"[bench] swm {swm['ms_per_step']:.3f} ms/step")

    rows = None
    if args.rows:
        log("[bench] Tsit5, ReparameterizedQG, nonlinear SWM, field statistics")
        rows = bench_rows(mixed, q0, dt)
        log("[bench] " + ", ".join(f"{k} {v.get('ms_per_step', v.get('us_per_call'))}" for k, v in rows.items()))

    cmm = None
    if args.cmm_solves > 0:
        log("[bench] masked-domain Helmholtz (octagon 3x256x256)")
        cmm = bench_cmm(args.cmm_solves, 3 if args.cpu_steps > 0 else 0)
        log(f"[bench] cmm {cmm['ms_per_solve']:.3f} ms/solve")

    cpu = None
    if args.cpu_steps > 0:
        log("[bench] cpu baseline (oracle)")
        cpu = cpu_baseline(n, q64, dt, args.cpu_steps, args.advection, mass)

    value = cells / (ms_mixed * 1e-3)
    plain_value = cells / (ms_plain * 1e-3)
    line = {
        "metric": "layer-cell-steps/s: 3-layer QG 1024² and 256-member ensemble vs plain GPU",
        "value": round(value, 1),
        "unit": "layer-cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_mixed, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 stencils; inversion: bf16x3 MFMA GEMMs (bf16x6 on Poisson-like planes' first inverse) "
                 "with f32 accumulation + one f32 refinement pass; f64 layer-to-mode projection and "
                 "mass-constraint means",
        "data": "synthetic (seeded k^-3 random streamfunction" + (", zero layer means)" if mass else ")"),
        "config": {"workload": f"{workload}_mixed", "n": n, "layers": nl, "dt": dt, "advection": args.advection,
                   "mass_constraint": mass, "inversion": "mixed",
                   "value_is": "the single 3-layer 1024^2 mixed step; the plain FP32 step and the 256-member "
                               "ensemble are the plain_fp32 and ensemble objects"},
        "plain_fp32": {"value": round(plain_value, 1), "ms_per_step": round(ms_plain, 4),
                       "speedup_mixed_vs_plain": round(ms_plain / ms_mixed, 3), "roofline": roof_plain},
        "bf16x6": {"value": round(cells / (ms_x6 * 1e-3), 1), "ms_per_step": round(ms_x6, 4),
                   "speedup_vs_plain": round(ms_plain / ms_x6, 3),
                   "form": "bf16 hi/mid/lo operands, 6 products, FP64 accumulation flushes, no refinement pass"},
        "roofline": roof,
        "roofline_stencil": roof_stencil,
        "kernel_profile_mixed": {"steps": args.profile_steps, "fit_to_step": round(fit_mixed, 4),
                                 **{k: {"launches": v["launches"],
                                        "event_avg_us": round(1e3 * v["total_ms"] / v["launches"], 2)}
                                    for k, v in prof_mixed.items()}},
        "ensemble": ens,
        "swm": swm,
        "cmm": cmm,
        "rows": rows,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
