"""Fixed workloads for hardware-counter and kernel-trace runs.

``python -m somax_amd.profiling --workload qg1024-mixed --steps 2`` first
copies ``--copy-mib`` MiB between two device buffers with one plain copy
kernel (its FETCH_SIZE / WRITE_SIZE give the counters' scale against a known
byte count), then advances the named workload by ``--steps`` SSP-RK3 steps.
Workloads:

* ``qg1024-weno-mass-mixed`` / ``qg1024-weno-mass-fp32`` (default): the
  3-layer QG double gyre at 1024 x 1024 interior cells with WENO-Z PV flux
  and the mass constraint, zero-mean initial streamfunction, mixed or plain
  FP32 inversion (the benchmark's headline workload).
* ``qg1024-mixed`` / ``qg1024-fp32``: the same grid with Arakawa advection
  and no mass constraint.
* ``ensemble`` / ``ensemble-merged``: the 256-member uneven ensemble,
  mixed inversion, one eager step (every launch visible to the profiler),
  one launch per class or merged launch groups.
* ``swm512``: the 2-layer shallow-water baroclinic jet, 512 x 512 channel,
  energy-enstrophy-conserving PV flux, SSP-RK3 at CFL 0.5.
* ``barotropic64-fp64`` / ``barotropic64-weno-fp64``: the 1-layer double
  gyre at 64 x 64 in FP64 (Arakawa or WENO-Z), SSP-RK3 at dt = 4000 s.

Each step is preceded by one untimed warm-up step.  Prints one line per
phase to stdout.

``python -m somax_amd.profiling summarize-trace TRACE.csv``,
``summarize-pmc FETCH.csv WRITE.csv COPY_MIB [OUT.json WORKLOAD]`` and
``summarize-counters COUNTERS.csv [COUNTERS.csv ...]`` turn rocprofv3 CSV
output into the markdown tables kept with the profiles (and the per-kernel
HBM bytes per launch that ``bench.py`` reports as ``traffic``).
"""

from __future__ import annotations

import argparse
import sys


def _grid(r) -> int:
    return int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"]) if "Grid_Size_X" in r else int(
        r["Grid_Size"])


def summarize_trace(path: str, top: int = 40) -> str:
    """Markdown table of a rocprofv3 ``*_kernel_trace.csv``: launches and mean
    duration per (kernel, grid size in work-items)."""
    import collections
    import csv

    agg = collections.defaultdict(list)
    with open(path) as f:
        for r in csv.DictReader(f):
            agg[(r["Kernel_Name"], _grid(r))].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    lines = ["| kernel | grid (work-items) | launches | mean us | total ms |", "|---|---:|---:|---:|---:|"]
    for (name, g), v in sorted(agg.items(), key=lambda kv: -sum(kv[1]))[:top]:
        lines.append(f"| `{name[:70]}` | {g} | {len(v)} | {sum(v) / len(v) / 1e3:.2f} | {sum(v) / 1e6:.3f} |")
    return "\n".join(lines)


# rocprof kernel symbol fragment -> the name the step profile (and bench.py) uses
_BENCH_NAMES = (("dst_gemm_bf16_kernel<1", "dst_gemm_bf16x1"), ("dst_gemm_bf16_kernel<3", "dst_gemm_bf16x3"),
                ("dst_gemm_bf16_pair_kernel<1", "dst_gemm_bf16x1"), ("dst_gemm_bf16_pair_kernel<3", "dst_gemm_bf16x3"),
                ("dst_gemm_bf16_kernel<6", "dst_gemm_bf16x6"), ("dst_gemm_f32_kernel", "dst_gemm_fp32"),
                ("qg_stage_", "qg_stage"), ("residual", "residual"), ("to_modal", "to_modal"),
                ("to_layer", "to_layer"), ("swm_stage_kernel", "swm_stage"))


def bench_name(kernel: str):
    for frag, name in _BENCH_NAMES:
        if frag in kernel:
            return name
    return None


def summarize_pmc(fetch_csv: str, write_csv: str, copy_mib: int, per_name: dict | None = None) -> str:
    """Markdown table of FETCH_SIZE / WRITE_SIZE (KiB per dispatch) per
    (kernel, grid), scaled by what each counter reads for the elementwise copy
    of ``copy_mib`` MiB (the multiply kernel with the largest grid).  When
    ``per_name`` is a dict it receives, per step-profile kernel name, the
    scaled HBM bytes per launch averaged over all dispatches of that name, and
    the two scales."""
    import collections
    import csv

    def load(path):
        agg = collections.defaultdict(list)
        with open(path) as f:
            for r in csv.DictReader(f):
                agg[(r["Kernel_Name"], _grid(r))].append(float(r["Counter_Value"]))
        return agg

    fe, wr = load(fetch_csv), load(write_csv)
    true_kib = copy_mib * 1024.0
    copy_key = max((k for k in fe if "MulFunctor" in k[0]), key=lambda k: k[1])
    f_scale = true_kib / (sum(fe[copy_key]) / len(fe[copy_key]))
    w_scale = true_kib / (sum(wr[copy_key]) / len(wr[copy_key]))
    lines = [f"copy kernel ({copy_mib} MiB read, {copy_mib} MiB written): FETCH_SIZE scale {f_scale:.3f}, "
             f"WRITE_SIZE scale {w_scale:.3f}", "",
             "| kernel | grid | dispatches | FETCH MiB (raw) | WRITE MiB (raw) | HBM MB per launch (scaled) |",
             "|---|---:|---:|---:|---:|---:|"]
    totals = collections.defaultdict(lambda: [0.0, 0])
    for k in sorted(fe, key=lambda k: -sum(fe[k]) / len(fe[k])):
        if k == copy_key or k not in wr:
            continue
        f = sum(fe[k]) / len(fe[k])
        w = sum(wr[k]) / len(wr[k])
        mb = (f * f_scale + w * w_scale) * 1024.0 / 1e6
        lines.append(f"| `{k[0][:70]}` | {k[1]} | {len(fe[k])} | {f / 1024:.2f} | {w / 1024:.2f} | {mb:.2f} |")
        name = bench_name(k[0])
        if name:
            n = min(len(fe[k]), len(wr[k]))
            totals[name][0] += (sum(fe[k][:n]) * f_scale + sum(wr[k][:n]) * w_scale) * 1024.0
            totals[name][1] += n
    if per_name is not None:
        per_name["fetch_scale"] = round(f_scale, 3)
        per_name["write_scale"] = round(w_scale, 3)
        per_name["bytes_per_launch"] = {n: round(t / c) for n, (t, c) in totals.items() if c}
    return "\n".join(lines)


def summarize_counters(paths) -> str:
    """Markdown table of any rocprofv3 ``--pmc`` counters: per (kernel, grid)
    the dispatches and the mean of each counter per dispatch, over one or
    more counter CSVs (one per pass)."""
    import collections
    import csv

    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    names = []
    for path in paths:
        with open(path) as f:
            for r in csv.DictReader(f):
                c = r["Counter_Name"]
                if c not in names:
                    names.append(c)
                agg[(r["Kernel_Name"], _grid(r))][c].append(float(r["Counter_Value"]))
    lines = ["| kernel | grid | dispatches | " + " | ".join(names) + " |",
             "|---|---:|---:|" + "---:|" * len(names)]
    for key in sorted(agg, key=lambda k: -max((sum(v) / len(v) for v in agg[k].values()), default=0.0)):
        row = agg[key]
        n = max(len(v) for v in row.values())
        cells = [f"{sum(row[c]) / len(row[c]):.4g}" if row.get(c) else "" for c in names]
        lines.append(f"| `{key[0][:70]}` | {key[1]} | {n} | " + " | ".join(cells) + " |")
    return "\n".join(lines)


def summarize_issue(paths, cus: int = 256, xcds: int = 8) -> str:
    """Markdown table of the instruction-issue counters per (kernel, grid):
    VALU busy = SQ_ACTIVE_INST_VALU / CUs / (GRBM_GUI_ACTIVE / XCDs) (the
    counter is in quad-cycles, one per wave64 VALU instruction on a 16-lane
    SIMD, and GRBM_GUI_ACTIVE sums the XCDs' active cycles), VALU / SALU / LDS
    instructions per wave, the share of wave cycles spent waiting for an
    instruction's operands (SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES), and, when the
    second pass is given, active lanes per VALU instruction and the LDS
    bank-conflict share of LDS cycles."""
    import collections
    import csv

    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    for path in paths:
        with open(path) as f:
            for r in csv.DictReader(f):
                agg[(r["Kernel_Name"], _grid(r))][r["Counter_Name"]].append(float(r["Counter_Value"]))

    def mean(row, c):
        v = row.get(c)
        return sum(v) / len(v) if v else None

    lines = ["| kernel | grid | dispatches | VALU busy | VALU / wave | SALU / wave | LDS / wave | waiting | "
             "lanes active | LDS conflicts |", "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for key in sorted(agg, key=lambda k: -(mean(agg[k], "SQ_ACTIVE_INST_VALU") or 0.0)):
        row = agg[key]
        gui, act, waves = mean(row, "GRBM_GUI_ACTIVE"), mean(row, "SQ_ACTIVE_INST_VALU"), mean(row, "SQ_WAVES")
        if not gui or act is None or not waves:
            continue
        n = max(len(v) for v in row.values())
        busy = act / cus / (gui / xcds)
        cells = [f"{100 * busy:.0f}%"]
        for c in ("SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS"):
            v = mean(row, c)
            cells.append(f"{v / waves:.0f}" if v is not None else "")
        wt, wc = mean(row, "SQ_WAIT_INST_ANY"), mean(row, "SQ_WAVE_CYCLES")
        cells.append(f"{100 * wt / wc:.0f}%" if wt is not None and wc else "")
        tc = mean(row, "SQ_THREAD_CYCLES_VALU")
        cells.append(f"{100 * tc / (act * 64):.0f}%" if tc is not None and act else "")
        bc, li = mean(row, "SQ_LDS_BANK_CONFLICT"), mean(row, "SQ_LDS_IDX_ACTIVE")
        cells.append(f"{100 * bc / li:.0f}%" if bc is not None and li else "")
        lines.append(f"| `{key[0][:70]}` | {key[1]} | {n} | " + " | ".join(cells) + " |")
    return "\n".join(lines)


def main(argv=None) -> int:
    if argv is None:
        argv = sys.argv[1:]
    if argv and argv[0] == "summarize-trace":
        print(summarize_trace(argv[1]))
        return 0
    if argv and argv[0] == "summarize-counters":
        print(summarize_counters(argv[1:]))
        return 0
    if argv and argv[0] == "summarize-issue":
        print(summarize_issue(argv[1:]))
        return 0
    if argv and argv[0] == "summarize-pmc":
        # summarize-pmc FETCH.csv WRITE.csv COPY_MIB [OUT.json WORKLOAD]
        per = {} if len(argv) > 5 else None
        print(summarize_pmc(argv[1], argv[2], int(argv[3]), per))
        if per is not None:
            import json

            per = {"workload": argv[5], "source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE", **per}
            with open(argv[4], "w") as f:
                json.dump(per, f, indent=1)
        return 0
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", default="qg1024-weno-mass-mixed",
                    choices=["qg1024-weno-mass-mixed", "qg1024-weno-mass-fp32", "qg1024-mixed", "qg1024-fp32",
                             "ensemble", "ensemble-merged", "swm512", "cmm256", "barotropic64-fp64",
                             "barotropic64-weno-fp64"])
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--dt", type=float, default=1200.0)
    args = ap.parse_args(argv)

    import numpy as np
    import torch

    from . import testcases as tc

    if args.copy_mib > 0:
        n = args.copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.mul(src, 1.0, out=dst)  # an elementwise kernel (not a DMA copy)
        torch.cuda.synchronize()
        print(f"copy: {args.copy_mib} MiB read + {args.copy_mib} MiB written", flush=True)
        del src, dst

    if args.workload in ("ensemble", "ensemble-merged"):
        from .ensemble import QGEnsemble

        models, qs = tc.config_ensemble(seed=7, precision="mixed")
        ens = QGEnsemble(models, concurrent=False, graph=False, merged=args.workload == "ensemble-merged")
        packed = ens.pack(qs)
        ens.step(packed, args.dt, 1)
        torch.cuda.synchronize()
        ens.step(packed, args.dt, args.steps)
        torch.cuda.synchronize()
        print(f"ensemble: {len(models)} members, {args.steps} steps", flush=True)
        return 0

    if args.workload == "swm512":
        from . import MultilayerSW2DState

        model, (h, u, v) = tc.baroclinic_jet_swm(512, seed=3, lateral_viscosity=50.0,
                                                  vorticity_flux="energy_enstrophy")
        h, u, v = (torch.from_numpy(np.asarray(a, dtype=np.float32)).cuda() for a in (h, u, v))
        st = MultilayerSW2DState(h=h, u=u, v=v)
        dt = tc.jet_dt(model)
        st = model.step(st, dt, 1)
        torch.cuda.synchronize()
        st = model.step(st, dt, args.steps)
        torch.cuda.synchronize()
        print(f"swm512: {args.steps} steps", flush=True)
        return 0

    if args.workload in ("barotropic64-fp64", "barotropic64-weno-fp64"):
        adv = "weno" if "weno" in args.workload else "arakawa"
        model = tc.config_barotropic(64, precision="fp64", advection=adv)
        q = torch.from_numpy(tc.random_pv(model, seed=8)).cuda()
        model._step_inplace(q, 4000.0, 1)
        torch.cuda.synchronize()
        model._step_inplace(q, 4000.0, args.steps)
        torch.cuda.synchronize()
        print(f"{args.workload}: {args.steps} steps", flush=True)
        return 0

    if args.workload == "cmm256":
        # The octagon of mqg_demo_octagonal.ipynb: 256 nodes, L = 5120 km,
        # three rigid-lid modes; one masked Helmholtz solve per "step".
        from .core import coupling_matrix
        from .elliptic import MaskedHelmholtz, octagon_node_mask

        N, L, f0 = 256, 5120e3, 9.375e-5
        A = coupling_matrix([400.0, 1100.0, 2600.0], [9.81, 0.025, 0.0125])
        A[0, 0] -= 1.0 / (400.0 * 9.81)  # rigid lid, as the notebook's LayerDomain
        lams = f0**2 * np.sort(np.linalg.eigvals(A).real)
        solver = MaskedHelmholtz(octagon_node_mask(N, N), L / (N - 1), L / (N - 1), lams, precision="mixed")
        rhs = torch.randn((3, N, N), device="cuda") * 1e-10
        solver.solve(rhs)
        torch.cuda.synchronize()
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(args.steps):
            solver.solve(rhs)
        t1.record()
        torch.cuda.synchronize()
        print(f"cmm256: {solver.num_boundary_points} boundary points, 3 planes, {args.steps} solves, "
              f"{t0.elapsed_time(t1) / max(args.steps, 1):.4f} ms per solve", flush=True)
        return 0

    precision = "mixed" if args.workload.endswith("mixed") else "fp32"
    weno_mass = "weno-mass" in args.workload
    model = tc.config_baroclinic(1024, precision=precision, advection="weno" if weno_mass else "arakawa",
                                 mass_correction=weno_mass)
    q = torch.from_numpy(tc.random_pv(model, seed=2024, zero_mean=weno_mass).astype(np.float32)).cuda()
    model._step_inplace(q, args.dt, 1)
    torch.cuda.synchronize()
    model._step_inplace(q, args.dt, args.steps)
    torch.cuda.synchronize()
    print(f"{args.workload}: {args.steps} steps", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
