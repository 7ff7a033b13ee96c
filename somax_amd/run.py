"""The runner's chunk protocol with its diagnostics on the device.

Mirror of the chunked integration of ``somax-sim`` (``somax/_src/cli/_run.py``):
``_state_diagnostics`` (:408-439), ``_has_non_finite`` (:499-501),
``_build_diagnostic_grid`` (:504-538), ``_chunked_integrate_with_diagnostics``
(:541-700), ``_assert_finite_state`` (:703-753), ``_flatten_diagnostics``
(:138-167) and ``IntegrationDivergedError`` (:38-45), with the same names
(without the leading underscore), return values and error texts.

The reference copies every field to the host (``np.asarray``) between chunks
to reduce it.  Here the reductions run in ``libsomaxhip.so``
(``somaxhip_field_stats``: min / max / NaN-skipping sum / NaN and non-finite
counts of up to 64 fields in two launches), and five doubles per field come
back.  Physical scalars come from the models' own ``diagnose`` (FP64 sums on
the device); its multi-element fields are summarised by the same kernel.
"""

from __future__ import annotations

import ctypes
import dataclasses
import logging
import math
import time
from types import SimpleNamespace
from typing import Any, Callable

import numpy as np

from . import _lib


def _torch():
    import torch

    return torch


class IntegrationDivergedError(RuntimeError):
    """An integration produced non-finite values (cli/_run.py:38-45); most
    often a time step above the CFL limit."""


# ---------------------------------------------------------------------------
# device reductions
# ---------------------------------------------------------------------------
_WS: dict = {}


def _workspace(device):
    torch = _torch()
    key = str(device)
    ws = _WS.get(key)
    if ws is None:
        nbytes = int(_lib.load().somaxhip_stats_workspace_bytes(_lib.STATS_MAX_FIELDS))
        ws = torch.empty(nbytes, dtype=torch.uint8, device=device)
        _WS[key] = ws
    return ws


def field_stats(fields) -> np.ndarray:
    """``float64[len(fields), 5]``: per field (min, max, sum, n_nan,
    n_nonfinite), with min / max / sum over the non-NaN values.

    ``fields``: device tensors (any shape; non-contiguous ones are made
    contiguous, other dtypes rounded to float32 first, so the min / max of
    an FP64 state are its float32 roundings).  One ``somaxhip_field_stats`` call per 64 fields and one
    small device-to-host copy; raises if the library is missing."""
    torch = _torch()
    fields = list(fields)
    if not fields:
        return np.zeros((0, 5))
    out = []
    for lo in range(0, len(fields), _lib.STATS_MAX_FIELDS):
        chunk = []
        for f in fields[lo:lo + _lib.STATS_MAX_FIELDS]:
            if not isinstance(f, torch.Tensor) or f.device.type != "cuda":
                raise TypeError("field_stats takes device tensors")
            if f.dtype != torch.float32:
                f = f.float()
            chunk.append(f.contiguous())
        dev = chunk[0].device
        nf = len(chunk)
        ptrs = (ctypes.c_void_p * nf)(*[c.data_ptr() if c.numel() else None for c in chunk])
        counts = (ctypes.c_size_t * nf)(*[c.numel() for c in chunk])
        res = torch.empty(5 * nf, dtype=torch.float64, device=dev)
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        _lib.check(
            _lib.load().somaxhip_field_stats(ptrs, counts, nf, res.data_ptr(), _workspace(dev).data_ptr(), stream),
            "somaxhip_field_stats",
        )
        out.append(res.cpu().numpy().reshape(nf, 5))  # keeps `chunk` alive until the kernels have run
    return np.concatenate(out, axis=0)


def _summary(row, size: int) -> dict:
    """The ``_state_diagnostics`` entry of one field from its device stats."""
    mn, mx, total, n_nan, n_bad = (float(x) for x in row)
    n_bad = int(n_bad)
    if n_bad == size:
        nan = float("nan")
        return {"min": nan, "mean": nan, "max": nan, "nan": n_bad}
    return {"min": mn, "mean": total / (size - int(n_nan)), "max": mx, "nan": n_bad}


def _state_fields(state):
    return [(f.name, getattr(state, f.name)) for f in dataclasses.fields(state)]


def state_diagnostics(state) -> dict[str, dict[str, float]]:
    """``{field: {"min", "mean", "max", "nan"}}`` (cli/_run.py:408-439): NaN-
    skipping min / mean / max (``np.nanmin`` etc.; +-Inf take part) and the
    count of non-finite values; all-NaN statistics when every value is
    non-finite."""
    items = _state_fields(state)
    rows = field_stats([v for _, v in items])
    return {name: _summary(rows[i], v.numel()) for i, (name, v) in enumerate(items)}


def has_non_finite(diag: dict[str, dict[str, float]]) -> bool:
    """cli/_run.py:499-501."""
    return any(stats["nan"] > 0 for stats in diag.values())


def assert_finite_state(stacked, *, mode: str) -> None:
    """Raise ``IntegrationDivergedError`` if a time-stacked state holds NaN or
    Inf (cli/_run.py:703-753): names the fields and counts, and whether the
    first saved step was already non-finite."""
    items = _state_fields(stacked)
    fields = [v for _, v in items]
    first = {}  # field index -> row of its first time slice
    for i, v in enumerate(list(fields)):
        if v.dim() >= 1 and v.shape[0] >= 1:
            first[i] = len(fields)
            fields.append(v[0])
    rows = field_stats(fields)
    bad, first_bad = [], []
    for i, (name, v) in enumerate(items):
        n_bad = int(rows[i, 4])
        if not n_bad:
            continue
        bad.append(f"{name} ({n_bad}/{v.numel()} non-finite)")
        if i in first:
            n0 = int(rows[first[i], 4])
            if n0:
                first_bad.append(f"{name} ({n0}/{v[0].numel()} at t0)")
    if not bad:
        return
    lines = [f"somax-sim {mode} integration produced non-finite values:", "  " + "; ".join(bad)]
    if first_bad:
        lines += ["  → already non-finite at the first saved step:", "    " + "; ".join(first_bad),
                  "  This usually means CFL violation at t0 or bad initial conditions."]
    else:
        lines.append("  Mid-integration blow-up. Most common cause: CFL violation "
                     "(time step too large for the grid spacing and the fastest wave).")
    lines.append("  Refusing to write artifacts. Check the timestepping (dt vs dx, wave speeds) and try a smaller dt.")
    raise IntegrationDivergedError("\n".join(lines))


def flatten_diagnostics(diag) -> dict[str, float]:
    """Scalar view of a diagnostics dataclass (cli/_run.py:138-167): 0-d
    values as themselves, 1-d values of at most 16 entries as
    ``<name>_layer_<i>``, larger arrays as ``<name>_mean/_max/_min`` (NaN
    propagates, as with ``ndarray.mean``).  The large arrays are reduced on
    the device."""
    torch = _torch()
    out: dict[str, float] = {}
    if diag is None or not dataclasses.is_dataclass(diag):
        return out
    big = []
    for f in dataclasses.fields(diag):
        value = getattr(diag, f.name)
        if value is None:
            continue
        if isinstance(value, torch.Tensor):
            if value.dim() == 0:
                out[f.name] = float(value)
            elif value.dim() == 1 and value.numel() <= 16:
                for i, x in enumerate(value.detach().cpu().tolist()):
                    out[f"{f.name}_layer_{i}"] = float(x)
            elif value.device.type == "cuda" and value.numel() > 0:
                big.append((f.name, value))
            elif value.numel():
                arr = value.detach().cpu().numpy()
                out[f"{f.name}_mean"], out[f"{f.name}_max"], out[f"{f.name}_min"] = (
                    float(arr.mean()), float(arr.max()), float(arr.min()))
            continue
        try:
            arr = np.asarray(value, dtype=np.float64)
        except (TypeError, ValueError):
            continue
        if arr.ndim == 0:
            out[f.name] = float(arr)
        elif arr.ndim == 1 and arr.size <= 16:
            for i, x in enumerate(arr.tolist()):
                out[f"{f.name}_layer_{i}"] = float(x)
        elif arr.size:
            out[f"{f.name}_mean"], out[f"{f.name}_max"], out[f"{f.name}_min"] = (
                float(arr.mean()), float(arr.max()), float(arr.min()))
    if big:
        rows = field_stats([v for _, v in big])
        for (name, v), row in zip(big, rows):
            if row[3] > 0:
                mean = mx = mn = float("nan")
            else:
                mean, mx, mn = float(row[2]) / v.numel(), float(row[1]), float(row[0])
            out[f"{name}_mean"], out[f"{name}_max"], out[f"{name}_min"] = mean, mx, mn
    return out


_HEADLINE = ("total_energy", "total_kinetic_energy", "energy", "kinetic_energy", "total_enstrophy", "enstrophy")
_ENERGY_KEYS = ("total_energy", "energy", "total_kinetic_energy", "kinetic_energy")


def physical_scalars(model, state) -> tuple[str, dict[str, float]]:
    """``(line, flat)`` from ``model.diagnose(state)`` (cli/_run.py:456-496);
    never raises (a failing ``diagnose`` gives its message and ``{}``)."""
    try:
        flat = flatten_diagnostics(model.diagnose(state))
    except Exception as exc:  # the reference reports and carries on
        return f"diagnose failed: {type(exc).__name__}: {exc}", {}
    if not flat:
        return "", flat
    parts = [f"{k}={flat[k]:.3g}" for k in _HEADLINE if k in flat]
    layers = sorted(k for k in flat if k.startswith("energy_layer_"))
    if layers and "total_energy" not in flat and "energy" not in flat:
        parts.append("energy_layers=[" + ",".join(f"{flat[k]:.3g}" for k in layers) + "]")
    return " ".join(parts), flat


def build_diagnostic_grid(save_ts, diagnostics_per_save: int):
    """``(diag_ts, save_indices)`` (cli/_run.py:504-538): every save interval
    split into ``diagnostics_per_save`` equal chunks; ``save_ts ==
    diag_ts[sorted(save_indices)]``."""
    save_ts = np.asarray(save_ts, dtype=np.float64)
    n = max(int(diagnostics_per_save), 1)
    if n == 1:
        return save_ts, set(range(save_ts.shape[0]))
    pts = [float(save_ts[0])]
    for a, b in zip(save_ts[:-1], save_ts[1:]):
        pts.extend(float(x) for x in np.linspace(a, b, n + 1)[1:])
    return np.asarray(pts), {i * n for i in range(save_ts.shape[0])}


# ---------------------------------------------------------------------------
# log formatting (cli/_units.py)
# ---------------------------------------------------------------------------
FIELD_UNITS = {"h": "m", "u": "m/s", "v": "m/s", "q": "1/s", "psi": "m²/s"}
_SIM_UNITS = ((60.0, "s", 1.0), (3600.0, "min", 60.0), (86400.0, "hr", 3600.0), (2_592_000.0, "day", 86400.0),
              (31_557_600.0, "month", 2_592_000.0), (math.inf, "yr", 31_557_600.0))


def format_time_seconds(seconds: float) -> str:
    if seconds < 60.0:
        return f"{int(seconds)} s" if seconds == int(seconds) else f"{seconds:.3g} s"
    for cut, label, div in _SIM_UNITS:
        if seconds < cut:
            return f"{seconds:.3g} s ({seconds / div:.2f} {label})"
    return f"{seconds:.3g} s"


def format_wallclock(seconds: float) -> str:
    if seconds < 1.0:
        return f"{seconds * 1000:.0f} ms"
    if seconds < 60.0:
        return f"{seconds:.2f} s"
    m, s = divmod(seconds, 60.0)
    if m < 60:
        return f"{int(m)}m{s:.0f}s"
    h, m = divmod(m, 60)
    return f"{int(h)}h{int(m)}m"


def format_state_stats(diag: dict[str, dict[str, float]]) -> str:
    parts = []
    for name, s in diag.items():
        unit = FIELD_UNITS.get(name, "")
        body = f"[{s['min']:.3g},{s['mean']:.3g},{s['max']:.3g}]" + (f" NaN={s['nan']}" if s["nan"] else "")
        parts.append(f"{name}{f'[{unit}]' if unit else ''}={body}")
    return " ".join(parts)


# ---------------------------------------------------------------------------
# the chunk loop
# ---------------------------------------------------------------------------
def _last(stacked, cls):
    return cls(**{f.name: getattr(stacked, f.name)[-1] for f in dataclasses.fields(cls)})


def chunked_integrate(model, state0, save_ts, dt: float, *, diagnostics_per_save: int = 1,
                      log: Callable[[str], Any] | None = None, mode: str = "run", **integrate_kw) -> SimpleNamespace:
    """``_chunked_integrate_with_diagnostics`` (cli/_run.py:541-700).

    One ``model.integrate`` call per diagnostic chunk (each saves its end
    state only); between chunks the state's min / mean / max / non-finite
    counts and the physical scalars are reduced on the device and written to
    ``log`` (default: ``logging.getLogger("somax_amd.run").debug``), with a
    warning when the energy grows more than tenfold from one chunk to the
    next.  A non-finite state aborts the run with ``IntegrationDivergedError``
    at the chunk where it appears.  Returns ``SimpleNamespace(ys=..., ts=...,
    stats={})``: the states at ``save_ts`` stacked on a leading time axis.
    ``integrate_kw`` (e.g. ``solver``) goes to ``model.integrate``."""
    torch = _torch()
    log = log or logging.getLogger("somax_amd.run").debug
    save_ts = np.asarray(save_ts, dtype=np.float64)
    diag_ts, save_idx = build_diagnostic_grid(save_ts, diagnostics_per_save)
    n = diag_ts.shape[0] - 1
    phys, _ = physical_scalars(model, state0)
    log(f"chunk 0/{n} sim_t={format_time_seconds(float(diag_ts[0]))} | {format_state_stats(state_diagnostics(state0))}"
        + (f" | physics: {phys}" if phys else "") + " | initial state")
    state, saved = state0, [state0]
    start = time.perf_counter()
    prev_energy = None
    for i in range(n):
        t0, t1 = float(diag_ts[i]), float(diag_ts[i + 1])
        tc = time.perf_counter()
        sol = model.integrate(state, t0, t1, dt, **integrate_kw)
        new = _last(sol.ys, type(state0))
        sd = state_diagnostics(new)
        wall = time.perf_counter() - tc
        phys, flat = physical_scalars(model, new)
        energy = next((float(flat[k]) for k in _ENERGY_KEYS if k in flat), None)
        warn = ""
        if (prev_energy is not None and energy is not None and abs(prev_energy) > 0 and not math.isnan(energy)
                and abs(energy) > 10 * abs(prev_energy)):
            warn = f" !! energy grew {energy / prev_energy:.1f}x from previous chunk"
        if energy is not None and not math.isnan(energy):
            prev_energy = energy
        log(f"chunk {i + 1}/{n} sim_t={format_time_seconds(t1)} | {format_state_stats(sd)}"
            + (f" | physics: {phys}" if phys else "") + f" | wall={format_wallclock(wall)}" + warn)
        if has_non_finite(sd):
            log(f"ABORT at chunk {i + 1}/{n}: non-finite state")
            raise IntegrationDivergedError(
                f"somax-sim {mode} integration produced non-finite values during chunk {i + 1}/{n} "
                f"(sim_t={format_time_seconds(t1)}).\n  {format_state_stats(sd)}\n"
                f"  Refusing to write artifacts. The non-finite values appeared between "
                f"sim_t={format_time_seconds(t0)} and sim_t={format_time_seconds(t1)} — earlier chunks were "
                f"finite. This is consistent with a slow numerical instability, not an immediate CFL violation.")
        if (i + 1) in save_idx:
            saved.append(new)
        state = new
    log(f"all {n} chunks completed in {format_wallclock(time.perf_counter() - start)}")
    cls = type(state0)
    ys = cls(**{f.name: torch.stack([getattr(s, f.name) for s in saved]) for f in dataclasses.fields(cls)})
    return SimpleNamespace(ys=ys, ts=save_ts, stats={})
