"""The runner diagnostics restated in oracle/run.py, pinned by the reference's
own cases (tests/test_cli_run.py:52-110) and by a brute-force loop, plus the
host logic of somax_amd.run that needs no GPU (diagnostic grid, formatting)."""

from __future__ import annotations

import math
import struct

import numpy as np
import pytest

from oracle import run as orun


def _brute(values):
    """IEEE-754 classification from the bits, then plain Python arithmetic."""
    lo, hi, total, cnt, bad = math.inf, -math.inf, 0.0, 0, 0
    for v in values:
        bits = struct.unpack("<I", struct.pack("<f", float(v)))[0]
        nonfinite = (bits & 0x7F800000) == 0x7F800000
        nan = nonfinite and (bits & 0x007FFFFF) != 0
        bad += nonfinite
        if not nan:
            lo, hi, total, cnt = min(lo, float(v)), max(hi, float(v)), total + float(v), cnt + 1
    if bad == len(values):
        return None, bad
    return (lo, total / cnt, hi), bad


@pytest.mark.parametrize("seed", range(4))
def test_summary_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(257).astype(np.float32)
    k = rng.integers(0, 5)
    x[rng.choice(257, size=k, replace=False)] = rng.choice([np.nan, np.inf, -np.inf], size=k)
    got = orun.field_summary(x)
    want, bad = _brute(x.tolist())
    assert got["nan"] == bad
    lo, mean, hi = want
    assert got["min"] == lo and got["max"] == hi
    assert got["mean"] == pytest.approx(mean, rel=1e-12) or (math.isnan(mean) and math.isnan(got["mean"])) \
        or (math.isinf(mean) and got["mean"] == mean)


def test_reference_cases():
    """tests/test_cli_run.py:52-110 expressed on the per-field counts."""
    ones = np.ones((3, 4, 4), np.float32)
    d = orun.state_diagnostics({"h": ones, "u": 0 * ones, "v": 0 * ones})
    assert all(s["nan"] == 0 for s in d.values())
    h = ones.copy()
    h[1, 0, 0] = np.nan
    d = orun.state_diagnostics({"h": h, "u": 0 * ones})
    assert d["h"]["nan"] == 1 and d["h"]["min"] == 1.0 and d["h"]["mean"] == 1.0
    d = orun.state_diagnostics({"h": np.full((2, 3, 3), np.inf, np.float32)})
    assert d["h"]["nan"] == 18 and all(math.isnan(d["h"][k]) for k in ("min", "mean", "max"))
    d = orun.state_diagnostics({"h": np.array([np.nan, np.inf, 1.0], np.float32)})
    assert d["h"] == {"min": 1.0, "mean": math.inf, "max": math.inf, "nan": 2}


def test_diagnostic_grid():
    """cli/_run.py:504-538: the save times are a subset at save_indices; the
    chunks subdivide each save interval equally."""
    from somax_amd.run import build_diagnostic_grid

    save = np.array([0.0, 600.0, 1500.0])
    ts, idx = build_diagnostic_grid(save, 1)
    assert np.array_equal(ts, save) and idx == {0, 1, 2}
    ts, idx = build_diagnostic_grid(save, 4)
    assert ts.shape == (9,) and idx == {0, 4, 8}
    assert np.array_equal(ts[sorted(idx)], save)
    assert np.allclose(np.diff(ts[:5]), 150.0) and np.allclose(np.diff(ts[4:]), 225.0)
    ts, idx = build_diagnostic_grid(save, 0)  # clamped to 1
    assert np.array_equal(ts, save)


def test_formatting():
    from somax_amd.run import format_state_stats, format_time_seconds, format_wallclock

    assert format_time_seconds(42) == "42 s" and format_time_seconds(600) == "600 s (10.00 min)"
    assert format_time_seconds(86400) == "8.64e+04 s (1.00 day)"
    assert format_wallclock(0.25) == "250 ms" and format_wallclock(75) == "1m15s"
    line = format_state_stats({"h": {"min": 1.0, "mean": 2.0, "max": 3.0, "nan": 0},
                               "q": {"min": 0.0, "mean": 0.5, "max": 1.0, "nan": 2}})
    assert line == "h[m]=[1,2,3] q[1/s]=[0,0.5,1] NaN=2"
