"""The runner's between-chunk field reduction, restated in NumPy.

TEST INFRASTRUCTURE ONLY (see the package docstring).

``state_diagnostics`` restates ``_state_diagnostics`` (somax/_src/cli/_run.py:
408-439): per field ``np.nanmin`` / ``np.nanmean`` / ``np.nanmax`` (NaN
skipped, +-Inf kept) in FP64, the count of ``~np.isfinite`` values, and
all-NaN statistics when every value is non-finite.  ``flatten`` restates the
multi-element branch of ``_flatten_diagnostics`` (:163-166): ``mean`` /
``max`` / ``min`` with NaN propagating.

Pinned by the reference's own cases (tests/test_cli_run.py:52-110: one NaN
counted as 1 of the field's size; an all-Inf field flagged; a clean field
passes) and by brute force: a Python loop over the values applying IEEE-754
classification directly (tests/test_oracle_run.py).
"""

from __future__ import annotations

import numpy as np


def field_summary(x) -> dict:
    """cli/_run.py:424-438 for one field, in FP64."""
    a = np.asarray(x, dtype=np.float64).ravel()
    bad = int(np.sum(~np.isfinite(a)))
    if bad == a.size:
        nan = float("nan")
        return {"min": nan, "mean": nan, "max": nan, "nan": bad}
    return {"min": float(np.nanmin(a)), "mean": float(np.nanmean(a)), "max": float(np.nanmax(a)), "nan": bad}


def state_diagnostics(fields: dict) -> dict:
    """``{name: field_summary(value)}`` in field order (cli/_run.py:408-439)."""
    return {k: field_summary(v) for k, v in fields.items()}


def flatten(x) -> tuple[float, float, float]:
    """``(mean, max, min)`` of a multi-element field (cli/_run.py:163-166)."""
    a = np.asarray(x, dtype=np.float64)
    return float(a.mean()), float(a.max()), float(a.min())
