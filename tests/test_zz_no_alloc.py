"""The no-allocation QG step (somaxhip_qg_step_noalloc, C ABI) against the
step with a caller workspace (somaxhip_qg_step): same kernels on the same
inputs, so the results must agree bit for bit.  Two calls in a row, the
second on a smaller grid (the shared scratch array keeps the first call's
contents), plus psi_out, and the refusal of a plan whose workspace exceeds
the array."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


def _noalloc(model, q, psi, dt, nsteps):
    import torch

    from somax_amd import _lib

    lib = _lib.load()
    prm = model._c_params()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return lib.somaxhip_qg_step_noalloc(model._ensure_plan(), q.data_ptr(), psi.data_ptr() if psi is not None else None,
                                        int(nsteps), float(dt), ctypes.byref(prm), stream)


def test_noalloc_step_equals_workspace_step(gpu):
    import torch

    from somax_amd import _lib
    from somax_amd import testcases as tc

    dt = 1200.0
    for n, precision in ((256, "mixed"), (64, "fp32")):
        _log(f"[noalloc] {n}^2 {precision}, WENO-Z + mass constraint, 3 steps")
        model = tc.config_baroclinic(n, precision=precision, advection="weno", mass_correction=True)
        q0 = torch.from_numpy(tc.random_pv(model, seed=n, zero_mean=True).astype(np.float32)).cuda()
        a, b = q0.clone(), q0.clone()
        pa, pb = torch.empty_like(q0), torch.full_like(q0, float("nan"))
        plan = model._ensure_plan()
        prm = model._c_params()
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(_lib.load().somaxhip_qg_step(plan, a.data_ptr(), pa.data_ptr(), 3, dt, ctypes.byref(prm),
                                                model._ws.data_ptr(), stream), "somaxhip_qg_step")
        _lib.check(_noalloc(model, b, pb, dt, 3), "somaxhip_qg_step_noalloc")
        torch.cuda.synchronize()
        assert torch.isfinite(b).all()
        assert torch.equal(a, b), f"{n}^2: max |diff| {float((a - b).abs().max()):.3e}"
        assert torch.equal(pa, pb)


def test_noalloc_refuses_a_plan_beyond_the_scratch_array(gpu):
    import torch

    from somax_amd import _lib
    from somax_amd import testcases as tc

    _log("[noalloc] 1024^2 plan: workspace beyond the scratch array is refused")
    model = tc.config_baroclinic(1024, precision="mixed")
    plan = model._ensure_plan()
    assert _lib.load().somaxhip_qg_workspace_bytes(plan) > _lib.NOALLOC_SCRATCH_BYTES
    q = torch.zeros((3, 1026, 1026), device="cuda")
    assert _noalloc(model, q, None, 1200.0, 1) == 1  # SOMAXHIP_ERR_ARG, nothing queued
    assert "scratch" in _lib.last_error()
    torch.cuda.synchronize()
    assert not q.any()
