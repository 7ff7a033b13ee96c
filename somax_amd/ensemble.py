"""Ensembles of QG models advanced together on the HIP path.

Extension with no reference counterpart: the reference advances one model
per ``SomaxModel.integrate`` call (somax/_src/core/model.py:53-88).  Here any
mix of ``BaroclinicQG`` / ``BarotropicQG`` models -- uneven grid sizes, layer
counts and parameters -- steps as one ensemble: members sharing a grid,
advection and precision form a class whose planes every kernel of the step
covers in one launch; classes run concurrently on their own streams, and the
step is captured as a hipGraph and replayed.  The arithmetic per member is
that of ``model.step``.

With ``merged=True`` the classes are packed into launch groups (those
sharing precision, advection, biharmonic flag and stage tile height, up to
eight classes each): every kernel phase of a Runge-Kutta stage is then one
launch over the whole group, whose blocks a work table assigns to the
classes, instead of one launch per class.

``QGEnsemble.autotune`` picks the launch configuration by measurement: the
transform layout (folded half-size transforms for every even grid, for the
large grids only, or for none), the member packing (one launch per class on
concurrent class streams or on one stream, or merged launch groups), and
then, class by class, the stage kernels' tile height (8 or 16 rows).  Every
candidate is a new ensemble, timed from its first step: launch shapes are
fixed when an ensemble first steps.  Packing, concurrency and tile height
leave every result bit unchanged; the transform layout changes the order of
the transforms' sums (both layouts meet the same parity bars against the
FP64 oracle).
"""

from __future__ import annotations

import ctypes
import math
import os

from . import _lib


def _torch():
    import torch

    return torch


class QGEnsemble:
    """A packed set of QG members (see the module docstring).

    The ensemble state is one float32 device buffer; ``pack`` builds it from
    per-member PV fields, ``member`` / ``unpack`` return views into it, and
    ``step`` advances it in place.
    """

    def __init__(self, models, *, concurrent: bool = True, graph: bool = True, merged: bool = False,
                 stage_rows=None, device="cuda"):
        torch = _torch()
        self.models = list(models)
        if not self.models:
            raise ValueError("an ensemble needs at least one member")
        lib = _lib.load()
        _lib.init_device()
        n = len(self.models)
        descs = (_lib.QGDesc * n)()
        keep = []
        for i, m in enumerate(self.models):
            d, k = m._desc()
            descs[i] = d
            keep.extend(k)
        params = (_lib.QGParams * n)(*[m._c_params() for m in self.models])
        flags = ((_lib.ENS_CONCURRENT if concurrent else 0) | (_lib.ENS_GRAPH if graph else 0)
                 | (_lib.ENS_MERGED if merged else 0))
        self.merged = bool(merged)
        self.concurrent = bool(concurrent)
        ens = ctypes.c_void_p()
        _lib.check(lib.somaxhip_ens_create(descs, params, n, flags, ctypes.byref(ens)), "somaxhip_ens_create")
        self._ens = ens
        self.device = device
        self.state_floats = int(lib.somaxhip_ens_state_floats(ens))
        self.offsets = [int(lib.somaxhip_ens_member_offset(ens, i)) for i in range(n)]
        self.num_classes = int(lib.somaxhip_ens_num_classes(ens))
        self.shapes = [(m._nl, m.grid.Ny, m.grid.Nx) for m in self.models]
        nbytes = int(lib.somaxhip_ens_workspace_bytes(ens))
        self._ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=device)
        if stage_rows is not None:
            rows = [stage_rows] * self.num_classes if isinstance(stage_rows, int) else list(stage_rows)
            if len(rows) != self.num_classes:
                raise ValueError(f"stage_rows: expected {self.num_classes} entries (one per class), got {len(rows)}")
            for c, r in enumerate(rows):
                self.set_stage_rows(c, r)

    def __del__(self):
        ens = getattr(self, "_ens", None)
        if ens is not None:
            try:
                _lib.load().somaxhip_ens_destroy(ens)
            except Exception:
                pass

    def __len__(self) -> int:
        return len(self.models)

    @property
    def layer_cells(self) -> int:
        """Interior layer-cells advanced per step (sum of nl * nx * ny)."""
        return sum(m._nl * m.grid.nx * m.grid.ny for m in self.models)

    # -- packed state ---------------------------------------------------------
    def pack(self, qs):
        """Packed float32 state from per-member PV fields (device tensors or arrays)."""
        torch = _torch()
        if len(qs) != len(self.models):
            raise ValueError(f"expected {len(self.models)} member fields, got {len(qs)}")
        out = torch.zeros(self.state_floats, dtype=torch.float32, device=self.device)
        for i, q in enumerate(qs):
            shape = self.shapes[i]
            t = torch.as_tensor(q, dtype=torch.float32, device=self.device).reshape(-1)
            if t.numel() != math.prod(shape):
                raise ValueError(f"member {i}: field has {t.numel()} values, expected {shape}")
            out[self.offsets[i]:self.offsets[i] + t.numel()] = t
        return out

    def member(self, packed, i: int):
        """View of member ``i``'s PV field ``(nl, Ny, Nx)`` in the packed state."""
        shape = self.shapes[i]
        return packed[self.offsets[i]:self.offsets[i] + math.prod(shape)].view(shape)

    def unpack(self, packed):
        return [self.member(packed, i) for i in range(len(self.models))]

    # -- stepping -------------------------------------------------------------
    def _check_packed(self, packed):
        torch = _torch()
        if not isinstance(packed, torch.Tensor) or packed.dtype != torch.float32 or packed.device.type != "cuda":
            raise TypeError("the packed state must be a float32 device tensor")
        if packed.numel() != self.state_floats or not packed.is_contiguous():
            raise ValueError(f"the packed state must be contiguous with {self.state_floats} values")

    @staticmethod
    def _stream():
        return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)

    def step(self, packed, dt: float, nsteps: int = 1) -> None:
        """Advance every member by ``nsteps`` SSP-RK3 steps of ``dt``, in place."""
        self._check_packed(packed)
        _lib.check(
            _lib.load().somaxhip_ens_step(self._ens, packed.data_ptr(), int(nsteps), float(dt), self._ws.data_ptr(),
                                         self._stream()),
            "somaxhip_ens_step",
        )

    def member_diagnostics(self, packed) -> list[dict]:
        """Per member ``{"min", "mean", "max", "nan"}`` of its PV field, as
        the runner's state diagnostics (NaN-skipping statistics, count of
        non-finite values), reduced on the device in one call per 64
        members: the ensemble's NaN flags."""
        from .run import _summary, field_stats

        self._check_packed(packed)
        views = self.unpack(packed)
        rows = field_stats(views)
        return [_summary(r, v.numel()) for r, v in zip(rows, views)]

    def set_params(self, params) -> None:
        """Replace the members' parameters (one params object per member)."""
        if len(params) != len(self.models):
            raise ValueError(f"expected {len(self.models)} parameter sets, got {len(params)}")
        for m, p in zip(self.models, params):
            m.params = p
        arr = (_lib.QGParams * len(self.models))(*[m._c_params() for m in self.models])
        _lib.check(_lib.load().somaxhip_ens_set_params(self._ens, arr, self._stream()), "somaxhip_ens_set_params")

    # -- configuration ------------------------------------------------------
    def class_info(self) -> list[dict]:
        """Per class (largest grid first): nx, ny, planes, members, fold, stage_rows."""
        lib = _lib.load()
        out = []
        for c in range(self.num_classes):
            v = [ctypes.c_int(0) for _ in range(6)]
            _lib.check(lib.somaxhip_ens_class_info(self._ens, c, *[ctypes.byref(x) for x in v]),
                       "somaxhip_ens_class_info")
            out.append(dict(zip(("nx", "ny", "planes", "members", "fold", "stage_rows"), (x.value for x in v))))
        return out

    def set_stage_rows(self, cls: int, rows: int) -> None:
        """Stage-kernel tile height of class ``cls`` (8 or 16); only before
        the first step."""
        _lib.check(_lib.load().somaxhip_ens_set_stage_rows(self._ens, int(cls), int(rows)),
                   "somaxhip_ens_set_stage_rows")

    def time_step(self, packed, dt: float, steps: int = 5) -> float:
        """Mean milliseconds of one ensemble step (HIP events; one untimed
        step first, which also captures the graph)."""
        torch = _torch()
        self.step(packed, dt, 1)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        self.step(packed, dt, steps)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    @classmethod
    def autotune(cls, models, qs, dt: float, steps: int = 5, graph: bool = True, log=None):
        """Build the fastest configuration of an ensemble of ``models`` for
        steps of ``dt`` from the initial fields ``qs`` (timed on copies).

        Stage 1 times every (transform layout, member packing) pair on a
        new ensemble each, the packings being one launch per class on
        concurrent streams or on one stream, and merged launch groups; stage
        2 tries the other tile height for each class of the winner in turn,
        on a new ensemble each, keeping a change only when the step gets
        faster.  Returns ``(ensemble, report)``; the report lists every
        timing and the chosen configuration."""
        log = log or (lambda msg: None)
        report = {"candidates": []}
        best = None
        saved = os.environ.get("SOMAXHIP_FOLD")

        def fold_env(fold):
            if fold == "default":
                os.environ.pop("SOMAXHIP_FOLD", None)
            else:
                os.environ["SOMAXHIP_FOLD"] = fold

        try:
            for fold in ("default", "1", "0"):
                for conc, merged in ((True, False), (False, False), (True, True)):
                    fold_env(fold)
                    ens = cls(models, concurrent=conc, graph=graph, merged=merged)
                    ms = ens.time_step(ens.pack(qs), dt, steps)
                    report["candidates"].append({"fold": fold, "concurrent": conc, "merged": merged,
                                                 "ms_per_step": round(ms, 4)})
                    log(f"autotune: fold={fold} concurrent={conc} merged={merged}: {ms:.3f} ms/step")
                    if best is None or ms < best[0]:
                        best = (ms, ens, fold, conc, merged)
                    del ens
            ms, ens, fold, conc, merged = best
            best = None
            rows = [info["stage_rows"] for info in ens.class_info()]
            fold_env(fold)
            # the tile height only shapes the Arakawa stage kernel
            tunable = any(getattr(m, "advection", "arakawa") == "arakawa" for m in models)
            for c, info in enumerate(ens.class_info() if tunable else []):
                trial = list(rows)
                trial[c] = 8 if rows[c] == 16 else 16
                cand = cls(models, concurrent=conc, graph=graph, merged=merged, stage_rows=trial)
                t = cand.time_step(cand.pack(qs), dt, steps)
                report["candidates"].append({"class": f"{info['nx']}x{info['ny']}", "stage_rows": trial[c],
                                             "ms_per_step": round(t, 4)})
                log(f"autotune: class {info['nx']}x{info['ny']} stage rows {trial[c]}: {t:.3f} ms/step")
                if t < ms:
                    ms, rows, ens = t, trial, cand
                del cand
        finally:
            if saved is None:
                os.environ.pop("SOMAXHIP_FOLD", None)
            else:
                os.environ["SOMAXHIP_FOLD"] = saved
        report["chosen"] = {"fold": fold, "concurrent": conc, "merged": merged, "ms_per_step": round(ms, 4),
                            "classes": ens.class_info()}
        return ens, report

    def profile(self, packed, dt: float, nsteps: int = 1) -> dict:
        """Eager serial step with every launch timed: {kernel: (launches, ms, flops, bytes)}."""
        self._check_packed(packed)
        stats = (_lib.KernelStat * 32)()
        n = ctypes.c_int(0)
        _lib.check(
            _lib.load().somaxhip_ens_step_profile(self._ens, packed.data_ptr(), int(nsteps), float(dt),
                                                 self._ws.data_ptr(), self._stream(), stats, 32, ctypes.byref(n)),
            "somaxhip_ens_step_profile",
        )
        return {
            stats[i].name.decode(): (stats[i].launches, stats[i].total_ms, stats[i].flops, stats[i].bytes)
            for i in range(n.value)
        }
