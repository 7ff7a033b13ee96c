"""Ensembles of QG models advanced together on the HIP path.

Extension with no reference counterpart: the reference advances one model
per ``SomaxModel.integrate`` call (somax/_src/core/model.py:53-88).  Here any
mix of ``BaroclinicQG`` / ``BarotropicQG`` models -- uneven grid sizes, layer
counts and parameters -- steps as one ensemble: members sharing a grid,
advection and precision form a class whose planes every kernel of the step
covers in one launch; classes run concurrently on their own streams, and the
step is captured as a hipGraph and replayed.  The arithmetic per member is
that of ``model.step``.
"""

from __future__ import annotations

import ctypes
import math

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

    def __init__(self, models, *, concurrent: bool = True, graph: bool = True, device="cuda"):
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
        flags = (_lib.ENS_CONCURRENT if concurrent else 0) | (_lib.ENS_GRAPH if graph else 0)
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
