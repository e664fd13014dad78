"""Operators on grids of different sizes applied together from one hipGraph.

No counterpart in the reference (one operator and one field per call,
src/em3d/operator.py:353-363).  The group keeps one input and one output
buffer per member; the first ``run`` captures every member's five FFT passes
into a graph (libem3d_hip ``em3d_group_*``) and every call replays it with a
single launch.  Kernel variants are chosen (``autotune``) before the capture,
because the graph keeps the kernels it recorded.
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from .operator import HipPlan


class PackedOperators:
    def __init__(self, operators, *, autotune: bool = True):
        import torch
        if not operators:
            raise ValueError("PackedOperators needs at least one operator")
        be = operators[0].backend
        for op in operators:
            if op.backend.precision is not be.precision or op.backend.device_index != be.device_index:
                raise ValueError("all operators of a group share one backend precision and device")
        self.operators = list(operators)
        self.backend = be
        self.tuned = [op.autotune(reps=3) if autotune else None for op in self.operators]
        self.inputs, self.outputs, keep = [], [], []
        plans, khats, chis, kinds, ins, outs, works = [], [], [], [], [], [], []
        for op in self.operators:
            plan = HipPlan.get(op._N, op._ncomp, 1, be.precision, be.device_index)
            khat, kind, chi = op._spectrum()
            u = torch.zeros(op._field_shape(), dtype=be.precision.torch_complex_dtype, device=be.torch_device)
            y = torch.zeros_like(u)
            w = op._workspace(plan, private=True)   # members run as concurrent graph branches
            self.inputs.append(u)
            self.outputs.append(y)
            keep += [plan, khat, chi, w]
            plans.append(plan.handle.value)
            khats.append(khat.data_ptr())
            chis.append(chi.data_ptr() if chi is not None else None)
            kinds.append(kind)
            ins.append(u.data_ptr())
            outs.append(y.data_ptr())
            works.append(w.data_ptr())
        n = len(self.operators)
        arr = lambda vals: (C.c_void_p * n)(*vals)
        self._keep = keep
        h = C.c_void_p()
        _lib.check(_lib.lib().em3d_group_create(n, arr(plans), arr(khats), arr(chis), (C.c_int * n)(*kinds),
                                                arr(ins), arr(outs), arr(works), C.byref(h)), "em3d_group_create")
        self.handle = h

    def run(self):
        """outputs[i] = A_i inputs[i] for every member (one graph launch)."""
        _lib.check(_lib.lib().em3d_group_matvec(self.handle, _lib.stream_handle(self.backend.torch_device)),
                   "em3d_group_matvec")
        return self.outputs

    def matvec(self, fields):
        """Fresh arrays A_i fields[i]."""
        if len(fields) != len(self.operators):
            raise ValueError(f"expected {len(self.operators)} fields, got {len(fields)}")
        for buf, f, op in zip(self.inputs, fields, self.operators):
            t, batch = op._prepare_input(f)
            if batch != 1 or t.dim() != buf.dim():
                raise ValueError("each field must have its operator's field shape")
            buf.copy_(t)
        return [y.clone() for y in self.run()]

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().em3d_group_destroy(h)
            except Exception:
                pass
            self.handle = None
