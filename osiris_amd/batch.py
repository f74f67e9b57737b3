"""Batched realizations: residual + Jacobian of many dual-porosity grids of
different sizes replayed as ONE hipGraph.

Each realization keeps its own assembly context (osr_asm), device state and
output buffers.  Layout and Jacobian row pointers are computed once, eagerly
(they synchronise, and the pattern depends only on the phase labels); the
property sweep, residual and Jacobian kernels of every realization are then
captured on one stream and replayed with a single launch.  Nothing in the
captured calls allocates, synchronises or records a memset node (workspaces
are created by an eager warm-up run before capture).
"""
from __future__ import annotations

import ctypes as C

from . import runtime as rt


class RealizationBatch:
    def __init__(self, ctx, fluid, cases):
        """cases: sequence of (grid, state) in the synth.Grid / State form."""
        import torch
        self.ctx = ctx
        self.items = []
        for grid, st in cases:
            asm = rt.Assembler(ctx, fluid, grid)
            d = asm.device_state(st)
            _, nrows = asm.layout(d)
            F = torch.empty(nrows, dtype=torch.float64, device="cuda")
            rowptr, cols, vals = asm.jacobian(d)   # eager: pattern, workspaces
            asm.residual(d, out=F)
            self.items.append(dict(asm=asm, d=d, sd=asm._sd(d), F=F, rowptr=rowptr, cols=cols, vals=vals))
        torch.cuda.synchronize()
        self.graph = None

    def _launch_all(self, stream=None):
        L = rt.lib()
        sp = rt._stream_ptr(stream)
        for it in self.items:
            rt._check(L.osr_residual_device(it["asm"].h, C.byref(it["sd"]), C.c_void_p(it["F"].data_ptr()), sp),
                      "osr_residual_device")
            rt._check(L.osr_jacobian_device(it["asm"].h, C.byref(it["sd"]), C.c_void_p(it["rowptr"].data_ptr()),
                                            C.c_void_p(it["cols"].data_ptr()), C.c_void_p(it["vals"].data_ptr()),
                                            sp), "osr_jacobian_device")

    def run_eager(self):
        self._launch_all()
        return [(it["F"], it["rowptr"], it["cols"], it["vals"]) for it in self.items]

    def capture(self):
        import torch
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self._launch_all(s)          # one more warm run on the capture stream
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._launch_all()
        self.graph = g
        return g

    def replay(self):
        if self.graph is None:
            self.capture()
        self.graph.replay()
        return [(it["F"], it["rowptr"], it["cols"], it["vals"]) for it in self.items]
