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

import numpy as np

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


class NewtonBatch:
    """One Newton-iteration evaluation of many dual-porosity realizations of
    different grid sizes, replayed as ONE hipGraph:

      1. the flash of every cell of every grid, both continua, as one batch
         (one osr_flash_device call; its pressures alias the grids' oil
         pressures, its feeds are the cells' overall compositions);
      2. per grid, the property sweep and residual;
      3. per grid and continuum, the phase-appearance test (PhaseCtrl with its
         saturation-pressure runs) on scratch copies of labels, compositions
         and saturations, so the state being linearised stays as given;
      4. per grid, the Jacobian: the reference-layout CSR (form="csr") or
         the reduced (nc+1)^2 block CSR assembled straight from the state
         (form="reduced").

    The Jacobian pattern follows the phase labels, so layouts and row
    pointers are computed once, eagerly, and an eager warm-up creates every
    workspace before capture.  All grids share one Jacobian workspace sized
    for the largest (their calls are serialised on the capture stream)."""

    FORMS = ("csr", "reduced")

    def __init__(self, ctx, fluid, cases, form="reduced", flash_mode="mixed", phase_test=True):
        """cases: sequence of (grid, state) in the synth.Grid / State form,
        states carrying their feeds z_m, z_f; all grids at one temperature."""
        import torch
        if form not in self.FORMS:
            raise ValueError(f"form must be one of {self.FORMS}")
        temps = {float(g.T) for g, _ in cases}
        if len(temps) != 1:
            raise ValueError("all grids of a batch must share one temperature")
        self.T = temps.pop()
        self.ctx, self.nc, self.form, self.flash_mode = ctx, ctx.nc, form, flash_mode
        self.phase_test = phase_test
        nc = self.nc
        sizes = [g.n for g, _ in cases]
        self.ncells = 2 * sum(sizes)
        f64 = dict(dtype=torch.float64, device="cuda")
        # flash batch: the grids' oil pressures are views into P_all
        self.P_all = torch.empty(self.ncells, **f64)
        self.z_all = torch.empty(nc, self.ncells, **f64)
        self.fout = rt.alloc_flash_out(nc, self.ncells, full=False)
        L = rt.lib()
        rt._bind_asm(L)
        if not getattr(L, "_pc_bound", False):
            L.osr_phase_ctrl_device.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_double] + [C.c_void_p] * 8
            L._pc_bound = True
        L.osr_asm_props_table.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        self.items = []
        off = 0
        for (grid, st), n in zip(cases, sizes):
            if st.z_m is None or st.z_f is None:
                raise ValueError("states must carry their feeds z_m, z_f")
            asm = rt.Assembler(ctx, fluid, grid)
            d = asm.device_state(st)
            for k, (po, z) in enumerate(((d["Po_m"], st.z_m), (d["Po_f"], st.z_f))):
                sl = slice(off + k * n, off + (k + 1) * n)
                self.P_all[sl].copy_(po)
                self.z_all[:, sl].copy_(torch.as_tensor(np.ascontiguousarray(z)))
                d["Po_m" if k == 0 else "Po_f"] = self.P_all[sl]
            off += 2 * n
            _, nrows = asm.layout(d)
            it = dict(asm=asm, d=d, sd=asm._sd(d), n=n, nrows=nrows,
                      F=torch.empty(nrows, **f64))
            if form == "csr":
                rp = torch.empty(nrows + 1, dtype=torch.int64, device="cuda")
                nnz = C.c_int64()
                rt._check(L.osr_jacobian_rowptr_device(asm.h, C.byref(it["sd"]), C.c_void_p(rp.data_ptr()),
                                                       C.byref(nnz), rt._stream_ptr(None)),
                          "osr_jacobian_rowptr_device")
                it.update(rowptr=rp, cols=torch.empty(nnz.value, dtype=torch.int32, device="cuda"),
                          vals=torch.empty(nnz.value, **f64))
            else:
                it["red"] = asm._reduced_out(d, None)
            if phase_test:
                props = C.c_void_p()
                rt._check(L.osr_asm_props_table(asm.h, C.byref(props)), "osr_asm_props_table")
                it["props"] = props.value
                it["scratch"] = [dict(label=torch.empty(n, dtype=torch.int8, device="cuda"),
                                      x=torch.empty(nc, n, **f64), y=torch.empty(nc, n, **f64),
                                      S=torch.empty(3, n, **f64)) for _ in range(2)]
                it["psat"] = torch.empty(2, n, **f64)
            self.items.append(it)
        ws = max(it["asm"].workspace_bytes() for it in self.items)
        self.ws = torch.empty(ws, dtype=torch.uint8, device="cuda")
        for it in self.items:
            it["asm"].set_workspace(self.ws)
        self.run_eager()            # creates every remaining workspace
        torch.cuda.synchronize()
        self.graph = None

    def nnz(self):
        """Stored Jacobian entries of one evaluation (scalars)."""
        if self.form == "csr":
            return sum(int(it["cols"].shape[0]) for it in self.items)
        return sum(int(it["red"][1].shape[0]) * (self.nc + 1) ** 2 for it in self.items)

    def _launch_all(self, stream=None):
        L = rt.lib()
        sp = rt._stream_ptr(stream)
        rt.flash(self.ctx, self.T, self.P_all, self.z_all, mode=self.flash_mode, out=self.fout, stream=stream)
        for it in self.items:
            asm, sd, n = it["asm"], it["sd"], it["n"]
            rt._check(L.osr_residual_device(asm.h, C.byref(sd), C.c_void_p(it["F"].data_ptr()), sp),
                      "osr_residual_device")
            if self.phase_test:
                d = it["d"]
                for k, sfx in enumerate(("_m", "_f")):
                    s = it["scratch"][k]
                    s["label"].copy_(d["label" + sfx])
                    s["x"].copy_(d["x" + sfx])
                    s["y"].copy_(d["y" + sfx])
                    s["S"].copy_(d["S" + sfx])
                    pg = it["props"] + 8 * (3 * 2 * n + k * n)       # field 3 (Pg), continuum k
                    rt._check(L.osr_phase_ctrl_device(
                        self.ctx.h, n, None, self.T, C.c_void_p(s["label"].data_ptr()),
                        C.c_void_p(d["Po" + sfx].data_ptr()), C.c_void_p(pg),
                        C.c_void_p(s["x"].data_ptr()), C.c_void_p(s["y"].data_ptr()),
                        C.c_void_p(s["S"].data_ptr()), C.c_void_p(it["psat"][k].data_ptr()), sp),
                        "osr_phase_ctrl_device")
            if self.form == "csr":
                rt._check(L.osr_jacobian_device(asm.h, C.byref(sd), C.c_void_p(it["rowptr"].data_ptr()),
                                                C.c_void_p(it["cols"].data_ptr()),
                                                C.c_void_p(it["vals"].data_ptr()), sp), "osr_jacobian_device")
            else:
                asm.reduced_direct(it["d"], it["F"], out=it["red"], stream=stream)

    def run_eager(self):
        self._launch_all()

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
        # the graph holds the context's flash, psat and phase-control
        # workspaces: pin them until the graph is dropped
        self.ctx.freeze(self)
        return g

    def replay(self):
        if self.graph is None:
            self.capture()
        self.graph.replay()

    def close(self):
        if self.graph is not None:
            self.ctx.unfreeze(self)
        self.graph = None
        for it in self.items:
            it["asm"].close()
        self.items = []
