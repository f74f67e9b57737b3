#!/usr/bin/env python3
"""Benchmark of the per-cell Newton hot path on one MI355X.

Workload (the batched-flash configuration): 8 pseudo-components, PR EOS,
2,000,000 independent cells, T ~ U(330, 420) K, z ~ Dirichlet(1.5), p placed
at +/-15 % of each cell's saturation pressure, so convergence is uneven
(single-phase cells stop after 2 passes, two-phase ones after tens to 500).
A step is one flash of every cell (SuccFalsh semantics, FP64) with inputs
already resident in HBM.  `value` is cells/s of the persistent work-queue
form (FP64 passes, slow cells finished by the lane-group kernel); the plain
one-lane-per-cell FP64 form and the mixed FP32/FP64-Newton form are timed
beside it, and the assembly leg times residual + Jacobian per Newton
iteration on the 64x64x8 dual-porosity grid.

Prints ONE JSON line (rank 0).  Only one GPU is used; --gpus > 1 exits.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_CELLS = 2_000_000
NC = 8
FLUID_SEED = 1
CELL_SEED = 7
PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 vector (= FP64 matrix) peak
PEAK_HBM_GBS = 8000.0


def flop_model(nc, iters, rr_iters):
    """Algorithmic flops of the flash: per SS pass two phase evaluations
    E(nc) = 2nc^2 + 10nc + 60 plus 3nc for the K update, and 6nc per
    Rachford-Rice Newton iteration (FMA = 2, transcendental = 1)."""
    E = 2 * nc * nc + 10 * nc + 60
    return float(np.sum(iters.astype(np.float64)) * (2 * E + 3 * nc) +
                 np.sum(rr_iters.astype(np.float64)) * 6 * nc)


def bytes_model(nc, n):
    """Algorithmic HBM bytes of one flash launch: read T, P, z; write x, y, K
    (3 nc doubles), L, Zl, Zg (3 doubles), label (1 byte), iteration
    counters (2 x 4 bytes)."""
    return n * (8 * (nc + 2) + 8 * (3 * nc + 3) + 1 + 8)


def read_traffic(path):
    """HBM traffic per launch from a committed rocprofv3 --pmc summary, if any."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch"):
                    return float(line.split("=")[1])
    except OSError:
        pass
    return None


def assembly_leg(rt, synth, torch, args):
    """Residual + Jacobian assembly per Newton iteration on the 64x64x8
    dual-porosity grid with 6 components (65,536 cells per continuum):
    property sweep, residual rows and the reference-layout CSR Jacobian."""
    fl = synth.fluid_pseudo(6, synth.PR, 3)
    ctx = rt.Context(fl)

    def flash_fn(T, P, z, sw):
        out = rt.flash(ctx, T, P, z, sw=sw, mode="persistent")
        return {k: v.cpu().numpy() for k, v in out.items()}

    grid, st = synth.dual_porosity_case(fl, 64, 64, 8, 0, flash_fn, p_range=(8e6, 20e6))
    asm = rt.Assembler(ctx, fl, grid)
    d = asm.device_state(st)
    off, nrows = asm.layout(d)
    F = torch.empty(nrows, dtype=torch.float64, device="cuda")
    rp, cols, vals = asm.jacobian(d)
    nnz = int(cols.shape[0])
    stream = torch.cuda.current_stream()
    L = rt.lib()
    sd = asm._sd(d)
    import ctypes as C

    def step():
        rt._check(L.osr_residual_device(asm.h, C.byref(sd), C.c_void_p(F.data_ptr()),
                                        rt._stream_ptr(None)), "residual")
        rt._check(L.osr_jacobian_device(asm.h, C.byref(sd), C.c_void_p(rp.data_ptr()),
                                        C.c_void_p(cols.data_ptr()), C.c_void_p(vals.data_ptr()),
                                        rt._stream_ptr(None)), "jacobian")
    for _ in range(2):
        step()
    torch.cuda.synchronize()
    steps = max(3, args.steps // 4)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        a0, a1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a0.record(stream)
        for _ in range(reps):
            fn()
        a1.record(stream)
        torch.cuda.synchronize()
        return a0.elapsed_time(a1) / reps

    # reduced (nc+1)^2 block CSR from this iteration's residual and CSR
    red = {}
    ms_red = timed(lambda: red.update(r=asm.reduced(d, F, rp, vals)), steps)
    nnzb = int(red["r"][1].shape[0])
    # adjoint of the residual w.r.t. log-permeability (cotangent F)
    ms_perm = timed(lambda: asm.perm_vjp(d, F), steps)
    # flash adjoint over the grid's cells (feed: the fracture liquid composition)
    n = grid.n
    zf = torch.as_tensor(st.x_f, device="cuda")
    outf = rt.flash(ctx, grid.T, d["Po_f"], zf)
    ms_adj = timed(lambda: rt.flash_adjoint(ctx, grid.T, d["Po_f"], zf, outf["K"], outf["label"],
                                            xbar=zf, Lbar=d["Po_f"] * 0 + 1.0), steps)
    # algorithmic HBM bytes of one residual+Jacobian: state read, F write,
    # CSR write (8 B value + 4 B column) -- properties stay on chip/L2
    nc = fl.nc
    state_b = 2 * n * (1 + 8 * (2 + 2 * nc + 3))
    out_b = 8 * nrows + 12 * nnz + 8 * (nrows + 1)
    return {
        "workload": "64x64x8 dual-porosity grid, 6 components PR, property sweep + residual + "
                    "reference-layout CSR Jacobian (forward mode) per Newton iteration",
        "cells": n, "blocks": 2 * n, "rows": int(nrows), "nnz": nnz,
        "ms_per_newton_iter": ms,
        "cells_per_s": n / (ms * 1e-3),
        "algorithmic_bytes": state_b + out_b,
        "achieved_GBs": (state_b + out_b) / (ms * 1e-3) / 1e9,
        "hbm_peak_GBs": PEAK_HBM_GBS,
        "reduced_bsr": {"ms": ms_red, "nnzb": nnzb, "block": f"{nc + 1}x{nc + 1}"},
        "perm_adjoint_ms": ms_perm,
        "flash_adjoint": {"ms": ms_adj, "cells": n},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=N_CELLS)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-assembly", action="store_true", help="skip the assembly leg")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="after the timed steps, copy a tensor of this many MiB once (a "
                         "kernel of known HBM traffic that scales rocprofv3 FETCH_SIZE / "
                         "WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from osiris_amd import runtime as rt, synth

    dev = torch.device("cuda:0")
    fl = synth.fluid_pseudo(NC, synth.PR, FLUID_SEED)
    ctx = rt.Context(fl)
    n = args.cells
    cells = synth.flash_cells(fl, n, CELL_SEED, rt.psat_numpy(ctx))
    T = torch.tensor(cells.T, device=dev)
    P = torch.tensor(cells.P, device=dev)
    z = torch.tensor(cells.z, device=dev)
    stream = torch.cuda.current_stream()

    def run(mode, out, steps, warmup):
        for _ in range(warmup):
            rt.flash(ctx, T, P, z, mode=mode, out=out)
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(steps)]
        t0 = time.perf_counter()
        for s, e in ev:
            s.record(stream)
            rt.flash(ctx, T, P, z, mode=mode, out=out)
            e.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kern = float(np.mean([s.elapsed_time(e) for s, e in ev])) * 1e-3
        return wall / steps, kern

    out = rt.alloc_flash_out(NC, n)
    sec_persist, kern_persist = run("persistent", out, args.steps, args.warmup)
    iters = out["iters"].cpu().numpy()
    rr = out["rr_iters"].cpu().numpy()
    labels = out["label"].cpu().numpy()
    gpu_res = {k: out[k].cpu().numpy() for k in ("x", "y", "K", "L", "label", "iters")}

    out_plain = rt.alloc_flash_out(NC, n)
    sec_plain, kern_plain = run("plain", out_plain, max(3, args.steps // 4), 1)
    out_mixed = rt.alloc_flash_out(NC, n)
    sec_mixed, kern_mixed = run("mixed", out_mixed, args.steps, 1)
    mixed_res = {k: out_mixed[k].cpu().numpy() for k in ("x", "y", "K", "L", "label", "iters")}

    if args.calib_copy_mb > 0:
        src = torch.ones(args.calib_copy_mb * 2**20 // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    flops = flop_model(NC, iters, rr)
    achieved = flops / kern_persist / 1e12
    cells_per_s = n / sec_persist

    result = {
        "metric": "flash+assembly cells/s per Newton iter, plain FP64 vs mixed/analytic @<=1e-8 err",
        "value": cells_per_s,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec_persist * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded PCG64 generator; no reference fixtures exist)",
        "config": {
            "workload": "batched flash with uneven convergence: 8 pseudo-components, PR, "
                        f"{n} independent cells, p within +/-15% of Psat, FP64 persistent "
                        "work-queue kernel",
            "n_cells": n, "nc": NC, "eos": "PR", "parallelism": "single GPU",
        },
        "paths": {
            "persistent_fp64": {"cells_per_s": cells_per_s, "kernel_ms": kern_persist * 1e3},
            "plain_fp64": {"cells_per_s": n / sec_plain, "kernel_ms": kern_plain * 1e3},
            "mixed_fp32_fp64": {"cells_per_s": n / sec_mixed, "kernel_ms": kern_mixed * 1e3},
            "speedup_persistent_vs_plain": sec_plain / sec_persist,
        },
        "workload_stats": {
            "mean_ss_passes": float(iters.mean()), "max_ss_passes": int(iters.max()),
            "frac_two_phase": float((labels == 0).mean()),
            "frac_capped_500": float((iters >= 500).mean()),
        },
        "roofline": {
            "bound": "mfma",
            "bound_note": "FP64 vector ALU; MI355X FP64 vector and FP64 matrix peaks are both 78.6 TFLOP/s",
            "kernel": f"k_flash_persistent<{NC}> + k_flash_group<{NC}, 8> (one flash step)",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "flops_per_launch": flops,
            "avg_launch_ms": kern_persist * 1e3,
            "algorithmic_bytes_per_launch": bytes_model(NC, n),
            "traffic": read_traffic(os.path.join(ROOT, "profiles", "flash_persistent_pr8_2M_pmc.txt")),
        },
    }

    if not args.no_assembly:
        result["assembly"] = assembly_leg(rt, synth, torch, args)

    if not args.no_cpu:
        # CPU baseline: the oracle (FP64 C restatement of SuccFalsh) on the
        # same 2M-cell step, all host threads up to 16, plus a 1-thread run
        # on a 100k-cell sample.  Parity of the timed GPU step against it.
        from oracle import oracle as O
        threads = min(16, os.cpu_count() or 1)
        t0 = time.perf_counter()
        ref = O.flash_batch(fl, cells.T, cells.P, cells.z, nthreads=threads)
        cpu_s = time.perf_counter() - t0
        m = 100_000 if n > 100_000 else n
        t0 = time.perf_counter()
        O.flash_batch(fl, cells.T[:m], cells.P[:m], cells.z[:, :m], nthreads=1)
        cpu1_s = time.perf_counter() - t0
        def parity(res):
            conv = (ref["iters"] < 500) & (res["iters"] < 500)
            errs = {}
            for k in ("x", "y", "K", "L"):
                fin = np.isfinite(ref[k]) & np.isfinite(res[k])
                e = np.abs(res[k] - ref[k]) / np.maximum(np.abs(ref[k]), 1e-12)
                e = np.where(fin, e, 0.0)
                errs[k] = float(e[..., conv].max())
            return {"label_mismatches": int((res["label"] != ref["label"]).sum()),
                    "max_rel_err_converged_cells": errs,
                    "cells_capped_either_side": int((~conv).sum())}
        result["cpu_baseline"] = {
            "value": n / cpu_s, "unit": "cells/s", "cores": threads, "kind": "port",
            "sample": f"full {n}-cell step, FP64 C oracle with OpenMP over cells",
            "single_core_cells_per_s": m / cpu1_s,
            "single_core_sample": f"first {m} cells, 1 thread",
        }
        result["parity_vs_oracle"] = {"persistent_fp64": parity(gpu_res),
                                      "mixed_fp32_fp64": parity(mixed_res)}
    print(json.dumps(result))


if __name__ == "__main__":
    main()
