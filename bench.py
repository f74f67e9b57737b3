#!/usr/bin/env python3
"""Benchmark of the per-cell Newton hot path on one MI355X.

Headline workload (the batched-realization configuration): 16 dual-porosity
grids with nx, ny ~ U{32..160}, nz ~ U{8..32} plus one 256x256x32 grid, 10
pseudo-components (PR), rock as the 10x10x1 workload, p ~ U(8, 20) MPa,
state = flash of (p, z ~ Dirichlet(2)).  A step is one Newton-iteration
evaluation of every matrix and fracture cell of every grid, replayed as one
hipGraph: batched flash of every cell, property sweep + residual, the
phase-appearance test (PhaseCtrl) and the Jacobian.  `value` is cells/s
(both continua counted) of the analytic path -- persistent work-queue FP64
flash and the reduced (nc+1)^2 block system assembled straight from the
state by forward mode; the plain FP64 path -- one-lane-per-cell FP64 flash
and the reference-layout CSR Jacobian -- and the analytic path with the
mixed FP32/FP64 flash are timed beside it on the 256x256x32 grid
(4,194,304 cells).

Secondary legs: the batched-flash configuration (8 pseudo-components, 2M
independent cells at +/-15 % of Psat; plain, persistent and mixed forms,
parity against the oracle), the mixed-precision configuration (12
pseudo-components, SRK, 4M independent cells; plain, persistent and mixed
forms, K-values, phase fractions and ln-fugacities per cell against the
oracle) and the 64x64x8 differentiation configuration (permeability
adjoint, flash adjoint).  Each persistent or mixed flash runs under the
launch knobs the autotuner (osiris_amd/tune.py) picks on that workload's own
cells; the knobs are reported in the line.  The CPU baseline is the oracle (FP64
C/C++ restatement) on bounded samples of the same workloads.

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


# VALU issue ceiling: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 VALU
# instruction per SIMD every 4 cycles
PEAK_VALU_WAVE_INSTR_PER_S = 256 * 4 * 2.4e9 / 4


def valu_view(path, ms):
    """VALU issue rate and FP64 rate of a launch from a committed rocprofv3
    --pmc summary (wave64 VALU instructions and FP64 flops per launch) and
    the launch time measured here."""
    got = {}
    try:
        with open(path) as f:
            for line in f:
                if "=" in line and line.split("=")[0] in ("valu_wave_instructions_per_launch", "fp64_flop_per_launch"):
                    k, v = line.strip().split("=")
                    got[k] = float(v)
    except OSError:
        return None
    if len(got) < 2:
        return None
    ins = got["valu_wave_instructions_per_launch"] / (ms * 1e-3)
    fl = got["fp64_flop_per_launch"] / (ms * 1e-3) / 1e12
    return {"wave_instructions_per_launch": got["valu_wave_instructions_per_launch"],
            "issue_rate_per_s": ins, "issue_peak_per_s": PEAK_VALU_WAVE_INSTR_PER_S,
            "issue_frac": ins / PEAK_VALU_WAVE_INSTR_PER_S,
            "fp64_flop_per_launch": got["fp64_flop_per_launch"], "fp64_tflops": fl,
            "fp64_peak_tflops": PEAK_FP64_TFLOPS, "fp64_frac": fl / PEAK_FP64_TFLOPS}


def assembly_leg(rt, synth, torch, args):
    """The differentiation workload: 64x64x8 dual-porosity grid with 6
    components (65,536 cells per continuum): residual + reference-layout CSR
    Jacobian per Newton iteration, the reduced system (from the CSR and
    straight from the state), the permeability adjoint and the flash
    adjoint."""
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
    # the same reduced system straight from the state (no CSR)
    ms_red_direct = timed(lambda: asm.reduced_direct(d, F, out=red["r"]), steps)
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
        "reduced_bsr": {"from_csr_ms": ms_red, "direct_ms": ms_red_direct, "nnzb": nnzb,
                        "block": f"{nc + 1}x{nc + 1}"},
        "perm_adjoint_ms": ms_perm,
        "flash_adjoint": {"ms": ms_adj, "cells": n},
    }


NEWTON_NC = 10
NEWTON_FLUID_SEED = 11
NEWTON_SEED = 4
NEWTON_BIG = (256, 256, 32)
# Flash of the analytic path: the persistent FP64 form.  Nearly every cell of
# this workload is two-phase and needs more than three passes, so the mixed
# form sends almost all of them through its FP64 Newton polish, which costs
# more than the FP64 passes it replaces (both are timed).
FAST_FLASH = "persistent"
KSP_ITS = 20   # BiCGStab iterations timed on the reduced system of the 256x256x32 grid
STEP_DT = 1.0       # time step (s) of the device time-step measurement
STEP_RTOL = 1e-8    # its BiCGStab tolerance
STEP_MAXIT = 300


def newton_shapes(seed):
    """Grid shapes of the batched-realization workload: 16 grids with
    nx, ny ~ U{32..160}, nz ~ U{8..32}, then the 256x256x32 grid."""
    g = np.random.Generator(np.random.PCG64(seed))
    shapes = [(int(g.integers(32, 161)), int(g.integers(32, 161)), int(g.integers(8, 33))) for _ in range(16)]
    return shapes + [NEWTON_BIG]


def reduced_bytes_model(nc, nblocks, nnzb):
    """Algorithmic HBM bytes of one reduced-system assembly: read the state
    (label, P, x, y, S: 8(2nc+4)+1 per block) and the static cell data
    (porosity, 3 permeabilities, 6 face kA/h: 80 per block); write the
    reduced blocks ((nc+1)^2 doubles + a 4-byte column each), the reduced
    right-hand side (nc+1 per block) and the elimination factors S, t
    ((nc+3)(nc+2) per block)."""
    np1 = nc + 1
    return (nblocks * (8 * (2 * nc + 4) + 1 + 80) + nnzb * (8 * np1 * np1 + 4) +
            nblocks * 8 * (np1 + (nc + 3) * (nc + 2)))


def time_graph(batch, torch, steps, warmup):
    """Capture the batch, replay `warmup` times, then time `steps` replays
    (HIP events on the replay stream, and the host clock)."""
    batch.capture()
    for _ in range(warmup):
        batch.replay()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        batch.replay()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    return e0.elapsed_time(e1) / steps, wall * 1e3


def newton_leg(rt, synth, torch, args, log):
    """Flash + assembly per Newton iteration on the batched-realization
    workload: the plain FP64 path and the mixed/analytic path on the
    256x256x32 grid, then the mixed/analytic path on all 17 grids in one
    graph (the headline)."""
    from osiris_amd.batch import NewtonBatch
    fl = synth.fluid_pseudo(NEWTON_NC, synth.PR, NEWTON_FLUID_SEED)
    ctx = rt.Context(fl)

    def flash_fn(T, P, z, sw):
        out = rt.flash(ctx, T, P, z, sw=sw, mode="persistent")
        return {k: v.cpu().numpy() for k, v in out.items()}

    shapes = newton_shapes(NEWTON_SEED)
    if args.newton_scale < 1:
        f = args.newton_scale
        shapes = [(max(4, int(nx * f)), max(4, int(ny * f)), max(2, int(nz * f))) for nx, ny, nz in shapes]
    t0 = time.perf_counter()
    cases = [synth.dual_porosity_case(fl, nx, ny, nz, 100 + i, flash_fn, p_range=(8e6, 20e6))
             for i, (nx, ny, nz) in enumerate(shapes)]
    log(f"newton: {len(cases)} grids generated in {time.perf_counter() - t0:.1f} s")
    big = cases[-1]
    nbig = 2 * big[0].n
    ncells = sum(2 * g.n for g, _ in cases)
    steps_big = max(2, args.steps // 4)

    # plain FP64 path on the 256x256x32 grid: one-lane-per-cell FP64 flash,
    # reference-layout CSR Jacobian
    b = NewtonBatch(ctx, fl, [big], form="csr", flash_mode="plain")
    ms_plain, wall_plain = time_graph(b, torch, steps_big, 1)
    nnz_plain = b.nnz()
    b.close()
    del b
    torch.cuda.empty_cache()
    log(f"newton: plain {ms_plain:.1f} ms on the big grid")

    # analytic path with the mixed FP32/FP64 flash on the same grid, its
    # launch knobs chosen by the autotuner on this grid's cells
    from osiris_amd import tune
    gb, sb = big
    # launch knobs of the reduced-system assembly, tuned on this grid's state
    # before any reduced-form graph is captured (a captured launch keeps its shape)
    asm_t = rt.Assembler(ctx, fl, gb)
    d_t = asm_t.device_state(sb)
    asm_t.layout(d_t)
    F_t = asm_t.residual(d_t)
    tuned_a = tune.tune_assembly(ctx, asm_t, d_t, F_t, reps=3)
    tune.apply(ctx, tuned_a)
    asm_t.close()
    del asm_t, d_t, F_t
    torch.cuda.empty_cache()
    log(f"newton: reduced-assembly knobs {tuned_a['config']} ({tuned_a['ms']:.1f} ms)")
    Pb = torch.tensor(np.concatenate([sb.Po_m, sb.Po_f]), device="cuda")
    zb = torch.tensor(np.concatenate([sb.z_m, sb.z_f], axis=1), device="cuda")
    tuned = tune.tune_flash(ctx, gb.T, Pb, zb, mode="mixed", reps=3)
    tune.apply(ctx, tuned)
    rt.flash(ctx, gb.T, Pb, zb, mode="mixed")
    mixed_stats = rt.flash_stats(ctx)
    b = NewtonBatch(ctx, fl, [big], form="reduced", flash_mode="mixed")
    ms_mixedflash_big, _ = time_graph(b, torch, steps_big, 1)
    b.close()
    del b
    # the persistent FP64 flash of the analytic path: its knobs tuned on the
    # same cells (kept for every later persistent run of this leg)
    tuned_p = tune.tune_flash(ctx, gb.T, Pb, zb, mode="persistent", reps=3)
    tune.apply(ctx, tuned_p)
    del Pb, zb
    torch.cuda.empty_cache()
    log(f"newton: analytic + mixed flash {ms_mixedflash_big:.1f} ms on the big grid ({tuned['config']})")

    # analytic path with the persistent FP64 flash, plus the reduced assembly alone
    b = NewtonBatch(ctx, fl, [big], form="reduced", flash_mode=FAST_FLASH)
    ms_mixed_big, wall_mixed_big = time_graph(b, torch, steps_big, 1)
    it = b.items[0]
    stream = torch.cuda.current_stream()
    reps = max(3, steps_big)
    it["asm"].reduced_direct(it["d"], it["F"], out=it["red"])
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        it["asm"].reduced_direct(it["d"], it["F"], out=it["red"])
    e1.record(stream)
    torch.cuda.synchronize()
    ms_red = e0.elapsed_time(e1) / reps
    nnzb_big = int(it["red"][1].shape[0])
    # linear solve of the reduced system (BiCGStab + block Jacobi, the
    # reference's KSP class): setup and a fixed number of iterations
    browptr, bcols, bvals, brhs = it["red"][:4]
    ksp = rt.Ksp(NEWTON_NC + 1, nbig)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nsing = ksp.setup(browptr, bcols, bvals)
    ms_ksp_setup = (time.perf_counter() - t0) * 1e3
    rhs = brhs.reshape(-1).contiguous()
    # throughput of the Krylov machinery: a fixed iteration count, with the
    # divergence test off (block Jacobi does not converge on this system;
    # the reason of a default-tolerance solve is reported beside)
    _, kdef = ksp.solve(rhs, rtol=1e-6, atol=0.0, maxit=200)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, kinfo = ksp.solve(rhs, rtol=1e-300, atol=0.0, dtol=1e300, maxit=KSP_ITS)
    torch.cuda.synchronize()
    ms_ksp_it = (time.perf_counter() - t0) * 1e3 / max(1, kinfo["iterations"])
    ksp.close()
    del ksp
    b.close()
    del b, it
    torch.cuda.empty_cache()
    log(f"newton: analytic {ms_mixed_big:.1f} ms on the big grid, reduced assembly {ms_red:.1f} ms")

    # one whole time step on the device (SolveNonLinear + UpdateAll(X, true)):
    # assembly, BiCGStab, acceptance, UpdateAll, PhaseCtrl, PCoeff, PJac and
    # the new sweep, on the 256x256x32 grid with a 1 s step (block Jacobi
    # converges on the accumulation-dominated system)
    import copy as _copy
    grid_s = _copy.copy(big[0])
    grid_s.dt = STEP_DT
    asm_s = rt.Assembler(ctx, fl, grid_s)
    stepper = rt.NewtonStep(asm_s)
    step_runs = []
    for _ in range(2):      # the first allocates the reduced blocks
        d_s = asm_s.device_state(big[1])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sres = stepper.step(d_s, rtol=STEP_RTOL, maxit=STEP_MAXIT)
        torch.cuda.synchronize()
        step_runs.append(((time.perf_counter() - t0) * 1e3, sres))
        del d_s
    ms_step, sres = step_runs[-1]
    stepper.close()
    asm_s.close()
    del stepper, asm_s
    torch.cuda.empty_cache()
    log(f"newton: device time step {ms_step:.1f} ms on the big grid ({sres['ksp']['iterations']} BiCGStab "
        f"iterations, {sres['ksp']['reason']}, accepted {sres['accepted']})")

    # the restart file of the 256x256x32 state (WriteRST / ReadRST byte image)
    asm_r = rt.Assembler(ctx, fl, big[0])
    d_r = asm_r.device_state(big[1])
    asm_r.layout(d_r)
    props_r = asm_r.props(d_r)
    img = rt.rst_write(asm_r, d_r, props_r)
    rt.rst_read(asm_r, d_r, img)
    reps = 5
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        rt.rst_write(asm_r, d_r, props_r)
    e1.record(stream)
    torch.cuda.synchronize()
    ms_rst_w = e0.elapsed_time(e1) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        rt.rst_read(asm_r, d_r, img)       # synchronises for the scalars
    ms_rst_r = (time.perf_counter() - t0) * 1e3 / reps
    rst_b = int(img.numel())
    asm_r.close()
    del asm_r, d_r, props_r, img
    torch.cuda.empty_cache()
    log(f"newton: restart image {rst_b / 1e9:.2f} GB written in {ms_rst_w:.2f} ms, read in {ms_rst_r:.2f} ms")

    # all grids in one graph (headline)
    b = NewtonBatch(ctx, fl, cases, form="reduced", flash_mode=FAST_FLASH)
    ms_all, wall_all = time_graph(b, torch, args.steps, args.warmup)
    flash_stats = rt.flash_stats(ctx)
    labels = b.fout["label"].cpu().numpy()
    nnzb_all = sum(int(it["red"][1].shape[0]) for it in b.items)
    # full-size property: a replay equals an eager evaluation bit for bit
    # (deterministic, atomic-free kernels), every grid and the batched flash.
    # Checksum on the device over the raw bits (int64 sums wrap mod 2^64,
    # so the order of summation does not matter): the plain sum and a sum
    # weighted by an odd multiplier per position, which changes whenever
    # any single element changes.
    def bits_checksum(t):
        flat = t.reshape(-1)
        if flat.dtype == torch.float64:
            flat = flat.view(torch.int64)
        s1 = torch.zeros((), dtype=torch.int64, device=flat.device)
        s2 = torch.zeros_like(s1)
        ch = 1 << 27
        for i in range(0, flat.numel(), ch):
            c = flat[i:i + ch].to(torch.int64)
            w = torch.arange(i, i + c.numel(), device=c.device, dtype=torch.int64) * 5308871522 + 1
            s1 += c.sum()
            s2 += (c * w).sum()
        return int(s1.item()), int(s2.item())

    def digest():
        return [bits_checksum(t) for t in [b.fout[k] for k in ("x", "y", "L", "label")] +
                [x for it in b.items for x in (it["F"], it["psat"], *it["red"])]]
    b.replay()
    torch.cuda.synchronize()
    d_replay = digest()
    b.run_eager()
    torch.cuda.synchronize()
    replay_equals_eager = digest() == d_replay
    b.close()
    del b
    torch.cuda.empty_cache()
    log(f"newton: all grids {ms_all:.1f} ms")

    red_bytes = reduced_bytes_model(NEWTON_NC, nbig, nnzb_big)
    return {
        "ncells": ncells, "ms_all": ms_all, "wall_ms_all": wall_all,
        "detail": {
            "workload": f"{len(cases) - 1} dual-porosity grids nx,ny~U{{32..160}}, nz~U{{8..32}} + "
                        f"{'x'.join(map(str, shapes[-1]))}, {NEWTON_NC} components PR, one hipGraph per "
                        "Newton-iteration evaluation: batched flash of every cell, property sweep + residual, "
                        "PhaseCtrl, Jacobian",
            "grids": [list(sh) for sh in shapes],
            "cells": ncells,
            "analytic_all_grids": {"ms_per_iter": ms_all, "cells_per_s": ncells / (ms_all * 1e-3),
                                   "reduced_blocks": nnzb_all, "flash": FAST_FLASH},
            "big_grid": {
                "cells": nbig,
                "plain_fp64_csr": {"ms_per_iter": ms_plain, "cells_per_s": nbig / (ms_plain * 1e-3),
                                   "jacobian_nnz": nnz_plain},
                "analytic_reduced": {"ms_per_iter": ms_mixed_big, "cells_per_s": nbig / (ms_mixed_big * 1e-3),
                                     "reduced_blocks": nnzb_big, "flash": FAST_FLASH,
                                     "tuned_knobs": tuned_p["config"]},
                "analytic_reduced_mixed_flash": {"ms_per_iter": ms_mixedflash_big,
                                                 "cells_per_s": nbig / (ms_mixedflash_big * 1e-3),
                                                 "flash": "mixed FP32 successive substitution + FP64 Newton",
                                                 "tuned_knobs": tuned["config"],
                                                 "flash_stages": {k: int(v) for k, v in mixed_stats.items()}},
                "speedup_analytic_vs_plain": ms_plain / ms_mixed_big,
                "reduced_assembly_ms": ms_red,
                "reduced_assembly_tuned_knobs": tuned_a["config"],
                "reduced_assembly_tuner_trials": [{"knobs": {k: v for k, v in c.items() if c[k] != tuned_a["config"][k]},
                                                   "ms": round(m, 2)} for c, m in tuned_a["trials"]],
                "linear_solve": {
                    "method": "BiCGStab + block Jacobi on the reduced system (osr_ksp_*)",
                    "setup_ms": ms_ksp_setup, "singular_diagonal_blocks": nsing,
                    "ms_per_iteration": ms_ksp_it, "iterations_timed": kinfo["iterations"],
                    "relative_residual_after": kinfo["rnorm"] / kinfo["bnorm"],
                    "solve_rtol_1e-6_maxit_200": {"reason": kdef["reason"], "iterations": kdef["iterations"],
                                                  "relative_residual": kdef["rnorm"] / kdef["bnorm"]},
                    "bsr_bytes_per_spmv": nnzb_big * (8 * (NEWTON_NC + 1) ** 2 + 4),
                    "spmv_note": "two SpMV and two block-Jacobi applications per iteration",
                },
                "restart_image": {
                    "what": "WriteRST / ReadRST byte image of the 256x256x32 state in HBM (osr_rst_write_device, "
                            "osr_rst_read_device), file transfer excluded",
                    "bytes": rst_b, "write_ms": ms_rst_w, "read_ms_wall": ms_rst_r,
                    "write_GBs_image_plus_state": 2 * rst_b / (ms_rst_w * 1e-3) / 1e9,
                },
                "device_time_step": {
                    "what": "SolveNonLinear (SNESKSPONLY) + UpdateAll(X, true) on the device "
                            "(osr_newton_step_device): pack, residual, reduced system, BiCGStab + block "
                            "Jacobi, expansion, acceptance (vnorm <= TOLMAIN), UpdateAll, PhaseCtrl, "
                            "CalcPCoeff, PJac, new sweep; wall time, host syncs included",
                    "dt_s": STEP_DT, "ksp_rtol": STEP_RTOL, "ksp_maxit": STEP_MAXIT,
                    "ms": ms_step, "first_call_ms": step_runs[0][0], "cells_per_s": nbig / (ms_step * 1e-3),
                    "accepted": sres["accepted"], "snes_reason": sres["snes_reason"], "vnorm": sres["vnorm"],
                    "ksp_iterations": sres["ksp"]["iterations"], "ksp_reason": sres["ksp"]["reason"],
                    "ksp_relative_residual": sres["ksp"]["rnorm"] / max(sres["ksp"]["bnorm"], 1e-300),
                },
            },
            "frac_two_phase_after_flash": float((labels == 0).mean()),
            "full_size_replay_equals_eager_bitwise": replay_equals_eager,
            "flash_stats": {k: int(v) for k, v in flash_stats.items()},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "reduced-system assembly of the 256x256x32 grid (osr_jacobian_bsr_reduced_direct_device: "
                      "k_props, k_red_locel, k_red_dprops, k_red_own, k_red_rhs)",
            "achieved": red_bytes / (ms_red * 1e-3) / 1e9,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": red_bytes / (ms_red * 1e-3) / 1e9 / PEAK_HBM_GBS,
            "algorithmic_bytes_per_launch": red_bytes,
            "avg_launch_ms": ms_red,
            "traffic": read_traffic(os.path.join(ROOT, "profiles", "reduced_assembly_pr10_256x256x32_pmc.txt")),
            # what actually bounds these kernels: VALU issue (counters of the
            # same launch, committed profile), against this run's launch time
            "valu": valu_view(os.path.join(ROOT, "profiles", "reduced_assembly_pr10_256x256x32_valu.txt"), ms_red),
        },
    }


def workload_stats(cells, labels, iters, K):
    """What the +/-15 %-of-Psat recipe yields under the reference's flash:
    the fraction of cells below their saturation pressure, the two-phase
    fraction overall and among those, and the near-critical fraction
    (two-phase cells whose converged K-values all lie within
    exp(+/-0.2))."""
    has = cells.psat > 0
    below = has & (cells.P < cells.psat)
    two = labels == 0
    lk = np.abs(np.log(np.where(K > 0, K, np.nan)))
    near = two & (np.nanmax(np.where(np.isfinite(lk), lk, np.inf), axis=0) < 0.2)
    return {"mean_ss_passes": float(iters.mean()), "max_ss_passes": int(iters.max()),
            "frac_two_phase": float(two.mean()), "frac_capped_500": float((iters >= 500).mean()),
            "frac_with_psat": float(has.mean()), "frac_below_psat": float(below.mean()),
            "frac_two_phase_among_below_psat": float(two[below].mean()) if below.any() else 0.0,
            "frac_near_critical": float(near.mean()),
            "note": "SuccFalsh starts from Wilson K-values and has no stability test: cells below their "
                    "bubble point whose Wilson Rachford-Rice has no root in (0, 1) stop after two passes "
                    "at L = 1 (single-phase liquid), so about a third of the below-Psat cells end two-phase "
                    "and the recipe's ~50 % target is not reached by the reference's own flash"}


def flash_leg(rt, synth, torch, args):
    """The batched-flash workload: plain, persistent and mixed forms on 2M
    cells (8 components, PR), flop-model roofline of the persistent form."""
    dev = torch.device("cuda:0")
    fl = synth.fluid_pseudo(NC, synth.PR, FLUID_SEED)
    ctx = rt.Context(fl)
    n = args.cells
    cells = synth.flash_cells(fl, n, CELL_SEED, rt.psat_numpy(ctx))
    T = torch.tensor(cells.T, device=dev)
    P = torch.tensor(cells.P, device=dev)
    z = torch.tensor(cells.z, device=dev)
    stream = torch.cuda.current_stream()
    steps = max(3, args.steps // 2)

    def run(mode, out, steps, warmup):
        for _ in range(warmup):
            rt.flash(ctx, T, P, z, mode=mode, out=out)
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(steps)]
        t0 = time.perf_counter()
        for s_, e in ev:
            s_.record(stream)
            rt.flash(ctx, T, P, z, mode=mode, out=out)
            e.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kern = float(np.mean([s_.elapsed_time(e) for s_, e in ev])) * 1e-3
        return wall / steps, kern

    from osiris_amd import tune
    tuned = {"persistent": tune.tune_flash(ctx, T, P, z, mode="persistent", reps=3)["config"]}
    tune.apply(ctx, {"config": tuned["persistent"]})
    out = rt.alloc_flash_out(NC, n)
    sec_persist, kern_persist = run("persistent", out, steps, 1)
    iters = out["iters"].cpu().numpy()
    rr = out["rr_iters"].cpu().numpy()
    labels = out["label"].cpu().numpy()
    gpu_res = {k: out[k].cpu().numpy() for k in ("x", "y", "K", "L", "label", "iters", "rr_iters")}
    out_plain = rt.alloc_flash_out(NC, n)
    sec_plain, kern_plain = run("plain", out_plain, max(3, steps // 4), 1)
    tuned["mixed"] = tune.tune_flash(ctx, T, P, z, mode="mixed", reps=3)["config"]
    tune.apply(ctx, {"config": tuned["mixed"]})
    out_mixed = rt.alloc_flash_out(NC, n)
    sec_mixed, kern_mixed = run("mixed", out_mixed, steps, 1)
    mixed_res = {k: out_mixed[k].cpu().numpy() for k in ("x", "y", "K", "L", "label", "iters", "rr_iters")}

    flops = flop_model(NC, iters, rr)
    achieved = flops / kern_persist / 1e12
    res = {
        "workload": "batched flash with uneven convergence: 8 pseudo-components, PR, "
                    f"{n} independent cells, p within +/-15% of Psat",
        "n_cells": n,
        "persistent_fp64": {"cells_per_s": n / sec_persist, "kernel_ms": kern_persist * 1e3},
        "plain_fp64": {"cells_per_s": n / sec_plain, "kernel_ms": kern_plain * 1e3},
        "mixed_fp32_fp64": {"cells_per_s": n / sec_mixed, "kernel_ms": kern_mixed * 1e3},
        "speedup_persistent_vs_plain": sec_plain / sec_persist,
        "speedup_mixed_vs_plain": sec_plain / sec_mixed,
        "tuned_knobs": tuned,
        "workload_stats": workload_stats(cells, labels, iters, gpu_res["K"]),
        "roofline": {
            "bound": "mfma",
            "bound_note": "FP64 vector ALU; MI355X FP64 vector and FP64 matrix peaks are both 78.6 TFLOP/s",
            "kernel": f"k_flash_persistent<{NC}> + k_flash_group<{NC}, 8> (one flash step)",
            "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS, "flops_per_launch": flops,
            "avg_launch_ms": kern_persist * 1e3,
            "algorithmic_bytes_per_launch": bytes_model(NC, n),
            "traffic": read_traffic(os.path.join(ROOT, "profiles", "flash_persistent_pr8_2M_pmc.txt")),
        },
    }
    return res, fl, cells, gpu_res, mixed_res


SRK_NC = 12
SRK_CELLS = 4_000_000
SRK_FLUID_SEED = 2
SRK_CELL_SEED = 11


def srk_mixed_leg(rt, synth, torch, args):
    """The mixed-precision configuration: 12 pseudo-components, SRK,
    4,000,000 independent cells at +/-15 % of Psat; the plain and persistent
    FP64 forms and the mixed FP32 SS + FP64 root/RR/Newton-polish form, and
    the mixed result against the FP64 GPU result on every cell."""
    dev = torch.device("cuda:0")
    fl = synth.fluid_pseudo(SRK_NC, synth.SRK, SRK_FLUID_SEED)
    ctx = rt.Context(fl)
    n = SRK_CELLS
    cells = synth.flash_cells(fl, n, SRK_CELL_SEED, rt.psat_numpy(ctx))
    T = torch.tensor(cells.T, device=dev)
    P = torch.tensor(cells.P, device=dev)
    z = torch.tensor(cells.z, device=dev)
    stream = torch.cuda.current_stream()
    steps = max(3, args.steps // 2)

    def run(mode, out, steps):
        rt.flash(ctx, T, P, z, mode=mode, out=out)
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(steps)]
        for s_, e in ev:
            s_.record(stream)
            rt.flash(ctx, T, P, z, mode=mode, out=out)
            e.record(stream)
        torch.cuda.synchronize()
        return float(np.mean([s_.elapsed_time(e) for s_, e in ev]))

    keys = ("x", "y", "K", "L", "label", "iters", "rr_iters")
    res = {}
    ms = {}
    from osiris_amd import tune
    tuned = {}
    for mode, k in (("plain", max(3, steps // 4)), ("persistent", steps), ("mixed", steps)):
        if mode != "plain":
            # launch knobs of this form chosen by the autotuner on these cells
            tuned[mode] = tune.tune_flash(ctx, T, P, z, mode=mode, reps=3)["config"]
            tune.apply(ctx, {"config": tuned[mode]})
        out = rt.alloc_flash_out(SRK_NC, n)
        ms[mode] = run(mode, out, k)
        if mode == "mixed":
            ms["mixed_stats"] = rt.flash_stats(ctx)
        res[mode] = {key: out[key].cpu().numpy() for key in keys}
        del out
    a, b = res["mixed"], res["persistent"]
    conv = (a["iters"] < 500) & (b["iters"] < 500)
    two = (b["label"] == 0) & conv

    def rel(u, v):
        e = np.abs(u - v) / np.maximum(np.abs(v), 1e-12)
        return np.where(np.isfinite(e), e, 0.0)

    def stats(e):
        return {"max": float(e.max()) if e.size else 0.0,
                "p99": float(np.quantile(e, 0.99)) if e.size else 0.0}

    vs_fp64 = {"label_mismatches": int((a["label"] != b["label"]).sum()),
               "cells_compared": int(conv.sum()),
               "K_rel_per_cell": stats(rel(a["K"], b["K"]).max(axis=0)[two]),
               "L_rel_two_phase": stats(rel(a["L"], b["L"])[two])}
    detail = {
        "workload": f"mixed-precision flash: {SRK_NC} pseudo-components, SRK, {n} independent cells, "
                    "p within +/-15% of Psat",
        "n_cells": n,
        "plain_fp64": {"cells_per_s": n / (ms["plain"] * 1e-3), "kernel_ms": ms["plain"]},
        "persistent_fp64": {"cells_per_s": n / (ms["persistent"] * 1e-3), "kernel_ms": ms["persistent"]},
        "mixed_fp32_fp64": {"cells_per_s": n / (ms["mixed"] * 1e-3), "kernel_ms": ms["mixed"],
                            "stages": ms["mixed_stats"]},
        "speedup_mixed_vs_plain": ms["plain"] / ms["mixed"],
        "speedup_persistent_vs_plain": ms["plain"] / ms["persistent"],
        "tuned_knobs": tuned,
        "frac_two_phase": float((b["label"] == 0).mean()),
        "workload_stats": workload_stats(cells, b["label"], b["iters"], b["K"]),
        "mixed_vs_gpu_fp64_every_cell": vs_fp64,
    }
    return detail, fl, cells, res


# The flash parity bars of tests/test_gpu_flash.py (compare_flash), applied
# by flash_bars to every cell of the full-size workloads.
FLASH_TOL, FLASH_TOL_K, FLASH_TOL_SLOW = 1e-8, 1e-7, 1e-6
FLASH_TOL_CAPPED, FLASH_TOL_CAPPED_MIXED = 1e-5, 5e-2
FLASH_BRANCH_FRAC, FLASH_TOL_BRANCH = 2e-3, 1e-6


SENS_EPS = (2.2e-16, 4.4e-16, 1e-15, 1e-14, 1e-13, 1e-12)


def oracle_k_sensitivity(O, fl, cells):
    """sens(c, target) -> (change, eps): the smallest relative perturbation
    eps of cell c's pressure, or of one of its feed fractions, under which the
    oracle's own K-values move by at least half of `target` (the GPU-oracle
    difference), and that change; (largest change seen, None) if none of
    SENS_EPS does -- how well K is determined by the FP64 inputs."""
    def sens(c, target):
        T, P, z = float(cells.T[c]), float(cells.P[c]), cells.z[:, c].copy()
        K0 = O.flash_cell(fl, T, P, z)["K"]
        worst = 0.0
        for eps in SENS_EPS:
            trials = [(P * (1 + eps), z), (P * (1 - eps), z)]
            for i in range(len(z)):
                z2 = z.copy()
                z2[i] *= 1 + eps
                trials.append((P, z2))
            for Pp, zp in trials:
                d = float(np.max(np.abs(O.flash_cell(fl, T, Pp, zp)["K"] / K0 - 1)))
                worst = max(worst, d)
                if d >= 0.5 * target:
                    return d, eps
        return worst, None
    return sens


def flash_bars(ref, res, mixed=False, sens=None):
    """Every cell against the oracle under the flash tests' bars: labels
    identical; x, y, L within 1e-8 (K 1e-7) on cells that stop within 100
    passes, except RR-branch cells (at most 0.2 % of them, within 1e-6);
    1e-6 on cells that need 100..499 passes; 1e-5 (mixed form 5e-2) where
    either side hits the 500-pass cap.  Counts per category, the largest
    error in each, the cells over their bar, and for the RR-branch cells how
    many stop after a different number of Rachford-Rice steps than the
    oracle."""
    it = np.maximum(res["iters"], ref["iters"])
    fast = it < 100
    slow = (it >= 100) & (it < 500)
    capped = it >= 500
    tol_capped = FLASH_TOL_CAPPED_MIXED if mixed else FLASH_TOL_CAPPED
    out = {"cells": int(it.size), "converged_lt_100_passes": int(fast.sum()), "slow_100_499": int(slow.sum()),
           "capped_either_side": int(capped.sum()),
           "label_mismatches": int((res["label"] != ref["label"]).sum())}
    branch = np.zeros(it.size, bool)
    bad = np.zeros(it.size, bool)
    fields = {}
    over_bar = 0
    finite_mismatch = 0
    for k in ("x", "y", "K", "L"):
        fin = np.isfinite(ref[k])
        finite_mismatch += int((fin != np.isfinite(res[k])).any(axis=0).sum() if fin.ndim == 2
                               else (fin != np.isfinite(res[k])).sum())
        e = np.where(fin, np.abs(res[k] - ref[k]) / np.maximum(np.abs(ref[k]), 1e-12), 0.0)
        ec = e.max(axis=0) if e.ndim == 2 else e
        tfast = FLASH_TOL_K if k == "K" else FLASH_TOL
        over = fast & (ec > tfast)
        branch |= over
        f = {}
        for name, mask, t in (("converged", fast & ~over, tfast), ("rr_branch", over, FLASH_TOL_BRANCH),
                              ("slow", slow, FLASH_TOL_SLOW), ("capped", capped, tol_capped)):
            m = float(ec[mask].max()) if mask.any() else 0.0
            nbad = int((ec[mask] > t).sum())
            bad |= mask & (ec > t)
            over_bar += nbad
            f[name] = {"max": m, "bar": t, "cells_over_bar": nbad}
        fields[k] = f
    out["fields"] = fields
    nb = int(branch.sum())
    allow = max(2, int(FLASH_BRANCH_FRAC * fast.sum()))
    out["rr_branch_cells"] = nb
    out["rr_branch_allowance"] = allow
    if "rr_iters" in res and "rr_iters" in ref and nb:
        out["rr_branch_cells_with_other_rr_count"] = int((res["rr_iters"][branch] != ref["rr_iters"][branch]).sum())
    out["finite_pattern_mismatches"] = finite_mismatch
    # the cells over their bar: passes and Rachford-Rice steps on both sides
    idx = np.flatnonzero(bad)
    eK = np.where(np.isfinite(ref["K"]), np.abs(res["K"] - ref["K"]) / np.maximum(np.abs(ref["K"]), 1e-12), 0.0)
    eK = eK.max(axis=0)
    out["cells_over_bar"] = int(idx.size)
    if idx.size:
        same_it = res["iters"][idx] == ref["iters"][idx]
        out["over_bar_same_pass_count"] = int(same_it.sum())
        if "rr_iters" in res and "rr_iters" in ref:
            out["over_bar_same_rr_count"] = int((res["rr_iters"][idx] == ref["rr_iters"][idx]).sum())
        # why each over-bar cell is over: its K-values are ill-conditioned
        # in the reference algorithm itself (the oracle's K moves by at least
        # half the difference under a relative input perturbation of at most
        # 1e-12, the size of a few passes' accumulated rounding), the two
        # sides stopped on different passes (the |dL| <= 1e-14 rule watches
        # L only), or the oracle hit the 500-pass cap where the mixed form's
        # Newton polish converged
        kinds = {"ill_conditioned": 0, "other_stop_pass": 0, "mixed_vs_capped_oracle": 0, "unexplained": 0}
        cells_out = []
        for c in idx[:200]:
            s, s_eps = sens(int(c), float(eK[c])) if sens is not None else (None, None)
            if mixed and ref["iters"][c] >= 500:
                kind = "mixed_vs_capped_oracle"
            elif s_eps is not None:
                kind = "ill_conditioned"
            elif res["iters"][c] != ref["iters"][c]:
                kind = "other_stop_pass"
            else:
                kind = "unexplained"
            kinds[kind] += 1
            if len(cells_out) < 12 or kind == "unexplained":
                cells_out.append({"cell": int(c), "gpu_passes": int(res["iters"][c]),
                                  "oracle_passes": int(ref["iters"][c]),
                                  "gpu_rr": int(res["rr_iters"][c]) if "rr_iters" in res else None,
                                  "oracle_rr": int(ref["rr_iters"][c]) if "rr_iters" in ref else None,
                                  "K_rel": float(eK[c]), "L": float(ref["L"][c]),
                                  "oracle_K_change": s, "at_input_perturbation": s_eps, "why": kind})
        out["over_bar_why"] = kinds
        out["over_bar_classified"] = int(min(idx.size, 200))
        out["over_bar_cells"] = cells_out
    out["meets_test_bars"] = bool(out["label_mismatches"] == 0 and over_bar == 0 and nb <= allow
                                  and finite_mismatch == 0)
    out["meets_test_bars_but_explained_cells"] = bool(
        out["label_mismatches"] == 0 and nb <= allow and finite_mismatch == 0
        and out.get("over_bar_why", {}).get("unexplained", 0) == 0 and idx.size <= 200)
    return out


def srk_oracle_parity(O, fl, cells, res, threads):
    """Every cell of the mixed-precision workload against the FP64 oracle:
    labels, K-values and phase fractions per cell, and the ln-fugacities
    ln(x_i phi_i P) of both phases on a sample of two-phase cells."""
    t0 = time.perf_counter()
    ref = O.flash_batch(fl, cells.T, cells.P, cells.z, nthreads=threads)
    cpu_s = time.perf_counter() - t0
    out = {"oracle_s": cpu_s, "oracle_threads": threads}
    for mode in ("mixed", "persistent"):
        r = res[mode]
        conv = (ref["iters"] < 500) & (r["iters"] < 500)
        two = conv & (ref["label"] == 0)
        eK = np.abs(r["K"] - ref["K"]) / np.maximum(np.abs(ref["K"]), 1e-300)
        eK = np.where(np.isfinite(eK), eK, 0.0).max(axis=0)
        eL = np.abs(r["L"] - ref["L"]) / np.maximum(np.abs(ref["L"]), 1e-300)
        rng = np.random.default_rng(0)
        idx = np.flatnonzero(two)
        idx = rng.choice(idx, size=min(2000, idx.size), replace=False) if idx.size else idx
        lnf = []
        for c in idx:
            T, P = float(cells.T[c]), float(cells.P[c])
            worst = 0.0
            for comp, gas in (("x", 0), ("y", 1)):
                _, l_mine = O.phase(fl, T, P, r[comp][:, c], gas)
                _, l_ref = O.phase(fl, T, P, ref[comp][:, c], gas)
                f_mine = np.log(r[comp][:, c]) + l_mine
                f_ref = np.log(ref[comp][:, c]) + l_ref
                worst = max(worst, float(np.max(np.abs(f_mine - f_ref))))
            lnf.append(worst)
        lnf = np.array(lnf)
        out[mode + "_test_bars"] = flash_bars(ref, r, mixed=(mode == "mixed"),
                                              sens=oracle_k_sensitivity(O, fl, cells))
        out[mode] = {
            "label_mismatches": int((r["label"] != ref["label"]).sum()),
            "cells_capped_either_side": int((~conv).sum()),
            "K_rel_per_cell_two_phase": {"max": float(eK[two].max()), "p99": float(np.quantile(eK[two], 0.99)),
                                         "frac_le_1e-8": float((eK[two] <= 1e-8).mean())},
            "L_rel_two_phase": {"max": float(eL[two].max()), "p99": float(np.quantile(eL[two], 0.99)),
                                "frac_le_1e-8": float((eL[two] <= 1e-8).mean())},
            "ln_fugacity_abs_sample": {"cells": int(idx.size), "max": float(lnf.max()) if lnf.size else 0.0,
                                       "p99": float(np.quantile(lnf, 0.99)) if lnf.size else 0.0},
        }
    return out


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_newton_sample(rt, synth, torch):
    """CPU baseline of the Newton-iteration step on the host cores: the
    oracle's flash of every cell, PhaseCtrl of both continua, residual and
    the reference's own Jacobian (IteratorJacobian restated, assembled row by
    row as FormJacobian does, PetscAUX.cpp:84-91, 135-143), on a 32x32x16
    dual-porosity grid of the same workload (10 components), once on one
    thread and once on every host core.  The GPU's CSR Jacobian and residual
    are then checked against the oracle on a 16x16x8 sample."""
    from oracle import oracle as O
    from osiris_amd.batch import NewtonBatch
    fl = synth.fluid_pseudo(NEWTON_NC, synth.PR, NEWTON_FLUID_SEED)

    def oflash(T, P, z, sw):
        r = O.flash_batch(fl, T, P, z, sw=sw)
        L, Zl, Zg = r["L"], r["Zl"], r["Zg"]
        nd = Zl * L / (Zl * L + Zg * (1 - L))
        r["so"], r["sg"] = nd * (1 - sw), (1 - nd) * (1 - sw)
        return r

    try:
        ncores = len(os.sched_getaffinity(0))
    except AttributeError:
        ncores = os.cpu_count() or 1
    ncores = min(ncores, 16)
    gridB, stB = synth.dual_porosity_case(fl, 32, 32, 16, 98, oflash, p_range=(8e6, 20e6))

    def step(grid, st, threads):
        n = grid.n
        T = np.full(2 * n, grid.T)
        P = np.concatenate([st.Po_m, st.Po_f])
        z = np.concatenate([st.z_m, st.z_f], axis=1)
        tp = [time.perf_counter()]
        O.flash_batch(fl, T, P, z, nthreads=threads)
        tp.append(time.perf_counter())
        pr = O.props(fl, grid, st)
        tp.append(time.perf_counter())
        for k, (lab, Po, x, y, S) in enumerate(((st.label_m, st.Po_m, st.x_m, st.y_m, st.S_m),
                                                (st.label_f, st.Po_f, st.x_f, st.y_f, st.S_f))):
            O.phase_ctrl(fl, np.full(n, grid.T), lab.copy(), Po, pr[k, :, 3].copy(), x.copy(), y.copy(),
                         S.copy(), nthreads=threads)
        tp.append(time.perf_counter())
        O.residual(fl, grid, st)
        tp.append(time.perf_counter())
        O.jacobian_ref(fl, grid, st, nthreads=threads)
        tp.append(time.perf_counter())
        return tp[-1] - tp[0], dict(zip(("flash", "property_sweep", "phase_test", "residual",
                                         "reference_jacobian"), np.diff(tp).tolist()))

    one_s, one_ph = step(gridB, stB, 1)
    all_s, all_ph = step(gridB, stB, ncores)
    nB = 2 * gridB.n
    base = {"value": nB / all_s, "unit": "cells/s", "cores": ncores, "kind": "port",
            "sample": f"32x32x16 dual-porosity grid of the same workload ({nB} cells, 10 components): oracle "
                      "flash of every cell + PhaseCtrl + residual + the reference's Jacobian "
                      "(IteratorJacobian restated, row by row), OpenMP over cells and blocks",
            "seconds": all_s, "phase_seconds": all_ph,
            "single_core": {"value": nB / one_s, "seconds": one_s, "phase_seconds": one_ph},
            "cpu_model": _cpu_model(),
            "note": "property sweep and residual run on one thread in both legs"}

    grid, st = synth.dual_porosity_case(fl, 16, 16, 8, 99, oflash, p_range=(8e6, 20e6))
    n = grid.n
    F_or = O.residual(fl, grid, st)
    rp, cols, vals = O.jacobian(fl, grid, st)
    # the GPU on the same sample (CSR form), against the oracle
    ctx = rt.Context(fl)
    b = NewtonBatch(ctx, fl, [(grid, st)], form="csr", flash_mode="plain")
    b.run_eager()
    torch.cuda.synchronize()
    it = b.items[0]
    Fg, vg = it["F"].cpu().numpy(), it["vals"].cpu().numpy()
    same_pattern = bool(np.array_equal(it["rowptr"].cpu().numpy(), rp) and
                        np.array_equal(it["cols"].cpu().numpy(), cols))
    Fs = O.residual_scale(fl, grid, st)
    fin = np.isfinite(F_or)
    f_err = float((np.abs(Fg[fin] - F_or[fin]) / np.maximum(Fs[fin], 1e-14 * np.nanmax(Fs) + 1e-300)).max())
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    rowmax = np.zeros(len(rp) - 1)
    np.maximum.at(rowmax, rows, np.abs(vals))
    dj = np.abs(vg - vals)
    rowerr = np.zeros(len(rp) - 1)
    np.maximum.at(rowerr, rows, dj)
    j_row = float((rowerr / np.maximum(rowmax, 1e-300)).max())
    big_entry = np.abs(vals) >= 1e-4 * rowmax[rows]
    j_entry = float((dj[big_entry] / np.abs(vals[big_entry])).max())
    j_floor = float((dj / np.maximum(np.abs(vals), 1e-8 * rowmax[rows])).max())
    b.close()
    # the analytic path of the headline (reduced blocks straight from the
    # state) on the same sample, against the Schur complement of the
    # oracle's CSR Jacobian and residual
    asm = rt.Assembler(ctx, fl, grid)
    d = asm.device_state(st)
    asm.layout(d)
    Fd = asm.residual(d)
    red = asm.reduced_direct(d, Fd)
    torch.cuda.synchronize()
    blocks, rhs, _, _, _, _ = O.reduced_from_csr(fl, grid, st, rp, cols, vals, F_or)
    browptr, bcols, bvals, brhs = (x.cpu().numpy() for x in red[:4])
    rerr = []
    for a in range(2 * n):
        deps = bcols[browptr[a]:browptr[a + 1]].tolist()
        ref = np.stack([blocks[(a, dd)] for dd in deps])             # [m, np1, np1]
        scale = np.maximum(np.abs(ref).max(axis=(0, 2)), np.abs(rhs[a]))
        e = np.abs(bvals[browptr[a]:browptr[a + 1]] - ref).max(axis=(0, 2))
        rerr.append(np.maximum(e, np.abs(brhs[a] - rhs[a])) / np.maximum(scale, 1e-300))
    rerr = np.concatenate(rerr)
    asm.close()
    return base, {"jacobian_pattern_identical": same_pattern,
                                "residual_max_row_scaled_err": f_err,
                                "jacobian_max_rowwise_rel_err": j_row,
                                "jacobian_max_rel_err_entries_above_1e-4_of_row": j_entry,
                                "jacobian_max_rel_err_floored_1e-8_of_row": j_floor,
                                "reduced_direct_vs_oracle_schur": {
                                    "max_row_scaled_err": float(rerr.max()),
                                    "p99": float(np.quantile(rerr, 0.99)),
                                    "rows_above_1e-8": int((rerr > 1e-8).sum()), "rows": int(rerr.size)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cells", type=int, default=N_CELLS)
    ap.add_argument("--newton-scale", type=float, default=1.0,
                    help="scale the grid dimensions of the batched-realization workload (debugging)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-flash", action="store_true", help="skip the batched-flash leg")
    ap.add_argument("--no-assembly", action="store_true", help="skip the 64x64x8 sensitivity leg")
    ap.add_argument("--no-newton", action="store_true", help="skip the Newton-iteration leg")
    ap.add_argument("--no-srk", action="store_true", help="skip the 12-component SRK mixed-precision leg")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="after the timed legs, copy a tensor of this many MiB once (a "
                         "kernel of known HBM traffic that scales rocprofv3 FETCH_SIZE / "
                         "WRITE_SIZE counters)")
    ap.add_argument("--quiet", action="store_true", help="no progress lines on stderr")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from osiris_amd import runtime as rt, synth

    def log(msg):
        if not args.quiet:
            print(msg, file=sys.stderr, flush=True)

    result = {
        "metric": "flash+assembly cells/s per Newton iter, plain FP64 vs mixed/analytic @<=1e-8 err",
        "value": None, "unit": "cells/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded PCG64 generator; no reference fixtures exist)",
    }
    if not args.no_newton:
        nl = newton_leg(rt, synth, torch, args, log)
        result["value"] = nl["ncells"] / (nl["ms_all"] * 1e-3)
        result["ms_per_step"] = nl["ms_all"]
        result["wall_ms_per_step"] = nl["wall_ms_all"]
        result["config"] = {"workload": nl["detail"]["workload"], "cells": nl["ncells"], "nc": NEWTON_NC,
                            "eos": "PR", "parallelism": "single GPU"}
        result["newton"] = nl["detail"]
        result["roofline"] = nl["roofline"]
    if not args.no_flash:
        fres, fl, cells, gpu_res, mixed_res = flash_leg(rt, synth, torch, args)
        result["flash_batched"] = fres
        log("flash leg done")
        if args.no_newton:
            result["value"] = fres["persistent_fp64"]["cells_per_s"]
            result["ms_per_step"] = fres["persistent_fp64"]["kernel_ms"]
            result["config"] = {"workload": fres["workload"], "n_cells": args.cells, "nc": NC, "eos": "PR",
                                "parallelism": "single GPU"}
            result["roofline"] = fres["roofline"]
    if not args.no_srk:
        sres, sfl, scells, sgpu = srk_mixed_leg(rt, synth, torch, args)
        result["flash_mixed_srk12"] = sres
        log("SRK mixed-precision leg done")
    if not args.no_assembly:
        result["sensitivities_64x64x8"] = assembly_leg(rt, synth, torch, args)
        log("sensitivity leg done")
    if args.calib_copy_mb > 0:
        src = torch.ones(args.calib_copy_mb * 2**20 // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    if not args.no_cpu:
        from oracle import oracle as O
        threads = min(16, os.cpu_count() or 1)
        base = {}
        if not args.no_newton:
            base, par = cpu_newton_sample(rt, synth, torch)
            result.setdefault("parity_vs_oracle", {})["newton_sample_csr"] = par
            log("cpu newton sample done")
        if not args.no_flash:
            n = args.cells
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
            result["flash_batched"]["cpu_baseline"] = {
                "value": n / cpu_s, "unit": "cells/s", "cores": threads, "kind": "port",
                "sample": f"full {n}-cell step, FP64 C oracle with OpenMP over cells",
                "single_core_cells_per_s": m / cpu1_s, "single_core_sample": f"first {m} cells, 1 thread"}
            result.setdefault("parity_vs_oracle", {})["flash_persistent_fp64"] = parity(gpu_res)
            result["parity_vs_oracle"]["flash_mixed_fp32_fp64"] = parity(mixed_res)
            sens = oracle_k_sensitivity(O, fl, cells)
            result["parity_vs_oracle"]["flash_persistent_fp64_test_bars"] = flash_bars(ref, gpu_res, sens=sens)
            result["parity_vs_oracle"]["flash_mixed_fp32_fp64_test_bars"] = flash_bars(ref, mixed_res, mixed=True,
                                                                                   sens=sens)
        if not args.no_srk:
            result.setdefault("parity_vs_oracle", {})["flash_mixed_srk12"] = srk_oracle_parity(
                O, sfl, scells, sgpu, threads)
            log("SRK oracle parity done")
        if base:
            result["cpu_baseline"] = base
        elif "cpu_baseline" in result.get("flash_batched", {}):
            result["cpu_baseline"] = result["flash_batched"]["cpu_baseline"]
    print(json.dumps(result))


if __name__ == "__main__":
    main()
