"""Launch-shape autotuner for the flash kernels.

The knobs are the `osr_tune_set` keys of a context (no counterpart in the
reference; they change scheduling only, never the arithmetic of a cell):

  flash.refill_min        idle lanes a persistent wave waits for before it
                          refills from the work queue (one atomic per refill)
  flash.park_after        FP64 passes after which a still-iterating cell is
                          handed to the lane-group kernel (0 = never)
  flash.blocks_per_cu     resident blocks per CU of the persistent grid
                          (0 = occupancy calculator)
  flash.mixed_warm_iters  FP64 passes of the mixed path before its FP32 stage
  flash.mixed_tol32       FP32 stage stop of the mixed path
  flash.mixed_max32       FP32 stage pass cap of the mixed path

and for the reduced (nc+1)^2 system assembled straight from the state:

  reduced.local_rows      1: local rows from closed-form EOS derivatives,
                          eliminated in one lane group per block; 0: forward
                          mode per (block, unknown), then the elimination
  reduced.dprops_wg       workgroup size of the direction sweep (64/128/256)
  reduced.own_wg          workgroup size of the balance-row kernel
  assembly.xcd_remap      1: each XCD walks a contiguous eighth of the grid

`tune_flash` times each candidate on the caller's own cells with HIP events
(median of a few launches) and walks the knobs one at a time, keeping the
best value of each (coordinate descent from the built-in defaults);
`tune_assembly` does the same with the reduced-system call of an assembler
on the caller's state.  The result is a plain dict that `apply` sets on any
context of the same fluid and `save`/`load` keep as JSON.  Apply it before
a graph is captured: a captured launch keeps the shape it was recorded with.
"""
from __future__ import annotations

import json

from . import runtime as rt

DEFAULTS = {
    "persistent": {"flash.refill_min": 48, "flash.park_after": 40, "flash.blocks_per_cu": 0},
    "mixed": {"flash.refill_min": 48, "flash.blocks_per_cu": 0, "flash.mixed_tol32": 1e-4,
              "flash.mixed_warm_iters": 3,
              "flash.mixed_max32": 150},
}

ASM_DEFAULTS = {"reduced.local_rows": 1, "reduced.dprops_wg": 256, "reduced.own_wg": 64,
                "assembly.xcd_remap": 0}

SPACE = {
    "reduced.local_rows": [0, 1],
    "reduced.dprops_wg": [64, 128, 256],
    "reduced.own_wg": [64, 128, 256],
    "assembly.xcd_remap": [0, 1],
    "flash.refill_min": [8, 16, 32, 48, 64],
    "flash.park_after": [0, 16, 24, 40, 64, 100, 160, 250],
    "flash.blocks_per_cu": [0, 1, 2, 3, 4, 6],
    "flash.mixed_tol32": [1e-4, 3e-5, 1e-5, 3e-6],
    "flash.mixed_max32": [40, 80, 150, 250],
    "flash.mixed_warm_iters": [2, 3, 5, 8, 16, 24, 40],
}


def _time(ctx, T, P, z, mode, out, reps):
    import torch
    stream = torch.cuda.current_stream()
    rt.flash(ctx, T, P, z, mode=mode, out=out)          # warm-up (and workspace allocation)
    times = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        rt.flash(ctx, T, P, z, mode=mode, out=out)
        e1.record(stream)
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
    times.sort()
    return times[len(times) // 2]


def tune_flash(ctx, T, P, z, mode="persistent", reps=5, sweeps=1, log=None):
    """Coordinate descent over the knobs of `mode` on cells (T, P, z) already
    on the device.  Returns {"mode", "config", "ms", "trials"}."""
    import torch
    if mode not in DEFAULTS:
        raise ValueError(f"mode must be one of {sorted(DEFAULTS)}")
    z = rt._as_dev(z, torch.float64)
    out = rt.alloc_flash_out(ctx.nc, z.shape[1], full=False)
    cfg = dict(DEFAULTS[mode])
    for k, v in cfg.items():
        ctx.tune(k, v)
    best = _time(ctx, T, P, z, mode, out, reps)
    trials = [(dict(cfg), best)]
    for _ in range(sweeps):
        for key in cfg:
            for v in SPACE[key]:
                if v == cfg[key]:
                    continue
                ctx.tune(key, v)
                ms = _time(ctx, T, P, z, mode, out, reps)
                trials.append(({**cfg, key: v}, ms))
                if log:
                    log(f"{key}={v}: {ms:.3f} ms (best {best:.3f})")
                if ms < best:
                    best = ms
                    cfg[key] = v
            ctx.tune(key, cfg[key])
    return {"mode": mode, "config": cfg, "ms": best, "trials": trials}


def _descend(ctx, cfg, timer, sweeps, log):
    for k, v in cfg.items():
        ctx.tune(k, v)
    best = timer()
    trials = [(dict(cfg), best)]
    for _ in range(sweeps):
        for key in cfg:
            for v in SPACE[key]:
                if v == cfg[key]:
                    continue
                ctx.tune(key, v)
                ms = timer()
                trials.append(({**cfg, key: v}, ms))
                if log:
                    log(f"{key}={v}: {ms:.3f} ms (best {best:.3f})")
                if ms < best:
                    best = ms
                    cfg[key] = v
            ctx.tune(key, cfg[key])
    return best, trials


def tune_assembly(ctx, asm, d, F, reps=3, sweeps=1, log=None):
    """Coordinate descent over the reduced system's launch knobs, timing the
    call asm.reduced_direct(d, F) (HIP events, median of `reps`) on the
    caller's state.  Returns {"mode": "reduced", "config", "ms", "trials"}."""
    import torch
    out = asm.reduced_direct(d, F)

    def timer():
        stream = torch.cuda.current_stream()
        asm.reduced_direct(d, F, out=out)
        times = []
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            asm.reduced_direct(d, F, out=out)
            e1.record(stream)
            e1.synchronize()
            times.append(e0.elapsed_time(e1))
        times.sort()
        return times[len(times) // 2]

    cfg = dict(ASM_DEFAULTS)
    best, trials = _descend(ctx, cfg, timer, sweeps, log)
    return {"mode": "reduced", "config": cfg, "ms": best, "trials": trials}


def apply(ctx, result):
    for k, v in result["config"].items():
        ctx.tune(k, v)


def save(result, path):
    with open(path, "w") as f:
        json.dump({"mode": result["mode"], "config": result["config"], "ms": result["ms"]}, f, indent=1)


def load(path):
    with open(path) as f:
        return json.load(f)


def main(argv=None):
    """Tune both launch forms on the batched-flash workload (8 pseudo-
    components, PR, p within +/-15 % of Psat) and print the results."""
    import argparse
    import torch
    from . import synth
    ap = argparse.ArgumentParser()
    ap.add_argument("--nc", type=int, default=8)
    ap.add_argument("--eos", default="PR")
    ap.add_argument("--cells", type=int, default=2_000_000)
    ap.add_argument("--out", default=None, help="write the results as JSON here")
    args = ap.parse_args(argv)
    fl = synth.fluid_pseudo(args.nc, getattr(synth, args.eos), 1)
    ctx = rt.Context(fl)
    cells = synth.flash_cells(fl, args.cells, 7, rt.psat_numpy(ctx))
    T = torch.tensor(cells.T, device="cuda")
    P = torch.tensor(cells.P, device="cuda")
    z = torch.tensor(cells.z, device="cuda")
    res = {}
    for mode in ("persistent", "mixed"):
        r = tune_flash(ctx, T, P, z, mode=mode, log=lambda m: print(mode, m, flush=True))
        res[mode] = {"config": r["config"], "ms": r["ms"]}
        print(mode, "best", r["config"], f"{r['ms']:.3f} ms", flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1)
    return res


if __name__ == "__main__":
    main()
