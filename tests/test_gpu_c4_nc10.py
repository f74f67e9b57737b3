"""The batched-realization Newton evaluation (NewtonBatch, one hipGraph) at
10 components against the oracle: three dual-porosity grids of different
shapes in one batch, mixed-precision flash, reduced (nc+1)^2 blocks
assembled straight from the state.

Bars:
  * flash of every cell of every grid: the per-category bars of
    test_gpu_flash.compare_flash (labels identical; 1e-8 on cells that
    converge within 100 passes, ...; the mixed form's capped cells 5e-2);
  * residual: 1e-8 of each row's scale (the sum of the magnitudes of its
    terms), floored at 1e-14 of the largest row scale (rows whose terms all
    vanish carry only rounding);
  * reduced blocks and right-hand sides: against the Schur complement of the
    oracle's CSR Jacobian and residual (oracle.reduced_from_csr), within
    1e-8 of the largest magnitude of their reduced row (blocks and rhs),
    floored at 1e-14 of the largest block entry;
  * the elimination factors S and t, through the expansion of a random
    reduced update: ds = -S dp - t within 1e-8 of the block's largest
    |ds| plus |dp|.
"""
import numpy as np
import pytest
import torch

from osiris_amd import synth
from osiris_amd.batch import NewtonBatch
from tests.asm_cases import oracle_flash_fn
from tests.test_gpu_flash import TOL_CAPPED_MIXED, compare_flash

pytestmark = pytest.mark.gpu

NC = 10
RTOL = 1e-8


def test_newton_batch_nc10_against_oracle(gpu, oracle):
    fl = synth.fluid_pseudo(NC, synth.PR, 11)
    shapes = [(6, 5, 3, 21), (8, 4, 2, 22), (5, 5, 4, 23)]
    cases = [synth.dual_porosity_case(fl, nx, ny, nz, seed, oracle_flash_fn(oracle, fl), p_range=(8e6, 20e6))
             for nx, ny, nz, seed in shapes]
    ctx = gpu.Context(fl)
    print("C4 nc=10: NewtonBatch of", [g.n for g, _ in cases], "cells, graph replay", flush=True)
    b = NewtonBatch(ctx, fl, cases, form="reduced", flash_mode="mixed")
    b.replay()
    torch.cuda.synchronize()

    # flash of every cell of the batch
    T = np.full(b.ncells, b.T)
    ref = oracle.flash_batch(fl, T, b.P_all.cpu().numpy(), b.z_all.cpu().numpy())
    compare_flash(b.fout, ref, what="C4 mixed flash", tol_capped=TOL_CAPPED_MIXED)

    rng = np.random.default_rng(5)
    for (grid, st), it in zip(cases, b.items):
        name = f"{grid.nx}x{grid.ny}x{grid.nz}"
        print(f"C4 nc=10 grid {name}: residual and reduced blocks vs oracle", flush=True)
        F_or = oracle.residual(fl, grid, st)
        Fs = oracle.residual_scale(fl, grid, st)
        F = it["F"].cpu().numpy()
        assert np.isfinite(F_or).all() and np.isfinite(F).all(), name
        e = np.abs(F - F_or) / np.maximum(Fs, 1e-14 * Fs.max())
        assert e.max() <= RTOL, (name, "residual", float(e.max()))

        rp, cols, vals = oracle.jacobian(fl, grid, st)
        blocks, rhs, S_o, t_o, P, Sc = oracle.reduced_from_csr(fl, grid, st, rp, cols, vals, F_or)
        browptr, bcols, bvals, brhs, S, t = (x.cpu().numpy() for x in it["red"])
        nb, np1 = 2 * grid.n, NC + 1
        gmax = max(np.abs(B).max() for B in blocks.values())
        worst = 0.0
        for a in range(nb):
            deps = bcols[browptr[a]:browptr[a + 1]].tolist()
            assert sorted(deps) == sorted(d for (aa, d) in blocks if aa == a), (name, a)
            gb = {d: bvals[browptr[a] + j] for j, d in enumerate(deps)}
            scale = np.maximum(np.max([np.abs(blocks[(a, d)]).max(axis=1) for d in deps], axis=0),
                               np.abs(rhs[a]))
            err = np.max([np.abs(gb[d] - blocks[(a, d)]).max(axis=1) for d in deps], axis=0)
            err = np.maximum(err, np.abs(brhs[a] - rhs[a])) / np.maximum(scale, 1e-14 * gmax)
            worst = max(worst, float(err.max()))
        print(f"  reduced blocks / rhs: worst row-scaled error {worst:.2e}", flush=True)
        assert worst <= RTOL, (name, "reduced blocks", worst)

        dp = rng.standard_normal((nb, np1))
        du = it["asm"].reduced_expand(it["d"], it["red"][4], it["red"][5],
                                      torch.as_tensor(dp).cuda()).cpu().numpy()
        worst = 0.0
        for blk in range(nb):
            assert np.array_equal(du[P[blk]], dp[blk]), (name, blk)
            ds = -S_o[blk] @ dp[blk] - t_o[blk]
            scale = max(np.abs(ds).max(), np.abs(dp[blk]).max())
            worst = max(worst, float(np.abs(du[Sc[blk]] - ds).max() / scale))
        print(f"  expansion (S, t): worst error {worst:.2e}", flush=True)
        assert worst <= RTOL, (name, "expansion", worst)
    b.close()
    ctx.close()
