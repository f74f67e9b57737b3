"""Batched realizations of different grid sizes captured in one hipGraph:
every replay (with eager work and synchronisations between replays) must
reproduce the eager residual and Jacobian bitwise (the kernels are
deterministic and atomic-free)."""
import numpy as np
import pytest
import torch

from osiris_amd import synth
from tests.asm_cases import oracle_flash_fn

pytestmark = pytest.mark.gpu


def test_batched_realizations_in_one_graph(gpu, oracle):
    from osiris_amd.batch import RealizationBatch
    fl = synth.fluid_pseudo(6, synth.PR, 3)
    shapes = [(6, 5, 3), (9, 4, 2), (5, 7, 4), (3, 3, 1)]
    cases = [synth.dual_porosity_case(fl, nx, ny, nz, 10 + i, oracle_flash_fn(oracle, fl), p_range=(8e6, 20e6))
             for i, (nx, ny, nz) in enumerate(shapes)]
    ctx = gpu.Context(fl)
    batch = RealizationBatch(ctx, fl, cases)
    eager = [[t.clone() for t in r] for r in batch.run_eager()]
    torch.cuda.synchronize()
    batch.capture()
    for rep in range(3):
        for r in batch.items:          # clobber the outputs between replays
            r["F"].fill_(np.nan)
            r["vals"].fill_(np.nan)
        torch.cuda.synchronize()
        out = batch.replay()
        torch.cuda.synchronize()
        for (F, rp, cols, vals), (F0, rp0, cols0, vals0) in zip(out, eager):
            assert torch.equal(F.nan_to_num(7.0), F0.nan_to_num(7.0)), rep
            assert torch.equal(cols, cols0) and torch.equal(vals.nan_to_num(7.0), vals0.nan_to_num(7.0)), rep
        # eager work on the default stream between replays
        _ = torch.ones(1000, device="cuda").sum().item()
    # and the graph results match the oracle on the smallest realization
    fl_, grid, st = fl, cases[3][0], cases[3][1]
    F_or = oracle.residual(fl_, grid, st)
    F_gpu = out[3][0].cpu().numpy()
    fin = np.isfinite(F_or)
    Fs = oracle.residual_scale(fl_, grid, st)
    err = np.abs(F_gpu[fin] - F_or[fin]) / np.maximum(Fs[fin], 1e-14 * np.nanmax(Fs) + 1e-300)
    assert err.max() <= 1e-8
