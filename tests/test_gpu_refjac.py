"""GPU parity of the reference-mode Jacobian (osr_jacobian_ref_device, the
reference's IteratorJacobian with its own pattern and hand rules) against
the oracle's restatement (or_jacobian_ref) on the same seeded states.

Parity bar: row pointer and column indices identical (the pattern is integer
work); every value within 1e-8 relative, floored at 1e-8 x the largest
magnitude of its row (the rows add up terms of very different size, and
entries that cancel to near zero carry the absolute rounding of their
terms).
"""
import numpy as np
import pytest
import torch

from osiris_amd import synth
from tests.asm_cases import c0_case, frozen, oracle_flash_fn

pytestmark = pytest.mark.gpu

RTOL = 1e-8


def _cases(oracle):
    yield "c0_10x10x1", c0_case(oracle)
    yield "c0_4x3x3", c0_case(oracle, nx=4, ny=3, nz=3, seed=1, sw_m=0.25, sw_f=0.05)
    yield "c0_cdif0", c0_case(oracle, nx=5, ny=4, nz=2, seed=2, const_diff=True, cdiff=0.0)
    fl6 = synth.fluid_pseudo(6, synth.PR, 3)
    g6, s6 = synth.dual_porosity_case(fl6, 6, 5, 3, 4, oracle_flash_fn(oracle, fl6), p_range=(8e6, 20e6))
    yield "pr6_6x5x3", (fl6, g6, s6)
    g6b, s6b = synth.dual_porosity_case(fl6, 12, 10, 4, 5, oracle_flash_fn(oracle, fl6), p_range=(5e6, 25e6))
    yield "pr6_12x10x4", (fl6, g6b, s6b)


def _compare(name, g, r):
    rp, cols, vals = r
    grp, gcols, gvals = g
    assert np.array_equal(grp, rp), f"{name}: row pointers differ"
    assert np.array_equal(gcols, cols), f"{name}: column indices differ"
    assert np.array_equal(np.isfinite(gvals), np.isfinite(vals)), f"{name}: finiteness differs"
    fin = np.isfinite(vals)
    a = np.where(fin, np.abs(vals), 0.0)
    rowmax = np.repeat(np.maximum.reduceat(a, rp[:-1]), np.diff(rp))
    err = np.where(fin, np.abs(gvals - vals) / np.maximum(a, 1e-8 * rowmax), 0.0)
    worst = int(np.argmax(err))
    row = int(np.searchsorted(rp, worst, side="right") - 1)
    assert err.max(initial=0) <= RTOL, (f"{name}: max rel err {err.max():.3e} at entry {worst} (row {row}, "
                                        f"col {cols[worst]}): gpu {gvals[worst]!r} oracle {vals[worst]!r}")
    return float(err.max(initial=0))


def test_reference_jacobian_parity(gpu, oracle):
    for name, (fl, grid, st) in _cases(oracle):
        for label, s in (("X0", st), ("frozen", frozen(oracle, fl, grid, st))):
            what = f"{name}/{label}"
            print(f"reference Jacobian: {what}, n {grid.n}, nc {fl.nc}", flush=True)
            ctx = gpu.Context(fl)
            asm = gpu.Assembler(ctx, fl, grid)
            d = asm.device_state(s)
            asm.layout(d)
            rp, cols, vals = asm.jacobian(d, reference=True)
            torch.cuda.synchronize()
            g = (rp.cpu().numpy(), cols.cpu().numpy(), vals.cpu().numpy())
            r = oracle.jacobian_ref(fl, grid, s)
            err = _compare(what, g, r)
            print(f"  nnz {len(r[1])}, max rel err {err:.2e}", flush=True)
            asm.close()
            ctx.close()


def test_reference_jacobian_zero_porosity(gpu, oracle):
    # Rows next to a zero-porosity block divide by its zero porosity in the
    # Fickian terms, as the reference's do: those entries are non-finite on
    # both sides; the rest is held to the bar.
    fl, grid, st = c0_case(oracle, nx=4, ny=3, nz=2, seed=3)
    grid.phi_m = grid.phi_m.copy()
    grid.phi_m[[1, 7]] = 0.0
    grid.phi_f = grid.phi_f.copy()
    grid.phi_f[5] = 0.0
    s = frozen(oracle, fl, grid, st)
    print("reference Jacobian: zero-porosity blocks, 4x3x2", flush=True)
    ctx = gpu.Context(fl)
    asm = gpu.Assembler(ctx, fl, grid)
    d = asm.device_state(s)
    asm.layout(d)
    rp, cols, vals = asm.jacobian(d, reference=True)
    torch.cuda.synchronize()
    g = (rp.cpu().numpy(), cols.cpu().numpy(), vals.cpu().numpy())
    _compare("zero porosity", g, oracle.jacobian_ref(fl, grid, s))
    asm.close()
    ctx.close()
