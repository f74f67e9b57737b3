"""GPU parity of the well terms against the CPU FP64 oracle: the residual,
the exact Jacobian and the reduced system with wells of every type (0..8)
at interior, side and corner blocks, under WellCondition on and off, inside
and outside their time windows, several wells in one block, producers shut
in below WELL_MIN_P and injectors above WELL_MAX_P.

Bars: Calc_Re (host code on both sides, the same operations) equal to 1e-15;
FlashInjection (device) within 1e-10 of the oracle's; residual rows within
1e-8 of the row's scale and Jacobian values within 1e-8 relative (floored at
1e-8 of the row's largest entry), as for the assembly without wells; the
reduced system assembled from the state against the one reduced from the
full CSR as in the reduced-system tests, and its solve against the full
Newton system.  States whose rate allocation divides 0 by 0 (a rate well
in a column with an immobile phase) produce NaN rows in the reference;
there the finite/non-finite pattern must match.
"""
import numpy as np
import pytest
import torch

from osiris_amd import synth
from tests.asm_cases import c0_case, column_state, frozen, oracle_flash_fn, with_wells, wrow
from tests.test_gpu_reduced import _check_solve, _same_reduction

pytestmark = pytest.mark.gpu

RTOL = 1e-8


def _two_phase(st, grid, cols, Po):
    """Columns (ix, iy) made two-phase with mobile water, oil and gas in both
    continua, each at its own pressure pair."""
    for (ix, iy), p in zip(cols, Po):
        st = column_state(st, grid, ix, iy, Po=p, label=0, x=[0.3, 0.3, 0.4], y=[0.8, 0.15, 0.05],
                          S=([0.3, 0.6, 0.1], [0.3, 0.5, 0.2]))
    return st


def _c0_all_types(oracle):
    fl, grid, st = c0_case(oracle, nx=5, ny=4, nz=3, seed=5, sw_m=0.45, sw_f=0.3)
    nc = fl.nc
    z = np.array([0.6, 0.1, 0.3])
    wells = [wrow(nc, 0, 2, 1, 0, 1, 8e6, 5.0),                  # interior, oil-rate cap
             wrow(nc, 1, 0, 0, 0, 2, 8e6, 3.0),                  # corner, WellCondition on
             wrow(nc, 2, 4, 3, 1, 2, 30e6, 5.0),                 # corner
             wrow(nc, 3, 0, 2, 0, 2, 13e6, 5.0, z=z),            # side, WellCondition on
             wrow(nc, 4, 3, 0, 0, 2, 2.0, 0.0),                  # side
             wrow(nc, 5, 1, 3, 0, 1, 2.0, 0.0),                  # side, WellCondition on
             wrow(nc, 6, 3, 2, 0, 2, 4.0, 0.0),                  # interior
             wrow(nc, 7, 1, 1, 0, 0, 30e6, 5.0),                 # interior, WellCondition on
             wrow(nc, 8, 4, 1, 0, 2, 40.0, 13e6, z=z),           # side, one block above WELL_MAX_P
             wrow(nc, 0, 2, 2, 0, 2, 9e6, 0.0),                  # two wells in one column:
             wrow(nc, 4, 2, 2, 1, 1, 1.5, 0.0),                  #   the later active one sets layer 1
             wrow(nc, 5, 2, 2, 0, 2, 1.0, 0.0, t0=7200.0)]       #   not yet open
    cond = [0, 1, 0, 1, 0, 1, 0, 1, 0, 0, 0, 0]
    cols = [(2, 1), (0, 0), (4, 3), (0, 2), (3, 0), (1, 3), (3, 2), (1, 1), (4, 1), (2, 2)]
    Po = [(10e6, 10.2e6), (11e6, 11.3e6), (9e6, 9.1e6), (10e6, 10.4e6), (12e6, 12.2e6), (10.5e6, 10.6e6),
          (11.5e6, 11.7e6), (9.5e6, 9.8e6), (12e6, 12.5e6), (10e6, 10.1e6)]
    st = _two_phase(st, grid, cols, Po)
    c_hot = (2 * grid.ny + 1) * grid.nx + 4                      # (4, 1) layer 2 above 14 MPa
    st.Po_m[c_hot] = 14.5e6
    st.Po_f[c_hot] = 14.6e6
    return fl, with_wells(grid, wells, sim_time=3600.0, re_f=50.0, cond=cond), st


def _cases(oracle):
    yield "c0_all_types", _c0_all_types(oracle)
    # the same wells on the seeded states as they are: single-phase columns
    # where rate wells divide by an immobile phase's mobility
    fl, g, _ = _c0_all_types(oracle)
    _, _, st = c0_case(oracle, nx=5, ny=4, nz=3, seed=5, sw_m=0.45, sw_f=0.3)
    yield "c0_all_types_raw", (fl, g, st)
    # six components: a two-phase and a gas injection stream, a producer
    fl6 = synth.fluid_pseudo(6, synth.PR, 3)
    g6, s6 = synth.dual_porosity_case(fl6, 6, 5, 3, 4, oracle_flash_fn(oracle, fl6), p_range=(8e6, 13e6))
    z2 = np.array([0.3, 0.1, 0.1, 0.1, 0.2, 0.2])
    zg = np.array([0.9, 0.1, 0.0, 0.0, 0.0, 0.0])
    w6 = [wrow(6, 3, 1, 1, 0, 2, 25e6, 0.0, z=z2), wrow(6, 8, 4, 3, 0, 1, 30.0, 13e6, z=zg),
          wrow(6, 8, 5, 0, 2, 2, 30.0, 13e6, z=z2), wrow(6, 0, 3, 2, 0, 2, 7.5e6, 0.0, skin=1.5, rw=0.08)]
    yield "pr6_injectors", (fl6, with_wells(g6, w6, sim_time=1e5, re_f=30.0, cond=[0, 1, 0, 0]), s6)


def _row_err(F, Fr, Fs):
    fin = np.isfinite(Fr)
    assert np.array_equal(fin, np.isfinite(F))
    err = np.abs(F - Fr) / np.maximum(Fs, 1e-14 * np.nanmax(Fs) + 1e-300)
    return err[fin].max(initial=0)


def test_well_data_parity(gpu, oracle):
    for name, (fl, grid, st) in _cases(oracle):
        asm = gpu.Assembler(gpu.Context(fl), fl, grid)
        re, inj = asm.well_data()
        rr, ir = oracle.well_data(fl, grid, st)
        np.testing.assert_allclose(re, rr, rtol=1e-15, err_msg=name)
        fin = np.isfinite(ir)
        assert np.array_equal(fin, np.isfinite(inj)), name
        np.testing.assert_allclose(inj[fin], ir[fin], rtol=1e-10, atol=1e-300, err_msg=name)


def test_residual_and_jacobian_with_wells(gpu, oracle):
    for name, (fl, grid, st) in _cases(oracle):
        for label, s in (("X0", st), ("frozen", frozen(oracle, fl, grid, st))):
            ctx = gpu.Context(fl)
            asm = gpu.Assembler(ctx, fl, grid)
            d = asm.device_state(s)
            asm.layout(d)
            F = asm.residual(d).cpu().numpy()
            Fr = oracle.residual(fl, grid, s)
            Fs = oracle.residual_scale(fl, grid, s)
            e = _row_err(F, Fr, Fs)
            print(name, label, "residual err", e, "non-finite rows", int((~np.isfinite(Fr)).sum()), flush=True)
            assert e <= RTOL, (name, label, e)
            if name != "c0_all_types_raw":
                assert np.isfinite(Fr).all(), name
            # the wells change the residual
            F0 = oracle.residual(fl, _no_wells(grid), s)
            assert np.count_nonzero(np.nan_to_num(Fr) != np.nan_to_num(F0)) > 0
            rp, cols, vals = asm.jacobian(d)
            rpo, colso, valso = oracle.jacobian(fl, grid, s)
            assert np.array_equal(rp.cpu().numpy(), rpo) and np.array_equal(cols.cpu().numpy(), colso), name
            v = vals.cpu().numpy()
            fin = np.isfinite(valso)
            assert np.array_equal(fin, np.isfinite(v)), name
            row = np.repeat(np.arange(len(rpo) - 1), np.diff(rpo))
            rowmax = np.zeros(len(rpo) - 1)
            np.maximum.at(rowmax, row, np.where(fin, np.abs(valso), 0))
            err = np.abs(v - valso)[fin] / np.maximum(np.abs(valso[fin]), 1e-8 * rowmax[row[fin]] + 1e-300)
            assert err.max(initial=0) <= RTOL, (name, label, err.max())


def _no_wells(grid):
    import copy
    g = copy.copy(grid)
    g.wells = None
    return g


def test_reduced_system_with_wells(gpu, oracle):
    for name, (fl, grid, st) in _cases(oracle):
        if name == "c0_all_types_raw":
            continue
        s = frozen(oracle, fl, grid, st)
        asm = gpu.Assembler(gpu.Context(fl), fl, grid)
        d = asm.device_state(s)
        _, nrows = asm.layout(d)
        F = asm.residual(d)
        rp, cols, vals = asm.jacobian(d)
        red = asm.reduced(d, F, rp, vals)
        red_direct = asm.reduced_direct(d, F)
        torch.cuda.synchronize()
        np1, nb = fl.nc + 1, 2 * grid.n
        print(name, "direct vs CSR reduction", _same_reduction(name, red, red_direct, np1), flush=True)
        _check_solve(name, asm, d, F, rp, cols, vals, red, nb, np1, nrows)
        _check_solve(name + " direct", asm, d, F, rp, cols, vals, red_direct, nb, np1, nrows)


def test_wells_reset_in_place_and_removed(gpu, oracle):
    fl, grid, st = _c0_all_types(oracle)
    asm = gpu.Assembler(gpu.Context(fl), fl, grid)
    d = asm.device_state(st)
    asm.layout(d)
    F1 = asm.residual(d).clone()
    # the same wells later in time with other conditions: same device tables
    w = grid.wells
    w2 = synth.Wells(w.welli.copy(), w.wellf.copy(), 1 - w.condition, sim_time=9000.0, re_f=w.re_f)
    asm.set_wells(w2)
    g2 = _no_wells(grid)
    g2.wells = w2
    F2 = asm.residual(d).cpu().numpy()
    Fr = oracle.residual(fl, g2, st)
    assert _row_err(F2, Fr, oracle.residual_scale(fl, g2, st)) <= RTOL
    assert not np.array_equal(F2, F1.cpu().numpy())
    # no wells: the residual without them, bit for bit
    asm.set_wells(None)
    F0 = asm.residual(d).cpu().numpy()
    plain = gpu.Assembler(gpu.Context(fl), fl, _no_wells(grid))
    dp = plain.device_state(st)
    plain.layout(dp)
    assert np.array_equal(F0, plain.residual(dp).cpu().numpy())


def test_example_deck_wells_on_the_device(gpu, oracle):
    """The reference's example deck with its two wells (a gas injector at a
    corner over layers 1-5 and a producer at the opposite corner), at 2 h,
    WellCondition off and on."""
    import os

    from osiris_amd import deck
    d = deck.load(os.path.join(os.path.dirname(__file__), "golden", "test.dat"))
    st = deck.initial_state(d, lambda T, P, z: oracle.flash_batch(d.fluid, T, P, z))
    ctx = gpu.Context(d.fluid)
    for cond in ([0, 0], [1, 1]):
        g = _no_wells(d.grid)
        g.wells = deck.wells_table(d, 7200.0, cond)
        asm = gpu.Assembler(ctx, d.fluid, g)
        dv = asm.device_state(st)
        asm.layout(dv)
        F = asm.residual(dv).cpu().numpy()
        Fr = oracle.residual(d.fluid, g, st)
        assert np.isfinite(Fr).all()
        assert _row_err(F, Fr, oracle.residual_scale(d.fluid, g, st)) <= RTOL, cond
        assert np.count_nonzero(Fr != oracle.residual(d.fluid, d.grid, st)) > 0
        rp, cols, vals = asm.jacobian(dv)
        rpo, colso, valso = oracle.jacobian(d.fluid, g, st)
        assert np.array_equal(cols.cpu().numpy(), colso)
        row = np.repeat(np.arange(len(rpo) - 1), np.diff(rpo))
        rowmax = np.zeros(len(rpo) - 1)
        np.maximum.at(rowmax, row, np.abs(valso))
        v = vals.cpu().numpy()
        assert np.max(np.abs(v - valso) / np.maximum(np.abs(valso), 1e-8 * rowmax[row] + 1e-300)) <= RTOL


def test_reference_jacobian_with_wells(gpu, oracle):
    """The reference-mode Jacobian (hand rules, its own pattern) with wells
    against the oracle's restatement, at the reference-mode bar: pattern
    identical, values within 1e-8 relative floored at 1e-8 of the row."""
    from tests.test_gpu_refjac import _compare
    for name, (fl, grid, st) in _cases(oracle):
        if name == "c0_all_types_raw":
            continue
        asm = gpu.Assembler(gpu.Context(fl), fl, grid)
        d = asm.device_state(st)
        asm.layout(d)
        rp, cols, vals = asm.jacobian(d, reference=True)
        torch.cuda.synchronize()
        g = (rp.cpu().numpy(), cols.cpu().numpy(), vals.cpu().numpy())
        r = oracle.jacobian_ref(fl, grid, st)
        _compare(name, g, r)
        # the wells change the reference-mode Jacobian
        r0 = oracle.jacobian_ref(fl, _no_wells(grid), st)
        assert np.count_nonzero(r[2] != r0[2]) > 0, name
