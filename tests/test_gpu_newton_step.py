"""One time step's nonlinear solve on the device (osr_newton_step_device):
SolveNonLinear (compsim/PetscAUX.cpp:566-730, SNESKSPONLY) and, when the
step is accepted, UpdateAll(X, true) with UpdateDependencies(true)
(PetscAUX.cpp:176-507): X -> state, PhaseCtrl/fPhaseCtrl with the gas
pressures of X0's sweep, CalcPCoeff, PJac and the new state's sweep.

The checker is the oracle's chain on the same inputs: InitGuess2, the full
reference-layout Jacobian and residual with a sparse direct solve, then
UpdateAll (or_unpack) on X0 + du, PhaseCtrl (or_phase_ctrl) with Pg of the
oracle's sweep at X0, the largest pressure and PJac of the new labels.

Bars: the update solves the oracle's full Newton system to the BiCGStab
tolerance (1e-5 of max |du|, the agreement the block-Jacobi solve reaches on
this system); vnorm = |du|_1 / nrows to 1e-13 of the GPU update's own norm;
UpdateAll from the GPU's update bitwise equal to the oracle's (pressures);
the step's PhaseCtrl bitwise equal to the GPU's PhaseCtrl run alone on the
oracle's UpdateAll; against the oracle's PhaseCtrl, the bars of its parity
test (cells within 1e-6 of the 0.95 threshold, or whose saturation-pressure
Newton ends differently on the two sides, may switch on one side only; at
most one such cell per continuum here), saturations 1e-12, compositions of
switched cells 1e-6 (the saturation-pressure Newton), of the others bitwise; PCoeff
the oracle's maximum exactly; the new layout the oracle's PJac; the new
matrix densities (the next step's rho_prev) within 1e-12 of the oracle's
sweep of the GPU's new state; the next step's residual at the new state
within 1e-8 of its row scale.  Rejections leave the state bitwise
unchanged: vnorm above TOLMAIN (a one-day step, vnorm 0.42) and a linear
solve that stops at its iteration cap (SNES_DIVERGED_LINEAR_SOLVE).
"""
import copy

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

from tests.asm_cases import c0_case, frozen

pytestmark = pytest.mark.gpu

FIELDS = ("label_m", "label_f", "Po_m", "Po_f", "x_m", "y_m", "x_f", "y_f", "S_m", "S_f")


def _state(oracle, dt):
    fl, grid, st = c0_case(oracle, nx=4, ny=3, nz=2, seed=1, sw_m=0.25, sw_f=0.05)
    grid.dt = dt
    s = frozen(oracle, fl, grid, st)
    s.pre_m = s.pre_f = None          # preProp of the state itself, as Pre_Mul leaves it
    return fl, grid, s


def _snapshot(d):
    return {k: d[k].clone() for k in FIELDS}


def _host_state(s, d):
    h = copy.copy(s)
    for k in FIELDS:
        setattr(h, k, d[k].cpu().numpy().reshape(np.shape(getattr(s, k))))
    return h


@pytest.mark.parametrize("dt, tolmain", [(1.0, 1e-2), (86400.0, 1.0)])
def test_accepted_step_matches_the_oracle_chain(gpu, oracle, dt, tolmain):
    """dt = 1 s passes the reference's TOLMAIN (vnorm 0.0089); the one-day
    step (vnorm 0.42) is accepted under a looser TOLMAIN and moves cells
    across PhaseCtrl's thresholds."""
    fl, grid, s = _state(oracle, dt=dt)
    ctx = gpu.Context(fl)
    asm = gpu.Assembler(ctx, fl, grid)
    d = asm.device_state(s)
    asm.layout(d)
    p0 = asm.props(d)
    pg0 = p0[3].clone()                        # the gas pressures of X0's sweep on the GPU
    rho0 = torch.cat([p0[5, :grid.n], p0[6, :grid.n]]).clone()   # its matrix densities
    nw = gpu.NewtonStep(asm)
    res = nw.step(d, tolmain=tolmain, rtol=1e-12, maxit=2000)
    torch.cuda.synchronize()
    print(res, flush=True)
    assert res["accepted"] and res["snes_reason"] == "converged_its", res
    assert res["ksp"]["reason"] in ("converged_rtol", "converged_atol"), res

    # the oracle's Newton system at X0
    X0 = oracle.pack(fl, grid, s)
    F = oracle.residual(fl, grid, s)
    rp, cols, vals = oracle.jacobian(fl, grid, s)
    J = sp.csr_matrix((vals, cols, rp), shape=(len(F), len(F)))
    du_ref = spla.spsolve(J.tocsc(), -F)
    du = nw.solution_update().cpu().numpy()
    assert res["nrows"] == len(X0) == len(du)
    assert np.abs(du - du_ref).max() <= 1e-5 * np.abs(du_ref).max()
    vn = np.abs(du).sum() / len(du)
    assert abs(res["vnorm"] - vn) <= 1e-13 * vn and vn <= tolmain

    # UpdateAll from the GPU's update, then PhaseCtrl with X0's gas pressures
    s1 = oracle.unpack(fl, grid, s, X0 + du)
    pr0 = oracle.props(fl, grid, s)
    got = _host_state(s, d)
    switched = 0
    for sfx in ("_m", "_f"):
        assert np.array_equal(getattr(got, "Po" + sfx), getattr(s1, "Po" + sfx)), sfx
    n = grid.n
    for k, sfx in enumerate(("_m", "_f")):
        lab0 = getattr(s1, "label" + sfx)
        args = [getattr(s1, a + sfx) for a in ("Po", "x", "y", "S")]
        L2, x2, y2, S2, ps = oracle.phase_ctrl(fl, grid.T, lab0, args[0], pr0[k, :, 3], *args[1:])
        # the same PhaseCtrl on the GPU, alone, from the oracle's UpdateAll
        # and the GPU's X0 gas pressures: the step must have done exactly this
        gl = torch.as_tensor(lab0.copy()).cuda()
        gx, gy, gS = (torch.as_tensor(np.ascontiguousarray(a)).cuda() for a in args[1:])
        gps = gpu.phase_ctrl(ctx, grid.T, gl, args[0], pg0[k * n:(k + 1) * n].cpu().numpy(), gx, gy, gS)
        gps = gps.cpu().numpy()
        assert np.array_equal(getattr(got, "label" + sfx), gl.cpu().numpy()), sfx
        for name, t in (("x", gx), ("y", gy), ("S", gS)):
            assert np.array_equal(getattr(got, name + sfx).reshape(-1), t.cpu().numpy().reshape(-1)), (name, sfx)
        # against the oracle with the PhaseCtrl parity bars: cells whose
        # saturation test lands within 1e-6 of the threshold, or whose
        # saturation-pressure Newton ends differently, may switch on one side
        test_p = np.where(lab0 == -1, pr0[k, :, 3], args[0])
        both = (ps > 0) & (gps > 0)
        near = np.zeros(n, bool)
        near[both] = np.abs(test_p[both] / (0.95 * ps[both]) - 1) < 1e-6
        unsure = near | ((ps > 0) != (gps > 0))
        gl = getattr(got, "label" + sfx)
        for c in np.nonzero((gl != L2) | unsure)[0]:
            print(sfx, "cell", c, "label", lab0[c], "gpu", gl[c], "oracle", L2[c], "psat gpu", gps[c],
                  "oracle", ps[c], flush=True)
        assert unsure.sum() <= 1, sfx
        ok = ~unsure
        assert np.array_equal(gl[ok], L2[ok]), sfx
        np.testing.assert_allclose(getattr(got, "S" + sfx)[:, ok], S2[:, ok], rtol=1e-12, atol=1e-15)
        sw = ok & (L2 != lab0)
        switched += int(sw.sum())
        print(sfx, "cells switched by PhaseCtrl", int(sw.sum()), flush=True)
        for name, ref in (("x", x2), ("y", y2)):
            g = getattr(got, name + sfx).reshape(ref.shape)
            np.testing.assert_allclose(g[:, sw], ref[:, sw], rtol=1e-6, atol=1e-12)
            keep = ok & (L2 == lab0)
            assert np.array_equal(g[:, keep], ref[:, keep]), (name, sfx)

    assert switched > 0 or dt < 86400.0

    # CalcPCoeff and PJac of the new state
    assert d["pcoeff"] == max(0.0, got.Po_m.max(), got.Po_f.max())
    got.pcoeff = d["pcoeff"]
    bs, off = oracle.layout(fl, grid, got)
    assert res["nrows_next"] == off[-1]

    # the new state's sweep (CalcIFT reads X0's densities) -> the next rho_prev
    got.rho_prev_m = s.rho_prev_m
    pr1 = oracle.props(fl, grid, got)
    rho = d["rho_prev_m"].cpu().numpy().reshape(2, -1)
    np.testing.assert_allclose(rho[0], pr1[0, :, 5], rtol=1e-12)
    np.testing.assert_allclose(rho[1], pr1[0, :, 6], rtol=1e-12)
    # the step's fused PhaseCtrl-and-sweep kernel gave bitwise the property
    # sweep of the new state run on its own (with X0's densities)
    d2 = dict(d)
    d2["rho_prev_m"] = rho0
    p1 = asm.props(d2)
    assert torch.equal(d["rho_prev_m"].reshape(-1), torch.cat([p1[5, :grid.n], p1[6, :grid.n]]))

    # the next step starts from that state: its residual is the oracle's
    got.rho_prev_m = rho
    F1 = asm.residual(d).cpu().numpy()
    F1_ref = oracle.residual(fl, grid, got)
    scale = oracle.residual_scale(fl, grid, got)
    assert np.array_equal(np.isfinite(F1), np.isfinite(F1_ref))
    err = np.abs(F1 - F1_ref) / np.maximum(scale, 1e-14 * np.nanmax(scale) + 1e-300)
    assert err[np.isfinite(F1_ref)].max() <= 1e-8, err.max()
    nw.close()


def test_rejected_steps_leave_the_state(gpu, oracle):
    # a one-day step: the update's mean |du| (0.42) is above TOLMAIN
    fl, grid, s = _state(oracle, dt=86400.0)
    ctx = gpu.Context(fl)
    asm = gpu.Assembler(ctx, fl, grid)
    d = asm.device_state(s)
    before, pc = _snapshot(d), d["pcoeff"]
    rho_before = d["rho_prev_m"].clone()
    nw = gpu.NewtonStep(asm)
    res = nw.step(d, rtol=1e-12, maxit=2000)
    torch.cuda.synchronize()
    print(res, flush=True)
    assert res["snes_reason"] == "converged_its" and res["vnorm"] > 1e-2 and not res["accepted"]
    assert res["nrows_next"] == res["nrows"]
    for k in FIELDS:
        assert torch.equal(d[k], before[k]), k
    assert d["pcoeff"] == pc and torch.equal(d["rho_prev_m"], rho_before)

    # the linear solve stops at its cap: SNES_DIVERGED_LINEAR_SOLVE, rejected
    fl, grid, s = _state(oracle, dt=1.0)
    asm2 = gpu.Assembler(ctx, fl, grid)
    d = asm2.device_state(s)
    before = _snapshot(d)
    nw2 = gpu.NewtonStep(asm2)
    res = nw2.step(d, rtol=1e-12, maxit=1)
    torch.cuda.synchronize()
    assert res["ksp"]["reason"] == "diverged_its" and res["snes_reason"] == "diverged_linear_solve"
    assert not res["accepted"]
    for k in FIELDS:
        assert torch.equal(d[k], before[k]), k
    # two steps from the same state are bitwise identical
    r_a = nw2.step(d, rtol=1e-12, maxit=2000)
    du_a = nw2.solution_update()
    d2 = asm2.device_state(s)
    r_b = nw2.step(d2, rtol=1e-12, maxit=2000)
    du_b = nw2.solution_update()
    torch.cuda.synchronize()
    assert r_a["accepted"] and r_b["accepted"] and r_a["vnorm"] == r_b["vnorm"]
    assert torch.equal(du_a, du_b)
    for k in FIELDS:
        assert torch.equal(d[k], d2[k]), k
    nw.close()
    nw2.close()
