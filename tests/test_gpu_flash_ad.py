"""Tangent and adjoint of the flash (implicit-function theorem on the
fugacity equations) against central finite differences of an exact
equilibrium solve, the adjoint against the tangent (dot-product identity),
and the torch.autograd wrapper.

Bars: tangent vs central differences (h = 1e-6 relative) within 1e-5 of the
largest output derivative of the cell; dot-product identity within 1e-9 of
the sum of |terms|; autograd gradients equal the adjoint kernel's.
"""
import numpy as np
import pytest
import torch

from osiris_amd import synth
from tests.flash_exact import solve

pytestmark = pytest.mark.gpu


def _cells(oracle, fl, n, seed):
    cells = synth.flash_cells(fl, n, seed, lambda T, P, k, x, y: oracle.psat_batch(fl, T, P, k, x, y))
    return cells


@pytest.mark.parametrize("nc,eos,seed", [(6, "PR", 3), (8, "PR", 1), (12, "SRK", 2)])
def test_adjoint_matches_tangent(gpu, oracle, nc, eos, seed):
    fl = synth.fluid_pseudo(nc, getattr(synth, eos), seed)
    cells = _cells(oracle, fl, 4000, 31)
    ctx = gpu.Context(fl)
    out = gpu.flash(ctx, cells.T, cells.P, cells.z)
    g = synth.rng(5)
    n = cells.P.shape[0]
    dP = torch.tensor(g.standard_normal(n) * cells.P * 1e-3, device="cuda")
    dz = torch.tensor(g.standard_normal((nc, n)) * cells.z * 1e-3, device="cuda")
    xb = torch.tensor(g.standard_normal((nc, n)), device="cuda")
    yb = torch.tensor(g.standard_normal((nc, n)), device="cuda")
    Lb = torch.tensor(g.standard_normal(n), device="cuda")
    dx, dy, dL = gpu.flash_tangent(ctx, cells.T, cells.P, cells.z, out["K"], out["label"], dP=dP, dz=dz)
    Pb, zb = gpu.flash_adjoint(ctx, cells.T, cells.P, cells.z, out["K"], out["label"], xb, yb, Lb)
    lhs = (xb * dx).sum(0) + (yb * dy).sum(0) + Lb * dL
    rhs = (zb * dz).sum(0) + Pb * dP
    scale = ((xb * dx).abs().sum(0) + (yb * dy).abs().sum(0) + (Lb * dL).abs() +
             (zb * dz).abs().sum(0) + (Pb * dP).abs())
    lab = out["label"].cpu().numpy()
    err = ((lhs - rhs).abs() / scale.clamp_min(1e-300)).cpu().numpy()
    fin = np.isfinite(err)
    assert fin[lab == 0].mean() > 0.999
    assert err[fin].max() < 1e-9, err[fin].max()
    assert (lab == 0).sum() > 100


@pytest.mark.parametrize("nc,eos,seed", [(6, "PR", 3), (12, "SRK", 2)])
def test_tangent_vs_central_differences(gpu, oracle, nc, eos, seed):
    fl = synth.fluid_pseudo(nc, getattr(synth, eos), seed)
    cells = _cells(oracle, fl, 3000, 41)
    ref = oracle.flash_batch(fl, cells.T, cells.P, cells.z)
    two = np.flatnonzero((ref["label"] == 0) & (ref["iters"] < 200) & (ref["L"] > 0.02) & (ref["L"] < 0.98))
    sel = two[:10]
    assert len(sel) == 10
    g = synth.rng(9)
    h = 1e-6
    ctx = gpu.Context(fl)
    for c in sel:
        T, P, z = cells.T[c], cells.P[c], cells.z[:, c].copy()
        u, L0, x0, y0 = solve(oracle, fl, T, P, z, np.log(ref["K"][:, c]))
        for direction in ("P", "z"):
            dP = P if direction == "P" else 0.0
            dz = np.zeros(nc) if direction == "P" else z * g.standard_normal(nc)
            up, Lp, xp, yp = solve(oracle, fl, T, P + h * dP, z + h * dz, u)
            um, Lm, xm, ym = solve(oracle, fl, T, P - h * dP, z - h * dz, u)
            fd = np.concatenate([(xp - xm) / (2 * h), (yp - ym) / (2 * h), [(Lp - Lm) / (2 * h)]])
            dx, dy, dL = gpu.flash_tangent(ctx, np.array([T]), np.array([P]), z[:, None], np.exp(u)[:, None],
                                           np.zeros(1, np.int8), dP=np.array([dP]), dz=dz[:, None])
            an = np.concatenate([dx.cpu().numpy()[:, 0], dy.cpu().numpy()[:, 0], dL.cpu().numpy()])
            err = np.abs(an - fd).max() / np.abs(fd).max()
            assert err < 1e-5, (c, direction, err)


def test_autograd_backward_is_the_adjoint(gpu, oracle):
    from osiris_amd import autograd as ad
    fl = synth.fluid_pseudo(6, synth.PR, 3)
    cells = _cells(oracle, fl, 2000, 51)
    ctx = gpu.Context(fl)
    P = torch.tensor(cells.P, device="cuda", requires_grad=True)
    z = torch.tensor(cells.z, device="cuda", requires_grad=True)
    x, y, L, lab = ad.flash(ctx, cells.T, P, z)
    w = torch.tensor(synth.rng(2).standard_normal(x.shape), device="cuda")
    loss = (w * x).sum() + (L * L).sum()
    gP, gz = torch.autograd.grad(loss, (P, z))
    out = gpu.flash(ctx, cells.T, cells.P, cells.z)
    Pb, zb = gpu.flash_adjoint(ctx, cells.T, cells.P, cells.z, out["K"], out["label"], xbar=w, Lbar=2 * L.detach())
    assert torch.equal(gP, Pb) and torch.equal(gz, zb)
    # and the tangent agrees with the gradient along a direction
    (xo, yo, Lo, lo), (dx, dy, dL) = ad.tangent(ctx, cells.T, cells.P, cells.z, dP=P.detach() * 1e-3)
    lhs = ((w * dx).sum(0) + 2 * Lo * dL)
    rhs = gP * P.detach() * 1e-3
    fin = torch.isfinite(lhs) & torch.isfinite(rhs)
    assert torch.allclose(lhs[fin], rhs[fin], rtol=1e-8, atol=1e-14)
