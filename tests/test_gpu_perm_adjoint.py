"""Adjoint of the residual with respect to log-permeability against central
finite differences of the oracle's residual (the oracle rebuilds the face
kA/h and the transfer shape factor from the perturbed permeabilities).

Objective: 1/2 |F_b|^2 over a set b of balance rows (component, total and
water rows); cotangent Fbar = F on b.  Bar: each checked gradient entry
within 1e-6 of the largest entry (central differences, h = 1e-6 on ln k).
"""
import copy

import numpy as np
import pytest
import torch

from tests.asm_cases import c0_case, frozen

pytestmark = pytest.mark.gpu


def _balance_mask(fl, grid, st, oracle, which):
    bs, off = oracle.layout(fl, grid, st)
    nc, n = fl.nc, grid.n
    mask = np.zeros(off[-1], bool)
    lab = np.concatenate([st.label_m, st.label_f])
    for blk in range(2 * n):
        k = blk // n
        if which == "fracture" and k == 0:
            continue
        r0 = nc if lab[blk] == 0 else 0
        mask[off[blk] + r0: off[blk] + r0 + nc + 1] = True
    return mask


@pytest.mark.parametrize("which", ["fracture", "both"])
def test_perm_adjoint_vs_central_differences(gpu, oracle, which):
    fl, grid, st = c0_case(oracle, nx=4, ny=3, nz=2, seed=1, sw_m=0.25, sw_f=0.05)
    s = frozen(oracle, fl, grid, st)
    mask = _balance_mask(fl, grid, s, oracle, which)

    def objective(g):
        F = oracle.residual(fl, g, s)
        return 0.5 * float(np.sum(F[mask] ** 2))

    ctx = gpu.Context(fl)
    asm = gpu.Assembler(ctx, fl, grid)
    d = asm.device_state(s)
    asm.layout(d)
    F = asm.residual(d)
    Fbar = torch.where(torch.as_tensor(mask, device="cuda"), F, torch.zeros_like(F))
    gm, gf = asm.perm_vjp(d, Fbar)
    gm, gf = gm.cpu().numpy(), gf.cpu().numpy()
    scale = max(np.abs(gm).max(), np.abs(gf).max())
    assert scale > 0
    h = 1e-6
    rng = np.random.default_rng(3)
    n = grid.n
    picks = [("perm_f", dd, c) for dd in range(3) for c in rng.choice(n, 3, replace=False)]
    picks += [("perm_m", dd, c) for dd in range(3) for c in rng.choice(n, 3, replace=False)]
    for name, dd, c in picks:
        gp = copy.deepcopy(grid)
        gn = copy.deepcopy(grid)
        getattr(gp, name)[dd, c] *= np.exp(h)
        getattr(gn, name)[dd, c] *= np.exp(-h)
        fd = (objective(gp) - objective(gn)) / (2 * h)
        an = (gm if name == "perm_m" else gf)[dd, c]
        assert abs(an - fd) <= 1e-6 * scale, (name, dd, c, an, fd, scale)
