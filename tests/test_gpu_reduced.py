"""Reduced (nc+1)^2 block-CSR: solving the reduced system and expanding the
update must solve the full Newton system J du = -F.

Bars: the normwise backward error of the expanded update,
|J du + F|_inf / (|J|_inf |du|_inf + |F|_inf), below 1e-12, and agreement
with a sparse direct solve of the full system to 1e-7 of max |du|.  (A
componentwise bar is meaningless here: at the initial state many rows --
water rows at zero mobility, constraint rows -- hold only rounding-level
terms, and the row scales span ~1e7.)
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

from osiris_amd import synth
from tests.asm_cases import c0_case, frozen, oracle_flash_fn

pytestmark = pytest.mark.gpu


def _cases(oracle):
    yield "c0_4x3x2", c0_case(oracle, nx=4, ny=3, nz=2, seed=1, sw_m=0.25, sw_f=0.05)
    fl6 = synth.fluid_pseudo(6, synth.PR, 3)
    yield "pr6_6x5x3", (fl6,) + synth.dual_porosity_case(fl6, 6, 5, 3, 4, oracle_flash_fn(oracle, fl6),
                                                          p_range=(8e6, 20e6))
    fl, grid, st = c0_case(oracle, nx=4, ny=4, nz=1, seed=7)
    grid.phi_m[[0, 5]] = 0.0
    grid.phi_f[[3]] = 0.0
    yield "c0_zero_porosity", (fl, grid, st)


def test_reduced_system_solves_the_full_one(gpu, oracle):
    for name, (fl, grid, st) in _cases(oracle):
        s = frozen(oracle, fl, grid, st)
        ctx = gpu.Context(fl)
        asm = gpu.Assembler(ctx, fl, grid)
        d = asm.device_state(s)
        _, nrows = asm.layout(d)
        F = asm.residual(d)
        rp, cols, vals = asm.jacobian(d)
        browptr, bcols, bvals, brhs, S, t = asm.reduced(d, F, rp, vals)
        np1 = fl.nc + 1
        nb = 2 * grid.n
        if name == "c0_zero_porosity":
            # the reference's 0/0 next to zero-porosity cells puts NaN into
            # neighbouring rows, so no solve is meaningful: check the
            # zero-porosity blocks reduce to identity blocks with zero rhs
            br, bc, bv = browptr.cpu().numpy(), bcols.cpu().numpy(), bvals.cpu().numpy()
            for blk in (0, 5, grid.n + 3):
                assert br[blk + 1] - br[blk] == 1 and bc[br[blk]] == blk
                assert np.array_equal(bv[br[blk]], np.eye(np1))
                assert np.all(brhs.cpu().numpy()[blk] == 0)
                assert np.all(S.cpu().numpy()[blk] == 0) and np.all(t.cpu().numpy()[blk] == 0)
            continue
        R = sp.bsr_matrix((bvals.cpu().numpy(), bcols.cpu().numpy().astype(np.int64), browptr.cpu().numpy()),
                          shape=(nb * np1, nb * np1)).tocsc()
        dp = spla.spsolve(R, brhs.cpu().numpy().reshape(-1))
        du = asm.reduced_expand(d, S, t, torch.as_tensor(dp.reshape(nb, np1)).cuda()).cpu().numpy()
        J = sp.csr_matrix((vals.cpu().numpy(), cols.cpu().numpy(), rp.cpu().numpy()), shape=(nrows, nrows))
        Fh = F.cpu().numpy()
        assert np.isfinite(Fh).all(), name
        res = J @ du + Fh
        eta = np.abs(res).max() / (abs(J).sum(axis=1).max() * np.abs(du).max() + np.abs(Fh).max())
        assert eta < 1e-12, (name, eta)
        du_full = spla.spsolve(J.tocsc(), -Fh)
        err = np.abs(du - du_full).max() / np.abs(du_full).max()
        assert err < 1e-7, (name, err)
