"""Why BiCGStab with block Jacobi does not converge on the one-day Newton
systems of the synthetic dual-porosity states, pinned on the oracle's own
reduced system (the Schur complement of its exact Jacobian, 8x8x4 grid,
10 components PR, p ~ U(8, 20) MPa independently per cell, the generator of
the Newton workload):

  * the system is solvable -- a sparse direct solve leaves a relative
    residual below 1e-10 -- but its 1-norm condition number exceeds 1e9;
  * the smallest singular direction (inverse iteration on the LU factors)
    concentrates on the oil-saturation unknown of two-phase fracture blocks
    whose pressure is below every face neighbour's: all their phase fluxes
    arrive from upstream neighbours, so their own saturations enter the
    balances only through the fracture accumulation (porosity 0.01) over one
    day.  A CPR pressure stage does not address that direction;
  * with a one-second step the accumulation dominates and the same
    preconditioned BiCGStab (scipy, block Jacobi on the (nc+1)^2 diagonal
    blocks) converges to 1e-8.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from osiris_amd import synth
from tests.asm_cases import frozen, oracle_flash_fn

NX, NY, NZ, NC = 8, 8, 4, 10


def _reduced(oracle, dt):
    fl = synth.fluid_pseudo(NC, synth.PR, 3)
    grid, st = synth.dual_porosity_case(fl, NX, NY, NZ, 7, oracle_flash_fn(oracle, fl), p_range=(8e6, 20e6), dt=dt)
    s = frozen(oracle, fl, grid, st)
    F = oracle.residual(fl, grid, s)
    rp, cols, vals = oracle.jacobian(fl, grid, s)
    blocks, rhs, *_ = oracle.reduced_from_csr(fl, grid, s, rp, cols, vals, F)
    nb, NP = 2 * grid.n, NC + 1
    keys = sorted(blocks)
    rows = np.array([a for a, d in keys])
    cl = np.array([d for a, d in keys])
    bv = np.stack([blocks[k] for k in keys])
    indptr = np.searchsorted(rows, np.arange(nb + 1))
    A = sp.bsr_matrix((bv, cl, indptr), shape=(nb * NP, nb * NP)).tocsr()
    dinv = np.linalg.inv(bv[rows == cl])
    return grid, s, A, rhs.reshape(-1), dinv


def _bicgstab_bj(A, b, dinv, maxit):
    nb, NP = dinv.shape[0], dinv.shape[1]
    M = spla.LinearOperator(A.shape, lambda r: np.einsum("bij,bj->bi", dinv, r.reshape(nb, NP)).reshape(-1))
    x, info = spla.bicgstab(A, b, M=M, rtol=1e-8, atol=0.0, maxiter=maxit)
    return np.linalg.norm(b - A @ x) / np.linalg.norm(b)


def test_one_day_system_is_near_singular_at_fracture_sinks(oracle):
    grid, s, A, b, dinv = _reduced(oracle, 86400.0)
    Ac = A.tocsc()
    lu = spla.splu(Ac)
    x = lu.solve(b)
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-10
    inv = spla.LinearOperator(A.shape, matvec=lu.solve, rmatvec=lambda v: lu.solve(v, trans="T"))
    cond = spla.onenormest(Ac) * spla.onenormest(inv)
    print(f"cond_1 {cond:.2e}")
    assert cond > 1e9
    v = np.random.default_rng(0).standard_normal(A.shape[0])
    for _ in range(6):
        v = lu.solve(v)
        v /= np.linalg.norm(v)
    n, NP = grid.n, NC + 1
    m = np.abs(v).reshape(2 * n, NP)
    top = np.argsort(m.max(axis=1))[::-1][:10]
    assert (m[top] ** 2).sum() > 0.5
    lab = np.concatenate([s.label_m, s.label_f])
    for blk in top:
        c = int(blk) - n
        assert c >= 0 and lab[blk] == 0, blk            # a two-phase fracture block
        assert int(np.argmax(m[blk])) == NC, blk         # its So unknown (x_0..x_7, P, Sw, So)
        iz, r = divmod(c, NX * NY)
        iy, ix = divmod(r, NX)
        nbrs = [o for ok, o in ((ix < NX - 1, c + 1), (ix > 0, c - 1), (iy < NY - 1, c + NX), (iy > 0, c - NX),
                                (iz < NZ - 1, c + NX * NY), (iz > 0, c - NX * NY)) if ok]
        assert all(s.Po_f[o] > s.Po_f[c] for o in nbrs), blk   # a pressure minimum: every neighbour upstream
    # the block-Jacobi BiCGStab makes no progress on it, nor does a CPR
    # two-stage preconditioner (quasi-IMPES pressure equation w = D^-T e_P
    # solved exactly, then block Jacobi on the remainder)
    assert _bicgstab_bj(A, b, dinv, 300) > 1e-2
    nb = 2 * n
    pidx = np.where(lab == 0, NC - 2, NC - 1)       # P among the primary unknowns
    W = dinv[np.arange(nb), pidx, :]
    Ab = A.tobsr(blocksize=(NP, NP))
    rows = np.repeat(np.arange(nb), np.diff(Ab.indptr))
    ap = np.einsum("ki,kij->kj", W[rows], Ab.data)[np.arange(len(rows)), pidx[Ab.indices]]
    Ap = sp.csr_matrix((ap, (rows, Ab.indices)), shape=(nb, nb)).tocsc()
    lup = spla.splu(Ap)
    R = sp.csr_matrix((W.reshape(-1), (np.repeat(np.arange(nb), NP), np.arange(nb * NP))), shape=(nb, nb * NP))
    Pr = sp.csr_matrix((np.ones(nb), (np.arange(nb) * NP + pidx, np.arange(nb))), shape=(nb * NP, nb))

    def cpr(r):
        x1 = Pr @ lup.solve(R @ r)
        rr = r - A @ x1
        return x1 + np.einsum("bij,bj->bi", dinv, rr.reshape(nb, NP)).reshape(-1)

    xc, _ = spla.bicgstab(A, b, M=spla.LinearOperator(A.shape, cpr), rtol=1e-8, atol=0.0, maxiter=300)
    assert np.linalg.norm(b - A @ xc) / np.linalg.norm(b) > 1e-2


def test_one_second_system_converges_with_block_jacobi(oracle):
    _, _, A, b, dinv = _reduced(oracle, 1.0)
    assert _bicgstab_bj(A, b, dinv, 200) <= 1e-8
