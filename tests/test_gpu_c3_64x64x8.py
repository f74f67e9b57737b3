"""The differentiation configuration at its own size: a 64x64x8 dual-porosity
grid (65,536 blocks), 6 components PR, p ~ U(8, 20) MPa, against the oracle:

  * the GPU's exact CSR Jacobian against central finite differences of the
    oracle's residual (h = 1e-6 relative, pressure in PCoeff units) on
    sampled columns of every block class (matrix and fracture, two-phase and
    single-phase; composition, pressure and saturation columns): within
    1e-5 of the column's largest entry, nothing outside the emitted pattern
    above 1e-6 -- the bar of the oracle's own finite-difference test;
  * the GPU's reference-mode Jacobian (IteratorJacobian's pattern and rules)
    against the oracle's restatement on the whole grid: identical row
    pointers and columns; every entry within 1e-8 relative, floored at 1e-8
    of its row's largest entry (the bar of the smaller reference-mode
    cases).  An entry over that bar must lie within 8x of what the oracle's
    own entry moves when either continuum's pressure moves by one ulp: near
    a cubic's inflection point the root's dZ/dP is ill-conditioned, and in
    such cells the oracle's entries move by up to 3e-5 relative under a
    one-ulp input change, so the GPU's different rounding of the same
    operations moves them as far.  No entry may differ by more than 1e-8 of
    its row's largest entry;
  * the permeability adjoint of 1/2 |F_b|^2 over the fracture balance rows
    against central differences of the oracle's residual in ln k (h = 1e-6)
    on sampled fracture and matrix cells: within 1e-6 of the largest
    gradient entry.
Water saturations sit off the tables' nodes (0.25 matrix, 0.05 fracture) so
a finite difference sees one table segment."""
import copy

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from osiris_amd import synth
from tests.asm_cases import column_variable, frozen, oracle_flash_fn

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c3(oracle):
    fl = synth.fluid_pseudo(6, synth.PR, 3)
    grid, st = synth.dual_porosity_case(fl, 64, 64, 8, 0, oracle_flash_fn(oracle, fl), p_range=(8e6, 20e6),
                                        sw_m=0.25, sw_f=0.05)
    return fl, grid, frozen(oracle, fl, grid, st)


def _replace(st, **kw):
    s = copy.copy(st)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def test_exact_jacobian_vs_oracle_finite_differences(gpu, oracle, c3):
    fl, grid, s = c3
    print("C3 64x64x8: GPU CSR Jacobian vs oracle central differences", flush=True)
    ctx = gpu.Context(fl)
    asm = gpu.Assembler(ctx, fl, grid)
    d = asm.device_state(s)
    asm.layout(d)
    rp, cols, vals = asm.jacobian(d)
    torch.cuda.synchronize()
    rp, cols, vals = rp.cpu().numpy(), cols.cpu().numpy(), vals.cpu().numpy()
    asm.close()
    bs, off = oracle.layout(fl, grid, s)
    N = int(off[-1])
    Jc = sp.csr_matrix((vals, cols, rp), shape=(N, N)).tocsc()
    n, nc = grid.n, fl.nc
    lab = np.concatenate([s.label_m, s.label_f])
    rng = np.random.default_rng(7)
    picks = []
    for k in (0, 1):
        for want in (0, 1):
            blks = np.flatnonzero((lab[k * n:(k + 1) * n] != 0) == bool(want)) + k * n
            if blks.size == 0:
                continue
            b = int(rng.choice(blks))
            p = 2 * nc if lab[b] == 0 else nc
            for j in (0, p, p + 1, int(bs[b]) - 1):
                picks.append(int(off[b]) + j)
    assert len(picks) >= 12
    for col in picks:
        name, idx = column_variable(s, nc, n, bs, off, col)
        arr = getattr(s, name)
        base = arr[idx]
        h = 1e-6 * base if name.startswith("Po") else 1e-6 * max(abs(base), 1e-3)

        def F_at(v):
            a = np.array(arr, copy=True)
            a[idx] = v
            return oracle.residual(fl, grid, _replace(s, **{name: a}))

        fd = (F_at(base + h) - F_at(base - h)) / (2 * h)
        if name.startswith("Po"):
            fd *= s.pcoeff
        jcol = Jc[:, col].toarray().ravel()
        scale = max(np.abs(fd).max(), np.abs(jcol).max(), 1e-300)
        err = np.abs(fd - jcol).max() / scale
        print(f"  column {col} ({name}{idx}): {err:.2e}", flush=True)
        assert err <= 1e-5, (col, name, idx, err)
        outside = np.ones(N, bool)
        outside[Jc[:, col].indices] = False
        assert np.abs(fd[outside]).max(initial=0) <= 1e-6 * scale, (col, name)
    ctx.close()


def test_reference_jacobian_whole_grid(gpu, oracle, c3):
    fl, grid, s = c3
    print("C3 64x64x8: GPU reference-mode Jacobian vs the oracle's restatement", flush=True)
    ctx = gpu.Context(fl)
    asm = gpu.Assembler(ctx, fl, grid)
    d = asm.device_state(s)
    asm.layout(d)
    rp, cols, vals = asm.jacobian(d, reference=True)
    torch.cuda.synchronize()
    grp, gcols, gvals = rp.cpu().numpy(), cols.cpu().numpy(), vals.cpu().numpy()
    asm.close()
    orp, ocols, ovals = oracle.jacobian_ref(fl, grid, s, nthreads=0)
    assert np.array_equal(grp, orp) and np.array_equal(gcols, ocols)
    assert np.isfinite(ovals).all() and np.isfinite(gvals).all()
    rowmax = np.repeat(np.maximum.reduceat(np.abs(ovals), orp[:-1]), np.diff(orp))
    d = np.abs(gvals - ovals)
    rowwise = float((d / rowmax).max())
    big = np.abs(ovals) >= 1e-6 * rowmax
    rel_big = float((d[big] / np.abs(ovals[big])).max())
    floored = d / np.maximum(np.abs(ovals), 1e-8 * rowmax)
    w = int(np.argmax(floored))
    print(f"  nnz {len(ocols)}: row-normwise {rowwise:.2e}; entries above 1e-6 of their row {rel_big:.2e}; "
          f"floored at 1e-8 of the row {floored.max():.2e} (entry {abs(ovals[w]) / rowmax[w]:.1e} of its row)",
          flush=True)
    bs, off = oracle.layout(fl, grid, s)
    lab = np.concatenate([s.label_m, s.label_f])
    rows = np.repeat(np.arange(len(orp) - 1), np.diff(orp))
    for e in np.argsort(floored)[-6:][::-1]:
        r, c = rows[e], ocols[e]
        b = int(np.searchsorted(off, r, side="right") - 1)
        cb = int(np.searchsorted(off, c, side="right") - 1)
        print(f"    row {r} (block {b}, label {lab[b]}, local {r - off[b]}) col {c} (block {cb}, label {lab[cb]}, "
              f"local {c - off[cb]}): gpu {gvals[e]!r} oracle {ovals[e]!r} rel {floored[e]:.2e} "
              f"of-row {abs(ovals[e]) / rowmax[e]:.1e}", flush=True)
    # the oracle's own movement under a one-ulp change of either pressure
    sens = np.zeros_like(ovals)
    eps = np.finfo(np.float64).eps
    for name in ("Po_m", "Po_f"):
        for f in (1 + eps, 1 - eps / 2):
            _, pc, pv = oracle.jacobian_ref(fl, grid, _replace(s, **{name: getattr(s, name) * f}), nthreads=0)
            assert np.array_equal(pc, ocols)
            np.maximum(sens, np.abs(pv - ovals), out=sens)
            del pc, pv
    over = floored > 1e-8
    unexplained = over & (d > 8 * sens)
    print(f"  over the bars {int(over.sum())} entries in {len(np.unique(rows[over]))} rows; "
          f"within 8x the oracle's one-ulp movement: {int((over & ~unexplained).sum())}; "
          f"unexplained {int(unexplained.sum())}", flush=True)
    for e in np.flatnonzero(unexplained)[:6]:
        print(f"    unexplained row {rows[e]} col {ocols[e]}: gpu {gvals[e]!r} oracle {ovals[e]!r} "
              f"one-ulp movement {sens[e]:.2e}", flush=True)
    assert not unexplained.any()
    assert rowwise <= 1e-8
    ctx.close()


def test_perm_adjoint_vs_oracle_finite_differences(gpu, oracle, c3):
    fl, grid, s = c3
    print("C3 64x64x8: permeability adjoint vs oracle central differences", flush=True)
    bs, off = oracle.layout(fl, grid, s)
    nc, n = fl.nc, grid.n
    lab = np.concatenate([s.label_m, s.label_f])
    mask = np.zeros(off[-1], bool)
    for blk in range(n, 2 * n):
        r0 = nc if lab[blk] == 0 else 0
        mask[off[blk] + r0: off[blk] + r0 + nc + 1] = True

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
    asm.close()
    scale = max(np.abs(gm).max(), np.abs(gf).max())
    assert scale > 0
    h = 1e-6
    rng = np.random.default_rng(11)
    # the cells with the largest fracture gradients, and random ones
    top = np.argsort(np.abs(gf).max(axis=0))[-2:]
    picks = [("perm_f", int(np.argmax(np.abs(gf[:, c]))), int(c)) for c in top]
    picks += [("perm_f", dd, int(c)) for dd, c in zip((0, 1, 2), rng.choice(n, 3, replace=False))]
    picks += [("perm_m", dd, int(c)) for dd, c in zip((0, 2), rng.choice(n, 2, replace=False))]
    for name, dd, c in picks:
        gp = copy.deepcopy(grid)
        gn = copy.deepcopy(grid)
        getattr(gp, name)[dd, c] *= np.exp(h)
        getattr(gn, name)[dd, c] *= np.exp(-h)
        fd = (objective(gp) - objective(gn)) / (2 * h)
        an = (gm if name == "perm_m" else gf)[dd, c]
        print(f"  {name}[{dd},{c}]: adjoint {an:.6e} fd {fd:.6e}", flush=True)
        assert abs(an - fd) <= 1e-6 * scale, (name, dd, c, an, fd, scale)
    ctx.close()
