"""Pins of the oracle's restatement of the reference's hand-written Jacobian
(IteratorJacobian, compsim/Petsc_Functions.cpp:6898-42073; oracle function
or_jacobian_ref).

The reference differentiates its residual term by term by hand.  Where its
rules are the true derivative, or_jacobian_ref must equal the oracle's exact
forward-mode Jacobian (itself pinned by finite differences in
test_oracle_assembly.py); where they are not, the difference must follow the
rule the reference applies, and nothing else may differ.  The checks:

  * pattern: the column set of every row is the one the reference's row code
    emits (own block / six neighbours / the other continuum's block, the
    leading diagonal placeholder of :7004-7006), and the count is PJac's
    preallocation arithmetic (Init_Calcs.cpp:1020-1105) corrected for
    single-phase neighbours and the placeholders;
  * configuration A (no capillary pressure, no diffusion, one layer so no
    gravity, IFT scaling off): the reference Jacobian equals the exact one
    with the sign of every neighbour-saturation column flipped (the row code
    emits +dRSide/dS there, e.g. :30075-30081) except at the entries of the
    listed copy-paste rules, and the fracture's kr = S rule (fCalc_Transes,
    d kro / d Sw = d kro / d Sg = -1) holds as an identity between columns;
  * the matrix-fracture transfer: switching the shape factor on changes the
    reference Jacobian by exactly the change of the exact one, with the sign
    flipped on the own pressure column and every column of the other
    continuum (the transfer's P-derivatives are added to dRsidedp, and the
    other continuum's columns are emitted as +dRSide/d(unknown), e.g.
    :31080-31206);
  * fugacity and constraint rows equal the exact derivative (no capillary
    pressure: the reference's dE-based dZ, :30339-30405, is exact).

Not pinned here (the parts where the reference's rules are neither the exact
derivative nor a sign rule of it, so only the restatement itself stands for
them): the inter-cell Fickian and diffusive-transfer derivatives (with their
1e4-scaled DifDiffusion), the gravity terms of the z faces, and the
capillary-pressure / IFT terms the reference leaves out.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from tests.asm_cases import c0_case, frozen

TOL = 1e-9


def _case(oracle, seed, all_two_phase, sigma=0.0, nx=6, ny=5):
    fl, grid, st = c0_case(oracle, nx=nx, ny=ny, nz=1, seed=seed)
    grid.swt = grid.swt.copy()
    grid.swt[:, 3] = 0
    grid.sgt = grid.sgt.copy()
    grid.sgt[:, 3] = 0
    grid.sigma = sigma
    grid.const_diff = True
    grid.cdiff = 0.0
    n, nc = grid.n, fl.nc
    if all_two_phase:
        g = np.random.default_rng(seed)
        for k in "mf":
            setattr(st, "label_" + k, np.zeros(n, np.int8))
            setattr(st, "x_" + k, g.dirichlet(np.full(nc, 2.0), n).T.copy())
            setattr(st, "y_" + k, g.dirichlet(np.full(nc, 2.0), n).T.copy())
            S = getattr(st, "S_" + k).copy()
            S[2] = g.uniform(0.05, 0.3, n) * (1 - S[0])
            S[1] = 1 - S[0] - S[2]
            setattr(st, "S_" + k, S)
    # large previous-sweep densities: parachor IFT >> 1, no kr scaling
    st.rho_prev_m = np.stack([np.full(n, 1e5), np.full(n, 1e2)])
    return fl, grid, st, frozen(oracle, fl, grid, st)


def _csr(rp, c, v):
    m = len(rp) - 1
    return sp.csr_matrix((v, c, rp), shape=(m, m))


class Layout:
    """Row and column classes of the reference layout (PJac, InitGuess2)."""

    def __init__(self, oracle, fl, grid, st, s):
        self.bs, self.off = oracle.layout(fl, grid, s)
        self.n, self.nc, self.nx, self.ny = grid.n, fl.nc, grid.nx, grid.ny
        self.labs = np.concatenate([st.label_m, st.label_f]).astype(int)

    def block(self, i):
        return np.searchsorted(self.off, i, side="right") - 1

    def row_class(self, r):
        b = self.block(r)
        lr = r - self.off[b]
        nc = self.nc
        if self.labs[b] == 0:
            if lr < nc:
                return b, "fug"
            if lr < 2 * nc - 1:
                return b, "comp"
            return b, ["tot", "wat", "sx", "sy", "ss"][lr - (2 * nc - 1)]
        if lr < nc - 1:
            return b, "comp"
        return b, ["tot", "wat", "sx", "ss"][lr - (nc - 1)]

    def col_class(self, b, col):
        """(relation of the column's block to row block b, column kind)."""
        cb = self.block(col)
        lc = col - self.off[cb]
        k, c = divmod(b, self.n)
        kb, cc = divmod(cb, self.n)
        rel = {0: "own", 1: "x+", -1: "x-", self.nx: "y+", -self.nx: "y-",
               self.nx * self.ny: "z+", -self.nx * self.ny: "z-"}[cc - c] if kb == k else "other"
        nl = self.labs[cb]
        p = 2 * self.nc if nl == 0 else self.nc
        if lc == p:
            kind = "P"
        elif lc < p:
            kind = "x"
        else:
            kind = ["Sw", "So" if nl != -1 else "Sg", "Sg"][lc - p - 1]
        return cb, rel, kind

    def neighbours(self, b):
        k, c = divmod(b, self.n)
        iz, rem = divmod(c, self.nx * self.ny)
        iy, ix = divmod(rem, self.nx)
        out = []
        if ix + 1 < self.nx: out.append(c + 1)
        if ix > 0: out.append(c - 1)
        if iy + 1 < self.ny: out.append(c + self.nx)
        if iy > 0: out.append(c - self.nx)
        return [k * self.n + o for o in out]

    def expected_columns(self, r):
        """Columns the reference's row code emits for row r."""
        b, rc = self.row_class(r)
        nc, off, bs, lab = self.nc, self.off, self.bs, self.labs[b]
        o = off[b]
        k, c = divmod(b, self.n)
        other = (1 - k) * self.n + c
        p = 2 * nc if lab == 0 else nc

        def pw(blk):
            q = 2 * nc if self.labs[blk] == 0 else nc
            return {off[blk] + q, off[blk] + q + 1}

        if rc == "fug":
            return set(range(o, o + 2 * nc + 1))
        if rc in ("comp", "tot"):
            cols = set(range(o, o + bs[b]))
            if lab == -1:
                cols.discard(o + p + 1)   # a gas block's own Sw is not emitted
            for nb in self.neighbours(b) + [other]:
                cols |= set(range(off[nb], off[nb] + bs[nb]))
            return cols
        if rc == "wat":
            cols = pw(b)
            for nb in self.neighbours(b) + [other]:
                cols |= pw(nb)
            return cols
        lr = r - o
        if rc == "sx":
            return set(range(o, o + nc)) | {o + lr}
        if rc == "sy":
            return set(range(o + nc, o + 2 * nc)) | {o + lr}
        return {o + p + 1, o + p + 2} | ({o + p + 3} if lab == 0 else set())


def pjac_total(lay, grid):
    """totalJac of PJac (Init_Calcs.cpp:1020-1105)."""
    nc, n = lay.nc, lay.n
    blockN = 18 * nc * nc + 35 * nc + 19
    face = 2 * nc * nc + 4 * nc + 2
    tot = 0
    for k in (0, 1):
        for c in range(n):
            iz, rem = divmod(c, grid.nx * grid.ny)
            iy, ix = divmod(rem, grid.nx)
            tot += blockN
            tot -= face * ((ix == 0) + (ix == grid.nx - 1) + (iy == 0) + (iy == grid.ny - 1)
                           + (iz == 0) + (iz == grid.nz - 1))
            lab = lay.labs[k * n + c]
            if lab == 1:
                tot -= 3 * nc * nc + 3 * nc + 1
            elif lab == -1:
                tot -= 3 * nc * nc + 4 * nc + 1
            if lay.labs[(1 - k) * n + c] != 0:
                tot -= nc * nc + nc
    return tot


@pytest.mark.parametrize("seed", [0, 4])
def test_pattern_is_the_reference_row_code(oracle, seed):
    fl, grid, st = c0_case(oracle, nx=10, ny=10, nz=1, seed=seed)
    s = frozen(oracle, fl, grid, st)
    rp, cols, vals = oracle.jacobian_ref(fl, grid, s)
    lay = Layout(oracle, fl, grid, st, s)
    assert np.all(np.isfinite(vals))
    for r in range(len(rp) - 1):
        got = cols[rp[r]:rp[r + 1]]
        assert np.all(np.diff(got) > 0), f"row {r}: columns not sorted/unique"
        exp = lay.expected_columns(r)
        assert set(got.tolist()) == exp, (r, lay.row_class(r), sorted(set(got.tolist()) ^ exp))
    # PJac assumes two-phase neighbours (2nc+4 columns each) and counts no
    # placeholder: each single-phase neighbour has nc+1 fewer columns in each
    # of the nc balance rows, and the Sigma-x (and Sigma-y) rows carry the
    # diagonal placeholder (one for a single-phase block, two otherwise).
    corr = 0
    for b in range(2 * grid.n):
        corr += 2 if lay.labs[b] == 0 else 1
        for nb in lay.neighbours(b):
            if lay.labs[nb] != 0:
                corr -= lay.nc * (lay.nc + 1)
    assert len(cols) == pjac_total(lay, grid) + corr


def _r1(B, lay):
    """Flip the sign of every neighbour-saturation column."""
    Bc = B.tocoo()
    sgn = np.ones(Bc.nnz)
    for t, (r, c) in enumerate(zip(Bc.row, Bc.col)):
        b, _ = lay.row_class(r)
        _, rel, kind = lay.col_class(b, c)
        if rel not in ("own", "other") and kind[0] == "S":
            sgn[t] = -1.0
    return sp.csr_matrix((Bc.data * sgn, (Bc.row, Bc.col)), shape=B.shape)


def _rowscale(*Ms):
    return np.max([abs(M).max(axis=1).toarray().ravel() for M in Ms], axis=0)


@pytest.mark.parametrize("seed,all2p", [(3, True), (4, False), (5, False)])
def test_configuration_a_equals_exact_up_to_the_reference_rules(oracle, seed, all2p):
    fl, grid, st, s = _case(oracle, seed, all2p)
    A = _csr(*oracle.jacobian_ref(fl, grid, s))
    lay = Layout(oracle, fl, grid, st, s)
    B = _r1(_csr(*oracle.jacobian(fl, grid, s)), lay)
    D = (A - B).tocoo()
    scale = _rowscale(A, B)
    bad = np.abs(D.data) > TOL * scale[D.row]
    hit = {"comp_xm_P": 0, "tot_ym_P": 0}
    for r, c in zip(D.row[bad], D.col[bad]):
        b, rc = lay.row_class(r)
        cb, rel, kind = lay.col_class(b, c)
        lab, nlab = lay.labs[b], lay.labs[cb]
        # copy-paste rules of the face code: the x- face's gas term assigns
        # the neighbour-P accumulator of the component rows (:31466, :31473);
        # the y- face of the total row assigns it twice
        if rc == "comp" and rel == "x-" and kind == "P":
            hit["comp_xm_P"] += 1
            continue
        if rc == "tot" and rel == "y-" and kind == "P":
            hit["tot_ym_P"] += 1
            continue
        # fracture kr = S with d kro / d Sw = d kro / d Sg = -1 (checked below)
        if b >= lay.n and kind in ("Sw", "Sg") and rc in ("comp", "tot"):
            continue
        # single-phase rows: the absent phase's own density and viscosity
        # enter the face averages' derivatives (Coef multiplies the
        # neighbour only), so the own P column and the columns of a
        # differently labelled neighbour differ
        if lab != 0 and rc in ("comp", "tot") and (
                (rel == "own" and kind == "P") or (rel not in ("own", "other") and nlab != lab)):
            continue
        # two-phase rows next to a gas-only neighbour: the composition
        # columns take the gas Coef from (neighbour label == 0)
        if lab == 0 and rc == "comp" and kind == "x" and (nlab == -1 or any(
                lay.labs[nb] == -1 for nb in lay.neighbours(b))):
            continue
        raise AssertionError(f"row {r} {rc} (block {b}, label {lab}) col {c} {rel} {kind}: "
                             f"ref {A[r, c]!r} exact {B[r, c]!r}")
    assert hit["comp_xm_P"] > 0 and hit["tot_ym_P"] > 0
    # fracture kr rule as an identity: against the exact derivative, a
    # neighbour's Sw and Sg columns both gain minus its So column (the oil
    # kr term), and the own Sw and Sg columns gain the same amount
    nc = lay.nc
    for b in range(lay.n, 2 * lay.n):
        if lay.labs[b] != 0:
            continue
        for lr in range(nc, 2 * nc):
            r = lay.off[b] + lr
            sc = scale[r]
            for cb in [b] + [nb for nb in lay.neighbours(b) if lay.labs[nb] == 0]:
                o = lay.off[cb]
                Sw, So, Sg = o + 2 * nc + 1, o + 2 * nc + 2, o + 2 * nc + 3
                dw, dg = A[r, Sw] - B[r, Sw], A[r, Sg] - B[r, Sg]
                if cb == b:
                    assert abs(dw - dg) <= TOL * sc
                else:
                    assert abs(dw + B[r, So]) <= TOL * sc and abs(dg + B[r, So]) <= TOL * sc


@pytest.mark.parametrize("seed,all2p", [(3, True), (4, False)])
def test_transfer_terms_follow_the_sign_rule(oracle, seed, all2p):
    fl, grid, st, s0 = _case(oracle, seed, all2p, sigma=0.0)
    A0 = _csr(*oracle.jacobian_ref(fl, grid, s0))
    B0 = _csr(*oracle.jacobian(fl, grid, s0))
    fl, grid, st, s1 = _case(oracle, seed, all2p, sigma=1.0 / 3.0)
    A1 = _csr(*oracle.jacobian_ref(fl, grid, s1))
    B1 = _csr(*oracle.jacobian(fl, grid, s1))
    lay = Layout(oracle, fl, grid, st, s1)
    dA = (A1 - A0).tocsr()
    dB = (_r1(B1, lay) - _r1(B0, lay)).tocoo()
    sgn = np.ones(dB.nnz)
    for t, (r, c) in enumerate(zip(dB.row, dB.col)):
        b, _ = lay.row_class(r)
        _, rel, kind = lay.col_class(b, c)
        if rel == "other" or (rel == "own" and kind == "P"):
            sgn[t] = -1.0
    pred = sp.csr_matrix((dB.data * sgn, (dB.row, dB.col)), shape=dA.shape)
    scale = _rowscale(A1)
    assert abs(dA).max() > 0
    D = (dA - pred).tocoo()
    bad = np.abs(D.data) > TOL * scale[D.row]
    assert not bad.any(), [(int(r), int(c)) for r, c in zip(D.row[bad][:5], D.col[bad][:5])]


@pytest.mark.parametrize("seed", [0, 4])
def test_fugacity_and_constraint_rows_are_exact(oracle, seed):
    fl, grid, st = c0_case(oracle, nx=10, ny=10, nz=1, seed=seed)
    grid.swt = grid.swt.copy()
    grid.swt[:, 3] = 0
    grid.sgt = grid.sgt.copy()
    grid.sgt[:, 3] = 0
    st.rho_prev_m = np.stack([np.full(grid.n, 1e5), np.full(grid.n, 1e2)])
    s = frozen(oracle, fl, grid, st)
    A = _csr(*oracle.jacobian_ref(fl, grid, s))
    B = _csr(*oracle.jacobian(fl, grid, s))
    lay = Layout(oracle, fl, grid, st, s)
    rows = [r for r in range(A.shape[0]) if lay.row_class(r)[1] in ("fug", "sx", "sy", "ss")]
    assert len(rows) > 100
    D = (A - B)[rows]
    scale = _rowscale(A[rows], B[rows])
    assert (abs(D).max(axis=1).toarray().ravel() <= 1e-12 * scale).all()
