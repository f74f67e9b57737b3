"""CPU transcription of the reduced system's lane-group elimination
(assembly.hip group_eliminate / group_write_st, as k_red_local calls them),
lane by lane for a wave of two 32-lane groups, checking

  * every cross-lane read (the pivot lane P of each step) stays inside the
    reading group, reads a lane that holds a column of the block (P - gbase <
    block size, never the residual lane 31), and the pivot lanes of one
    block are exactly its secondary columns, each once;
  * the result: S = J_ls^-1 J_lp and t = J_ls^-1 F_l of random local
    systems of both static shapes (two-phase nc+3 rows, single-phase 2),
    including ones that need row exchanges, against numpy.linalg.solve;
  * two groups of different shapes in one wave do not read each other.
"""
import numpy as np
import pytest

GROUP = 32


def red_sec_col(lab, nc, i):
    if lab == 0:
        return nc - 2 + i if i < 2 + nc else 2 * nc + 3
    return nc - 1 if i == 0 else nc + 2


def red_role(lab, nc, j):
    if lab == 0:
        if j < nc - 2:
            return j
        if j < 2 * nc:
            return -1 - (j - (nc - 2))
        if j <= 2 * nc + 2:
            return nc - 2 + (j - 2 * nc)
        return -1 - (nc + 2)
    if j < nc - 1:
        return j
    if j == nc - 1:
        return -1
    if j <= nc + 1:
        return nc - 1 + (j - nc)
    return -2


def eliminate(cols, nc, lab, nl, gbase, bs, reads):
    """cols: [64][nc+3] per-lane columns of the whole wave; only the lanes of
    the group at gbase change.  reads collects (reader group, source lane)."""
    if nl == nc + 3:
        lab = 0
    lanes = range(gbase, gbase + GROUP)
    for q in range(nl):
        P = gbase + red_sec_col(lab, nc, q)
        reads.append((gbase, P))
        piv_of = {}
        for L in lanes:
            piv, best = q, abs(cols[L][q])
            for r in range(q + 1, nl):
                if abs(cols[L][r]) > best:
                    best, piv = abs(cols[L][r]), r
            piv_of[L] = piv
        piv = piv_of[P]
        for L in lanes:
            cols[L][q], cols[L][piv] = cols[L][piv], cols[L][q]
        f = [cols[P][r] for r in range(nl)]
        for L in lanes:
            cn = cols[L][q] * (1.0 / f[q])
            for r in range(nl):
                if r != q:
                    cols[L][r] -= f[r] * cn
            cols[L][q] = cn


def write_st(cols, nc, lab, nl, gbase, bs):
    NLM, NP = nc + 3, nc + 1
    S = np.zeros((NLM, NP))
    t = np.zeros(NLM)
    for j in range(GROUP):
        col = cols[gbase + j]
        if j == GROUP - 1:
            t[:] = [col[r] if r < nl else 0.0 for r in range(NLM)]
        elif j < bs:
            role = red_role(lab, nc, j)
            if role >= 0:
                S[:, role] = [col[r] if r < nl else 0.0 for r in range(NLM)]
    return S, t


def _system(rng, nc, lab, permute):
    bs = 2 * nc + 4 if lab == 0 else nc + 3
    nl = nc + 3 if lab == 0 else 2
    sec = [red_sec_col(lab, nc, i) for i in range(nl)]
    prim = [j for j in range(bs) if red_role(lab, nc, j) >= 0]
    J = rng.standard_normal((nl, bs))
    A = J[:, sec]
    A += 4 * np.eye(nl) if not permute else 0.0
    if permute:
        # put the large entries off the diagonal so partial pivoting swaps rows
        A[:] = np.roll(np.eye(nl) * 8, 1, axis=0) + 0.1 * A
    J[:, sec] = A
    F = rng.standard_normal(nl)
    return bs, nl, sec, prim, J, F


@pytest.mark.parametrize("nc", [3, 6, 10, 12])
@pytest.mark.parametrize("permute", [False, True])
def test_group_elimination_transcribed(nc, permute):
    rng = np.random.default_rng(nc + 17 * permute)
    for lab_a, lab_b in ((0, 1), (-1, 0), (0, 0), (1, -1)):
        cols = [[0.0] * (nc + 3) for _ in range(64)]
        systems = {}
        for gbase, lab in ((0, lab_a), (GROUP, lab_b)):
            bs, nl, sec, prim, J, F = _system(rng, nc, lab, permute)
            for j in range(bs):
                cols[gbase + j][:nl] = J[:, j].tolist()
            cols[gbase + GROUP - 1][:nl] = F.tolist()
            systems[gbase] = (lab, bs, nl, sec, prim, J, F)
        reads = []
        # the two groups take different static shapes in one wave
        for gbase, (lab, bs, nl, sec, prim, J, F) in systems.items():
            eliminate(cols, nc, lab, nl, gbase, bs, reads)
        for gbase, (lab, bs, nl, sec, prim, J, F) in systems.items():
            srcs = [P for g, P in reads if g == gbase]
            assert all(gbase <= P < gbase + bs and P != gbase + GROUP - 1 for P in srcs)
            assert sorted(P - gbase for P in srcs) == sorted(sec)
            S, t = write_st(cols, nc, lab, nl, gbase, bs)
            A = J[:, sec]
            Sref = np.linalg.solve(A, J[:, prim])
            tref = np.linalg.solve(A, F)
            np.testing.assert_allclose(S[:nl, :len(prim)], Sref, rtol=1e-11, atol=1e-12)
            np.testing.assert_allclose(t[:nl], tref, rtol=1e-11, atol=1e-12)
            assert not S[nl:].any() and not t[nl:].any()
