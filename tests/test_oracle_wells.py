"""Pins the oracle's well terms -- Calc_Re, FlashInjection and the rates the
wells add to the balance rows -- with conditions that do not share its code:
Peaceman's closed form on a square block and its image-well analogues,
mirror and rotation symmetry of the boundary cases, phase equilibrium of the
injection stream under an independent EOS, the allocation of a specified
rate over a well's completions (the completions' rates sum to it), the
linearity of a pressure-specified rate in Pwf - P, the rows and blocks a well
touches, and finite differences of the residual for the well rows'
own-block Jacobian columns."""
import copy

import numpy as np
import pytest
import scipy.sparse as sp

from tests import eos_np
from tests.asm_cases import bal_rows, c0_case, cell, column_state, column_variable, frozen, with_wells, wrow

WAT_M_RO = 55.5084


@pytest.fixture(scope="module")
def case(oracle):
    return c0_case(oracle, nx=5, ny=4, nz=3, seed=3)


# ---------------------------------------------------------------- Calc_Re
def _re_at(oracle, fl, grid, st, dx, dy, blocks):
    g = copy.copy(grid)
    g.dx = np.full(grid.nx, dx)
    g.dy = np.full(grid.ny, dy)
    g = with_wells(g, [wrow(fl.nc, 0, ix, iy, 0, 0, 7e6, 0) for ix, iy in blocks])
    re, _ = oracle.well_data(fl, g, st)
    return re


def test_calc_re_closed_forms_on_square_blocks(oracle, case):
    fl, grid, st = case
    nx, ny = grid.nx, grid.ny
    d = 7.5
    interior = [(2, 1), (1, 2), (3, 2)]
    sides = [(2, ny - 1), (2, 0), (nx - 1, 1), (0, 2)]
    corners = [(0, ny - 1), (0, 0), (nx - 1, ny - 1), (nx - 1, 0)]
    re = _re_at(oracle, fl, grid, st, d, d, interior + sides + corners)
    # Peaceman (1978): r_o = e^{-pi/2} dx on a square block
    np.testing.assert_allclose(re[:3], np.exp(-np.pi / 2) * d, rtol=1e-14)
    # one no-flow side: three faces (T = 1) and one image at distance d,
    # whose distances to the faces' neighbours are sqrt(2) d, 2 d, sqrt(2) d:
    # ln r = (3 ln d + ln 4 - 2 pi) / 3
    np.testing.assert_allclose(re[3:7], d * (4 * np.exp(-2 * np.pi)) ** (1 / 3), rtol=1e-14)
    # a corner: two faces and three images (at d, sqrt(2) d, d), image
    # distance ratios sqrt(2), sqrt(5/2), 2 per face: ln r = (2 ln d + ln 20 - 2 pi) / 2
    np.testing.assert_allclose(re[7:], d * np.sqrt(20) * np.exp(-np.pi), rtol=1e-14)


def test_calc_re_mirror_and_rotation_symmetry(oracle, case):
    fl, grid, st = case
    nx, ny = grid.nx, grid.ny
    blocks = [(2, 1), (2, ny - 1), (2, 0), (nx - 1, 1), (0, 1), (0, ny - 1), (0, 0), (nx - 1, ny - 1),
              (nx - 1, 0)]
    a, b = 10.0, 25.0
    r = _re_at(oracle, fl, grid, st, a, b, blocks)
    s = _re_at(oracle, fl, grid, st, b, a, blocks)
    # mirror images of the reservoir give the same radius
    np.testing.assert_allclose(r[1], r[2], rtol=1e-14)       # Iy = ny-1 / Iy = 0
    np.testing.assert_allclose(r[3], r[4], rtol=1e-14)       # Ix = nx-1 / Ix = 0
    np.testing.assert_allclose(r[5:], r[5], rtol=1e-14)      # the four corners
    # a quarter turn swaps dx and dy and the x and y sides
    np.testing.assert_allclose(s[0], r[0], rtol=1e-14)
    np.testing.assert_allclose(s[1], r[3], rtol=1e-14)
    np.testing.assert_allclose(s[3], r[1], rtol=1e-14)
    np.testing.assert_allclose(s[5:], r[5:], rtol=1e-14)
    # the no-flow boundary brings the images in: a side block's radius lies
    # between the interior's and a corner's ... and the interior's is the
    # rectangular-block form of the same product
    Tx, Ty = a / b, b / a
    ln_int = (2 * Tx * np.log(b) + 2 * Ty * np.log(a) - 2 * np.pi) / (2 * Tx + 2 * Ty)
    np.testing.assert_allclose(r[0], np.exp(ln_int), rtol=1e-14)


# ---------------------------------------------------------- FlashInjection
def _inj(oracle, fl, grid, st, z, P, t=3):
    nc = fl.nc
    a, b = (P, 0.0) if t == 3 else (10.0, P)
    g = with_wells(grid, [wrow(nc, t, 1, 1, 0, 0, a, b, z=np.asarray(z, float))])
    _, inj = oracle.well_data(fl, g, st)
    return inj[0]


def test_flash_injection_two_phase_equilibrium(oracle, case):
    fl, grid, st = case
    nc = fl.nc
    T = grid.T
    for z, P, t in (([0.6, 0.1, 0.3], 10e6, 3), ([0.3, 0.2, 0.5], 5e6, 8)):
        z = np.asarray(z)
        o = _inj(oracle, fl, grid, st, z, P, t)
        x, y = o[:nc], o[nc:2 * nc]
        OF, GF, OD, GD = o[2 * nc:2 * nc + 4]
        L = OF
        assert 0 < L < 1 and GF == pytest.approx(1 - L, abs=1e-15)
        # the Rachford-Rice Newton stops on a step below 1e-4 (NEWTON_TOL
        # 1e-8 on its square), so the phase split closes to ~1e-8
        assert x.sum() == pytest.approx(1, abs=1e-7) and y.sum() == pytest.approx(1, abs=1e-7)
        np.testing.assert_allclose(L * x + (1 - L) * y, z, atol=1e-7)
        Zl, lpl = eos_np.lnphi_fd(fl, T, P, x, gas=False)
        Zg, lpg = eos_np.lnphi_fd(fl, T, P, y, gas=True)
        np.testing.assert_allclose(np.log(x) + lpl, np.log(y) + lpg, atol=1e-6)
        np.testing.assert_allclose([OD, GD], [P / (Zl * eos_np.R * T), P / (Zg * eos_np.R * T)], rtol=1e-9)


def test_flash_injection_single_phase_and_viscosities(oracle, case):
    fl, grid, st = case
    nc = fl.nc
    # a lean gas stays one vapour phase: all gas, y = z
    o = _inj(oracle, fl, grid, st, [0.9, 0.1, 0.0], 10e6)
    assert o[2 * nc] == 0 and o[2 * nc + 1] == 1
    np.testing.assert_allclose(o[nc:2 * nc], [0.9, 0.1, 0.0], atol=1e-15)
    # the stream's viscosities are the LBC viscosities of its phases: those
    # the property sweep gives a fracture block holding the phase at P
    P = 10e6
    o = _inj(oracle, fl, grid, st, [0.6, 0.1, 0.3], P)
    x, y = o[:nc], o[nc:2 * nc]
    s = copy.deepcopy(st)
    s.label_f[0], s.label_f[1] = 1, -1
    s.Po_f[0] = s.Po_f[1] = P
    s.x_f[:, 0], s.y_f[:, 0] = x, x
    s.x_f[:, 1], s.y_f[:, 1] = y, y
    pr = oracle.props(fl, grid, s)
    np.testing.assert_allclose(o[2 * nc + 2], pr[1, 0, 5], rtol=1e-12)    # oil density
    np.testing.assert_allclose(o[2 * nc + 3], pr[1, 1, 6], rtol=1e-12)    # gas density
    np.testing.assert_allclose(o[2 * nc + 4], pr[1, 0, 7], rtol=1e-12)    # oil viscosity
    np.testing.assert_allclose(o[2 * nc + 5], pr[1, 1, 8], rtol=1e-12)    # gas viscosity


# ----------------------------------------------------------- well rows
def _dF(oracle, fl, grid, st, rows, **kw):
    F0 = oracle.residual(fl, grid, st)
    F1 = oracle.residual(fl, with_wells(grid, rows, **kw), st)
    return F1 - F0


def test_pwf_well_rows_blocks_and_linearity(oracle, case):
    fl, grid, st = case
    nc, n = fl.nc, grid.n
    ix, iy = 2, 1
    s = column_state(st, grid, ix, iy, Po=12e6)
    bs, off = oracle.layout(fl, grid, s)
    own = set()
    for k in range(2):
        for iz in (0, 1):
            own.update(bal_rows(fl, s, off, k * n + cell(grid, ix, iy, iz), n).tolist())
    mask = np.zeros(off[-1], bool)
    mask[sorted(own)] = True
    hc = np.zeros(off[-1], bool)     # their hydrocarbon rows (the water is immobile here)
    for k in range(2):
        for iz in (0, 1):
            hc[bal_rows(fl, s, off, k * n + cell(grid, ix, iy, iz), n)[:nc]] = True
    d = {}
    for pwf in (6e6, 8e6, 10e6):
        d[pwf] = _dF(oracle, fl, grid, s, [wrow(nc, 0, ix, iy, 0, 1, pwf, 0)])
        # only the completed blocks' balance rows change, and all their
        # hydrocarbon rows do
        assert np.all(d[pwf][~mask] == 0)
        assert np.all(d[pwf][hc] != 0)
    # the rates are PI (Pwf - P): affine in Pwf
    np.testing.assert_allclose(d[8e6], (d[6e6] + d[10e6]) / 2, rtol=1e-9, atol=1e-12 * np.abs(d[6e6]).max())
    # a producer whose Pwf is above the block pressure produces nothing
    assert np.all(_dF(oracle, fl, grid, s, [wrow(nc, 0, ix, iy, 0, 1, 13e6, 0)]) == 0)
    # below WELL_MIN_P = 7 MPa a type-0 producer shuts in
    s2 = column_state(st, grid, ix, iy, Po=6.9e6)
    assert np.all(_dF(oracle, fl, grid, s2, [wrow(nc, 0, ix, iy, 0, 1, 5e6, 0)]) == 0)
    # outside its time window a well is off; at the window's ends it is on
    assert np.all(_dF(oracle, fl, grid, s, [wrow(nc, 0, ix, iy, 0, 1, 8e6, 0, t0=7200, t1=1e9)],
                      sim_time=3600) == 0)
    assert np.any(_dF(oracle, fl, grid, s, [wrow(nc, 0, ix, iy, 0, 1, 8e6, 0, t0=3600, t1=1e9)],
                      sim_time=3600) != 0)


def _sum_rows(d, fl, st, off, n, grid, ix, iy, layers):
    tot = np.zeros(fl.nc + 1)
    for k in range(2):
        for iz in layers:
            tot += d[bal_rows(fl, st, off, k * n + cell(grid, ix, iy, iz), n)]
    return tot


def test_rate_wells_allocate_the_specified_rate(oracle, case):
    """The fracture rows weigh the matrix completions with Re_f in place of
    the matrix's Calc_Re radius (Petsc_Functions.cpp:185 sets Re = Re_f for
    the whole fracture block), so the completions' rates add up to the
    specified rate exactly when the two radii agree: Re_f is set to Calc_Re
    of the well's column here."""
    fl, grid, st = case
    nc, n = fl.nc, grid.n
    ix, iy = 3, 2
    re_m, _ = oracle.well_data(fl, with_wells(grid, [wrow(nc, 0, ix, iy, 0, 0, 7e6, 0)]), st)
    kw = dict(re_f=float(re_m[0]))
    # gas injection at a specified rate (type 8): the completions share the
    # rate by their geometric factors, so the gas into the column is the
    # rate, the oil OF/GF of it, each component the stream's share
    z = np.array([0.6, 0.1, 0.3])
    s = column_state(st, grid, ix, iy, Po=10e6)
    bs, off = oracle.layout(fl, grid, s)
    rate = 40.0
    rows = [wrow(nc, 8, ix, iy, 0, 2, rate, 13e6, z=z)]
    g = with_wells(grid, rows, **kw)
    _, inj = oracle.well_data(fl, g, s)
    xo, yg, OF, GF = inj[0, :nc], inj[0, nc:2 * nc], inj[0, 2 * nc], inj[0, 2 * nc + 1]
    tot = _sum_rows(_dF(oracle, fl, grid, s, rows, **kw), fl, s, off, n, grid, ix, iy, range(3))
    qo = rate * OF / GF
    np.testing.assert_allclose(tot[nc - 1], -(rate + qo), rtol=1e-12)
    np.testing.assert_allclose(tot[:nc - 1], -(xo[:nc - 1] * qo + yg[:nc - 1] * rate), rtol=1e-12)
    assert tot[nc] == 0
    # above WELL_MAX_P = 14 MPa the injector shuts in
    s_hi = column_state(st, grid, ix, iy, Po=14.5e6)
    assert np.all(_dF(oracle, fl, grid, s_hi, rows, **kw) == 0)
    # water injection at a specified rate (type 2 under WellCondition): split
    # by the completions' water mobility, summing to the rate in any state
    rows = [wrow(nc, 2, ix, iy, 0, 1, 20e6, 7.0)]
    sw = column_state(st, grid, ix, iy, Po=10e6)
    # water mobile in every completion
    for iz in range(grid.nz):
        c = cell(grid, ix, iy, iz)
        sw.S_m[:, c] = [0.45, 0.35, 0.2] if sw.label_m[c] == 0 else [0.45, 0.55, 0.0]
        sw.S_f[:, c] = [0.3, 0.5, 0.2] if sw.label_f[c] == 0 else [0.3, 0.7, 0.0]
    bs, off = oracle.layout(fl, grid, sw)
    d = _dF(oracle, fl, grid, sw, rows, cond=[1], **kw)
    tot = _sum_rows(d, fl, sw, off, n, grid, ix, iy, (0, 1))
    # the specified mass rate enters the water row in moles (WAT_M_RO mol/kg)
    np.testing.assert_allclose(tot[nc], -7.0 * WAT_M_RO, rtol=1e-12)
    assert np.all(tot[:nc] == 0)
    # oil production at a specified rate (type 4) from an undersaturated oil
    # column (one state in every completion of both continua): the oil out
    # of the column is the rate
    so = column_state(st, grid, ix, iy, Po=12e6, label=1, x=[0.3, 0.3, 0.4], y=[0.3, 0.3, 0.4],
                 S=[0.25, 0.75, 0.0])
    bs, off = oracle.layout(fl, grid, so)
    rows = [wrow(nc, 4, ix, iy, 0, 2, 3.0, 0.0)]
    tot = _sum_rows(_dF(oracle, fl, grid, so, rows, **kw), fl, so, off, n, grid, ix, iy, range(3))
    np.testing.assert_allclose(tot[nc - 1], 3.0, rtol=1e-12)
    np.testing.assert_allclose(tot[:nc - 1], 3.0 * np.array([0.3, 0.3]), rtol=1e-12)


def _all_types(nc, ix, iy):
    z = np.array([0.6, 0.1, 0.3])
    return [wrow(nc, 0, ix, iy, 0, 1, 8e6, 5.0),
            wrow(nc, 1, ix, iy, 0, 1, 8e6, 5.0),
            wrow(nc, 2, ix, iy, 0, 1, 20e6, 5.0),
            wrow(nc, 3, ix, iy, 0, 1, 20e6, 5.0, z=z),
            wrow(nc, 4, ix, iy, 0, 1, 3.0, 0.0),
            wrow(nc, 5, ix, iy, 0, 1, 3.0, 0.0),
            wrow(nc, 6, ix, iy, 0, 1, 3.0, 0.0),
            wrow(nc, 7, ix, iy, 0, 1, 20e6, 5.0),
            wrow(nc, 8, ix, iy, 0, 1, 40.0, 13e6, z=z)]


@pytest.mark.parametrize("wtype", range(9))
@pytest.mark.parametrize("cond", (0, 1))
def test_well_rows_own_jacobian_columns_match_finite_differences(oracle, wtype, cond):
    fl, grid, st = c0_case(oracle, nx=4, ny=3, nz=2, seed=1, sw_m=0.45, sw_f=0.3)
    nc, n = fl.nc, grid.n
    ix, iy = 1, 1
    # the continua at different pressures: the transfer term's upstream
    # choice has a kink at equal pressures
    # (a two-phase column with mobile water, oil and gas off the saturation
    # tables' nodes: a gas-rate well in a column without mobile gas divides
    # 0 by 0, as the reference does)
    s0 = frozen(oracle, fl, grid, column_state(st, grid, ix, iy, Po=(10e6, 10.2e6), label=0, x=[0.3, 0.3, 0.4],
                                          y=[0.8, 0.15, 0.05], S=([0.3, 0.6, 0.1], [0.3, 0.5, 0.2])))
    g = with_wells(grid, [_all_types(nc, ix, iy)[wtype]], cond=[cond])
    rp, cols, vals = oracle.jacobian(fl, g, s0)
    bs, off = oracle.layout(fl, g, s0)
    N = off[-1]
    J = sp.csr_matrix((vals, cols, rp), shape=(N, N))
    F0 = oracle.residual(fl, grid, s0)
    active = np.any(oracle.residual(fl, g, s0) != F0)
    # a type-7 well injects only under WellCondition
    assert active == (wtype != 7 or cond == 1), "the well is active"
    for k in range(2):
        blk = k * n + cell(grid, ix, iy, 0)
        rows = bal_rows(fl, s0, off, blk, n)
        for col in range(off[blk], off[blk + 1]):
            name, idx = column_variable(s0, nc, n, bs, off, col)
            arr = getattr(s0, name)
            base = arr[idx]
            h = 1e-6 * base if name.startswith("Po") else 1e-6 * max(abs(base), 1e-3)

            def F_at(v):
                a = np.array(arr, copy=True)
                a[idx] = v
                return oracle.residual(fl, g, _replace(s0, **{name: a}))

            fd = (F_at(base + h) - F_at(base - h)) / (2 * h)
            if name.startswith("Po"):
                fd *= s0.pcoeff
            jc = J[:, col].toarray().ravel()
            scale = max(np.abs(fd[rows]).max(), np.abs(jc[rows]).max(), 1e-300)
            assert np.abs(fd[rows] - jc[rows]).max() <= 1e-5 * scale, (wtype, cond, col, name)


def _replace(st, **kw):
    s = copy.copy(st)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


# ------------------------------------------- reference-mode Jacobian
def _dense(rpc):
    rp, cols, vals = rpc
    N = len(rp) - 1
    return sp.csr_matrix((vals, cols, rp), shape=(N, N))


@pytest.mark.parametrize("wtype", (0, 1, 2, 3, 4, 7, 8))
def test_reference_jacobian_well_columns(oracle, wtype):
    """The reference's well entries (IteratorJacobian, matrix :32543,
    :32808-32813, :35425, :36029; fracture :15033, :17868, :18474): the
    productivity indices PI = Q / (Pwf - P), read off the residual's rates,
    in the block's P column (times PCoeff; component row i: x_i PIo + y_i
    PIg with the injection stream's y_i in an injector's block, total PIo +
    PIg, water PIw) and, for a producer, -Qo, -Qg in the block's x_i, y_i
    columns; nothing elsewhere."""
    fl, grid, st = c0_case(oracle, nx=4, ny=3, nz=2, seed=1, sw_m=0.45, sw_f=0.3)
    nc, n = fl.nc, grid.n
    ix, iy = 1, 1
    s0 = column_state(st, grid, ix, iy, Po=(10e6, 10.2e6), label=0, x=[0.3, 0.3, 0.4],
                      y=[0.8, 0.15, 0.05], S=([0.3, 0.6, 0.1], [0.3, 0.5, 0.2]))
    z = np.array([0.6, 0.1, 0.3])
    rows = {0: wrow(nc, 0, ix, iy, 0, 1, 8e6, 0.0), 1: wrow(nc, 1, ix, iy, 0, 1, 8e6, 0.0),
            2: wrow(nc, 2, ix, iy, 0, 1, 20e6, 0.0), 3: wrow(nc, 3, ix, iy, 0, 1, 13e6, 0.0, z=z),
            4: wrow(nc, 4, ix, iy, 0, 1, 3.0, 0.0), 7: wrow(nc, 7, ix, iy, 0, 1, 20e6, 0.0),
            8: wrow(nc, 8, ix, iy, 0, 1, 40.0, 13e6, z=z)}
    cond = [1] if wtype in (7, 8) else [0]
    g = with_wells(grid, [rows[wtype]], cond=cond)
    bs, off = oracle.layout(fl, g, s0)
    dF = oracle.residual(fl, g, s0) - oracle.residual(fl, grid, s0)
    dJ = (_dense(oracle.jacobian_ref(fl, g, s0)) - _dense(oracle.jacobian_ref(fl, grid, s0))).tocoo()
    pr = oracle.props(fl, g, s0)
    _, inj = oracle.well_data(fl, g, s0)
    pwf = rows[wtype][1][0] if wtype != 8 else rows[wtype][1][1]
    expect = {}
    for k in range(2):
        for iz in (0, 1):
            c = cell(grid, ix, iy, iz)
            blk = k * n + c
            r = bal_rows(fl, s0, off, blk, n)
            x = (s0.x_m if k == 0 else s0.x_f)[:, c]
            y = (s0.y_m if k == 0 else s0.y_f)[:, c]
            injector = wtype in (3, 8)
            xo, yg = (inj[0, :nc], inj[0, nc:2 * nc]) if injector else (x, y)
            # rates from the residual: [comp rows..., total] = -A [Qo, Qg]
            A = np.vstack([np.column_stack([xo[:nc - 1], yg[:nc - 1]]), [1.0, 1.0]])
            q, *_ = np.linalg.lstsq(A, -dF[r[:nc]], rcond=None)
            Qo, Qg, Qw = q[0], q[1], -dF[r[nc]]
            Pw, Po, Pg = pr[k, c, 1], pr[k, c, 2], pr[k, c, 3]
            pressure_rates = wtype in (0, 1, 2, 3, 7) or (wtype == 8 and cond[0])
            if pressure_rates:
                PIo = Qo / (pwf - (Pg if injector else Po))
                PIg = Qg / (pwf - Pg)
                PIw = Qw / (pwf - Pw)
            else:
                PIo = PIg = PIw = 0.0
            pcol = off[blk] + 2 * nc            # two-phase block: [x, y, P, ...]
            for i in range(nc - 1):
                expect[(r[i], pcol)] = s0.pcoeff * (x[i] * PIo + yg[i] * PIg)
                if not injector:
                    expect[(r[i], off[blk] + i)] = -Qo
                    expect[(r[i], off[blk] + nc + i)] = -Qg
            expect[(r[nc - 1], pcol)] = s0.pcoeff * (PIo + PIg)
            expect[(r[nc], pcol)] = s0.pcoeff * PIw
    got = {}
    for a, b, v in zip(dJ.row, dJ.col, dJ.data):
        if v != 0:
            got[(a, b)] = got.get((a, b), 0.0) + v
    scale = max(abs(v) for v in expect.values())
    assert sum(abs(v) > 1e-9 * scale for v in expect.values()) >= 4, "the well sets entries"
    for key, v in expect.items():
        assert got.get(key, 0.0) == pytest.approx(v, rel=1e-7, abs=1e-12 * scale), (wtype, key)
    extra = {kk: v for kk, v in got.items() if kk not in expect and abs(v) > 1e-12 * scale}
    assert not extra, (wtype, extra)
