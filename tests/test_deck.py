"""The deck subset (osiris_amd/deck.py) read as the reference's Data_Input
reads it (compsim/Data_Input.cpp, MIfstream.cpp).

CPU: the reference's example deck (tests/golden/test.dat, a copy of
compsim/test.dat) yields the values written in it, with the reference's unit
conversions (RESTEMP + 273.15, DT and TOTALT in hours) and defaults (SIGMA 1,
TOR 1, the KSP tolerances), its saturation tables equal the independent
transcription the synthetic workloads use, and its components the library
entries; token, atof/atoi and n*v repeat semantics; every distribution form;
the forms the reference cannot read (JVAR, INITCOND 2, components without a
parachor) are refused.  The deck's initial state (the oracle's flash) gives a
finite residual and Jacobian in the oracle.

GPU: the same deck through the device path -- initial flash, residual,
reference-layout Jacobian and the reduced system -- against the oracle at
the bars of the assembly parity tests (labels equal; x, y 1e-8; residual 1e-8
of its row scale; Jacobian entries 1e-8 of their row's largest entry); the
reduced solve has normwise backward error below 1e-12.
"""
import os

import numpy as np
import pytest

from osiris_amd import deck, synth

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "test.dat")


def test_example_deck_values():
    d = deck.load(GOLDEN)
    g, f = d.grid, d.fluid
    assert (g.nx, g.ny, g.nz) == (10, 10, 5)
    assert np.all(g.dx == 100) and np.all(g.dy == 100) and np.all(g.dz == 50)
    assert np.all(g.phi_m == 0.1932) and np.all(g.phi_f == 0.02)
    assert np.all(g.perm_m == 10) and np.all(g.perm_f == 1000) and g.perm_m.shape == (3, 500)
    assert np.all(g.tor == 1.0)
    assert g.lf == 0.01 and d.run["re_f"] == 1e6
    assert g.T == 80 + 273.15 and g.wat_ro == 1000 and g.wat_mu == 1
    assert g.dt == 0.01 * 3600 and d.run["total_time"] == 17520 * 3600 and d.run["dt_max"] == 100
    assert g.refP == 101325 and g.cpor == 0 and g.dcpor == 0
    assert g.sigma == 1.0                      # no SIGMA keyword: SigmaMF = 1
    assert g.const_diff and g.cdiff == 0.0     # CDIF 0: constant (zero) diffusion
    assert np.array_equal(g.swt, synth.SWT_TEST_DAT) and np.array_equal(g.sgt, synth.SGT_TEST_DAT)
    assert d.run["ksp"] == {"rtol": 1e-18, "atol": 1e-50, "dtol": 1e3, "maxit": 5000}
    assert d.run["solver_type"] == "bcgs"
    assert d.names == ["C1", "C2", "nC5", "nC10", "nC15"] and f.nc == 5 and f.eos == synth.PR
    assert np.array_equal(f.Tc, [190.56, 305.32, 469.7, 617.7, 708])
    assert np.array_equal(f.Pc, [4599000, 4872000, 3370000, 2110000, 1480000])
    assert np.array_equal(f.omega, [0.008, 0.0995, 0.2515, 0.4923, 0.6863])
    assert np.array_equal(f.MW, [16.043, 30.070, 72.150, 142.285, 212.419])
    assert np.array_equal(f.parachor, [74.05, 112.91, 236.00, 440.69, 647.43])
    assert not f.kij.any()                     # no BIC keyword
    ini = d.init
    assert ini["kind"] == 3
    assert np.all(ini["Po_m"] == 12e6) and np.all(ini["Po_f"] == 12e6)
    assert np.all(ini["Sw_m"] == 0.2) and np.all(ini["Sw_f"] == 0.0)
    for k in ("z_m", "z_f"):
        assert ini[k].shape == (5, 500) and np.all(ini[k] == np.array([0, 0.1, 0.1, 0.3, 0.5])[:, None])
    assert d.wells == [
        {"type": 8, "i": 0, "j": 0, "k_first": 0, "k_last": 4,
         "values": [50, 14e6, 0, 0.1, 1, 0, 0, 0, 0, 1 * 3600, 17520 * 3600]},
        {"type": 0, "i": 9, "j": 9, "k_first": 0, "k_last": 0, "values": [7e6, 10, 0, 0.1, 1 * 3600, 17520 * 3600]}]


def test_token_and_number_semantics():
    assert deck.tokens("GRID\t3  2\r\n1\x01x\x7fy é") == ["GRID", "3", "2", "1", "x", "y"]
    assert deck.atof("12abc") == 12.0 and deck.atof("abc") == 0.0 and deck.atof("1e3x") == 1000.0
    assert deck.atof("-.5") == -0.5 and deck.atof("+2.") == 2.0
    assert deck.atoi("7.5") == 7 and deck.atoi("x") == 0 and deck.atoi("-3") == -3
    assert deck.ecl_star("3*0.5") == (3, 0.5) and deck.ecl_star("4") == (1, 4.0)
    assert deck.ecl_star("2**7") == (2, 7.0)


BASE = """GRID 3 2 2
DI VAR 1 2 3
DJ CON 4
DK VAR 5 6
POR VAR 0.1 0.2 0 0.3 0.4 0.5 0.6 0.7 0.8 0.9 0.15 0.25
FPOR KVAR 0.01 0.02
TOR IVAR 1 2 3
PERMI CON 1  PERMJ CON 2  PERMK KVAR 3 4
FPERMI CON 10 FPERMJ CON 20 FPERMK CON 30
Lf 0.005 Re_f 10
MODEL SRK
NC 3 0
COMPNAME C1 CO2 nC10
BIC 0.1 0.2 0.3
SWT 2
0.2 0 1 100
1 1 0 0
SGT 2
0 0 1 0
0.8 1 0 0
WELLS 0
RESTEMP 100
WATERPROPS 990 0.5
SIGMA 0.25
DT 2 TOTALT 48
"""


def _with_init(init):
    return BASE + init


def test_distributions_and_initial_conditions():
    d = deck.parse(_with_init("INITCOND 0\nIPRESS 5*1e7 7*2e7\nFIPRESS 12*3e7\nIWS 0*0.3\nFIWS 12*0\n"
                              "IGC 12*0.2 12*0.3 12*0.5 9\nFIGC 36*0.3333\n"))
    g, f = d.grid, d.fluid
    assert np.array_equal(g.dx, [1, 2, 3]) and np.array_equal(g.dy, [4, 4]) and np.array_equal(g.dz, [5, 6])
    # VAR in k-outer, i-inner order; a zero becomes 1e-11
    assert np.array_equal(g.phi_m, [0.1, 0.2, 1e-11, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 0.15, 0.25])
    assert np.array_equal(g.phi_f, [0.01] * 6 + [0.02] * 6)            # KVAR
    assert np.array_equal(g.tor, [1, 2, 3] * 4)                        # IVAR
    assert np.array_equal(g.perm_m[2], [3] * 6 + [4] * 6)
    assert np.array_equal(g.perm_f[:, 0], [10, 20, 30])
    assert f.eos == synth.SRK and g.sigma == 0.25 and g.T == 373.15 and g.dt == 7200 and not g.const_diff
    k = np.zeros((3, 3))
    k[1, 0] = k[0, 1] = 0.1
    k[2, 0] = k[0, 2] = 0.2
    k[2, 1] = k[1, 2] = 0.3
    assert np.array_equal(f.kij, k)
    ini = d.init
    assert np.array_equal(ini["Po_m"], [1e7] * 5 + [2e7] * 7)
    assert np.all(ini["Sw_m"] == 0.3)                                  # 0*v fills the rest
    assert np.array_equal(ini["z_m"], np.array([[0.2] * 12, [0.3] * 12, [0.5] * 12]))   # the 9 is not read
    d1 = deck.parse(_with_init("INITCOND 1\nIPRESS 1e7 2e7\nFIPRESS 3e7 4e7\nIWS 0.1 0.2\nFIWS 0 0\n"
                               "IGC 0.5 0.3 0.2 0.6 0.3 0.1\nFIGC 0.4 0.4 0.2 0.1 0.1 0.8\n"))
    ini = d1.init
    assert np.array_equal(ini["Po_f"], [3e7] * 6 + [4e7] * 6)
    assert np.array_equal(ini["z_m"][:, 0], [0.5, 0.3, 0.2]) and np.array_equal(ini["z_m"][:, 11], [0.6, 0.3, 0.1])


def test_first_occurrence_and_refusals():
    # FileSearch finds the first token equal to the keyword
    d = deck.parse(_with_init("SIGMA 9\nINITCOND 3 IPRESS 1e7 FIPRESS 1e7 IWS 0 FIWS 0 IGC 0.5 0.3 0.2 "
                              "FIGC 0.5 0.3 0.2\n"))
    assert d.grid.sigma == 0.25
    ok = "INITCOND 3 IPRESS 1e7 FIPRESS 1e7 IWS 0 FIWS 0 IGC 0.5 0.3 0.2 FIGC 0.5 0.3 0.2\n"
    for bad, what in ((BASE.replace("POR VAR 0.1 0.2 0 0.3 0.4 0.5 0.6 0.7 0.8 0.9 0.15 0.25",
                                    "POR JVAR 0.1 0.2") + ok, "JVAR"),
                      (_with_init("INITCOND 2\n"), "INITCOND 2"),
                      (BASE.replace("COMPNAME C1 CO2 nC10", "COMPNAME C1 O2 nC10") + ok, "parachor"),
                      (BASE.replace("NC 3 0", "NC 3 1") + ok, "user components"),
                      (BASE.replace("COMPNAME C1 CO2 nC10", "COMPNAME C1 XX nC10") + ok, "unknown component"),
                      (BASE.replace("Lf 0.005 ", "") + ok, "Lf"),
                      (_with_init("INITCOND 3 IPRESS"), "IPRESS")):
        with pytest.raises(deck.DeckError, match=what):
            deck.parse(bad)


def test_example_deck_drives_the_oracle(oracle):
    d = deck.load(GOLDEN)
    st = deck.initial_state(d, lambda T, P, z: oracle.flash_batch(d.fluid, T, P, z))
    assert np.all(st.label_m == 1) and np.all(st.label_f == 1)       # 12 MPa, no methane: liquid
    assert st.pcoeff == 12e6
    np.testing.assert_allclose(st.S_m[:, 0], [0.2, 0.8, 0.0])
    F = oracle.residual(d.fluid, d.grid, st)
    rp, cols, vals = oracle.jacobian(d.fluid, d.grid, st)
    assert np.isfinite(F).all() and np.isfinite(vals).all()
    assert len(F) == 2 * 500 * (5 + 3)


@pytest.mark.gpu
def test_example_deck_on_the_device(gpu, oracle):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    import torch
    d = deck.load(GOLDEN)
    ctx = gpu.Context(d.fluid)

    def dev_flash(T, P, z):
        out = gpu.flash(ctx, T, P, z, mode="persistent")
        return {k: v.cpu().numpy() for k, v in out.items()}
    st = deck.initial_state(d, dev_flash)
    ref = deck.initial_state(d, lambda T, P, z: oracle.flash_batch(d.fluid, T, P, z))
    for k in ("label_m", "label_f"):
        assert np.array_equal(getattr(st, k), getattr(ref, k)), k
    for k in ("x_m", "y_m", "x_f", "y_f", "S_m", "S_f"):
        a, b = getattr(st, k), getattr(ref, k)
        assert np.all(np.abs(a - b) <= 1e-8 * np.maximum(np.abs(b), 1e-12)), k
    asm = gpu.Assembler(ctx, d.fluid, d.grid)
    dv = asm.device_state(ref)
    _, nrows = asm.layout(dv)
    F = asm.residual(dv).cpu().numpy()
    Fr = oracle.residual(d.fluid, d.grid, ref)
    Fs = oracle.residual_scale(d.fluid, d.grid, ref)
    err = np.abs(F - Fr) / np.maximum(Fs, 1e-14 * Fs.max() + 1e-300)
    assert err.max() <= 1e-8, err.max()
    rp, cols, vals = asm.jacobian(dv)
    rpo, colso, valso = oracle.jacobian(d.fluid, d.grid, ref)
    assert np.array_equal(rp.cpu().numpy(), rpo) and np.array_equal(cols.cpu().numpy(), colso)
    row = np.repeat(np.arange(len(rpo) - 1), np.diff(rpo))
    rowmax = np.zeros(len(rpo) - 1)
    np.maximum.at(rowmax, row, np.abs(valso))
    v = vals.cpu().numpy()
    assert np.max(np.abs(v - valso) / np.maximum(np.abs(valso), 1e-8 * rowmax[row] + 1e-300)) <= 1e-8
    Ft = torch.as_tensor(F).cuda()
    browptr, bcols, bvals, brhs, S, t = asm.reduced_direct(dv, Ft)
    np1, nb = d.fluid.nc + 1, 2 * d.grid.n
    R = sp.bsr_matrix((bvals.cpu().numpy(), bcols.cpu().numpy().astype(np.int64), browptr.cpu().numpy()),
                      shape=(nb * np1, nb * np1)).tocsc()
    dp = spla.spsolve(R, brhs.cpu().numpy().reshape(-1))
    du = asm.reduced_expand(dv, S, t, torch.as_tensor(dp.reshape(nb, np1)).cuda()).cpu().numpy()
    J = sp.csr_matrix((valso, colso, rpo), shape=(nrows, nrows))
    eta = np.abs(J @ du + Fr).max() / (abs(J).sum(axis=1).max() * np.abs(du).max() + np.abs(Fr).max())
    assert eta < 1e-12, eta
