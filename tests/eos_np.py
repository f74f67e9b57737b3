"""Independent numpy statement of the PR/SRK cubic EOS, used to pin the oracle.

Written from the textbook thermodynamics (residual Gibbs energy of a
two-parameter cubic, e.g. Michelsen & Mollerup, "Thermodynamic Models", ch. 3),
not from the oracle's code: Z comes from numpy's companion-matrix roots and
ln(phi_i) from differentiating n*G_res/RT numerically in the mole numbers.
The constants Omega_a, Omega_b and m(omega) are the reference's
(compsim/Init_Calcs.cpp:80-95) because they define the model being restated.
"""
from __future__ import annotations

import numpy as np

R = 8.3144621


def ab(fluid, T):
    Tc, Pc, w = (np.asarray(v, float) for v in (fluid.Tc, fluid.Pc, fluid.omega))
    if fluid.eos == 1:  # SRK
        Oa, Ob = 0.42747, 0.08664
        m = 0.48508 + 1.55171 * w - 0.15613 * w ** 2
    else:
        Oa, Ob = 0.457235, 0.077796
        m = 0.3796 + 1.485 * w - 0.1644 * w ** 2 + 0.01667 * w ** 3
    a = Oa * R ** 2 * Tc ** 2 / Pc * (1 + m * (1 - np.sqrt(T / Tc))) ** 2
    b = Ob * R * Tc / Pc
    return a, b


def deltas(eos):
    return (1.0, 0.0) if eos == 1 else (1 + np.sqrt(2), 1 - np.sqrt(2))


def real_roots(A, B, eos):
    d1, d2 = deltas(eos)
    # (Z - B)(Z + d1 B)(Z + d2 B) expanded, minus A (Z - B) ... written as
    # Z^3 + c2 Z^2 + c1 Z + c0 for P = RT/(v-b) - a/((v+d1 b)(v+d2 b)).
    c2 = (d1 + d2 - 1) * B - 1
    c1 = A + d1 * d2 * B ** 2 - (d1 + d2) * B * (B + 1)
    c0 = -(A * B + d1 * d2 * B ** 2 * (B + 1))
    r = np.roots([1.0, c2, c1, c0])
    return np.sort(r[np.abs(r.imag) < 1e-9 * max(1.0, np.abs(r).max())].real)


def pick(roots, gas):
    if gas:
        return roots.max()
    pos = roots[roots > 0]
    return pos.min() if pos.size else roots.max()


def gres_n(fluid, T, P, n, gas):
    """n * G_res/(R T) for mole numbers n (PR/SRK), the molar residual Gibbs
    energy at fixed T, P times the total moles."""
    a, b = ab(fluid, T)
    kij = np.asarray(fluid.kij, float)
    aij = np.sqrt(np.outer(a, a)) * (1 - kij)
    nt = n.sum()
    am = n @ aij @ n / nt ** 2
    bm = n @ b / nt
    A = am * P / (R * T) ** 2
    B = bm * P / (R * T)
    Z = pick(real_roots(A, B, fluid.eos), gas)
    d1, d2 = deltas(fluid.eos)
    g = (Z - 1) - np.log(Z - B) - A / (B * (d1 - d2)) * np.log((Z + d1 * B) / (Z + d2 * B))
    return nt * g, Z


def lnphi_fd(fluid, T, P, x, gas, h=1e-6):
    x = np.asarray(x, float)
    out = np.zeros(len(x))
    for i in range(len(x)):
        e = np.zeros(len(x))
        e[i] = h
        gp, _ = gres_n(fluid, T, P, x + e, gas)
        gm, _ = gres_n(fluid, T, P, x - e, gas)
        out[i] = (gp - gm) / (2 * h)
    _, Z = gres_n(fluid, T, P, x, gas)
    return Z, out
