"""Exact-equilibrium two-phase split for derivative checks (test helper).

Solves ln K_i - ln phiL_i(x) + ln phiG_i(y) = 0 for ln K with scipy's
fsolve, the Rachford-Rice root by brentq in (0, 1), and ln phi from the
oracle's phase routine -- no code shared with the tangent/adjoint kernels,
whose analytic Jacobians this is used to check by central differences.
"""
import numpy as np
from scipy.optimize import brentq, fsolve


def split(oracle, fl, T, P, z, u):
    K = np.exp(u)

    def f(L):
        return float(np.sum(z * (1 - K) / (L + (1 - L) * K)))

    L = brentq(f, 1e-15, 1 - 1e-15, xtol=1e-17, rtol=8.9e-16, maxiter=500)
    x = z / (L + K * (1 - L))
    return L, x, K * x


def solve(oracle, fl, T, P, z, u0):
    def g(u):
        L, x, y = split(oracle, fl, T, P, z, u)
        _, lpl = oracle.phase(fl, T, P, x, False)
        _, lpg = oracle.phase(fl, T, P, y, True)
        return u - lpl + lpg

    u, info, ier, msg = fsolve(g, u0, xtol=1e-14, full_output=True)
    # "xtol too small" is fsolve reaching machine precision: judge by g itself
    assert np.abs(g(u)).max() < 1e-11, (ier, msg, np.abs(g(u)).max())
    L, x, y = split(oracle, fl, T, P, z, u)
    return u, L, x, y
