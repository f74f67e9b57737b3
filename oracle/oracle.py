"""ctypes front end of the CPU FP64 oracle (oracle/libosr_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the cpu_baseline leg of ``bench.py``.  The product path never imports
this module.  Arrays are numpy, structure-of-arrays ``[component, cell]``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libosr_oracle.so")
_lib = None

_dp = C.POINTER(C.c_double)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.or_solve_z.restype = C.c_double
        L.or_solve_z.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int]
        L.or_bubble.restype = C.c_double
        L.or_dew.restype = C.c_double
        L.or_lu_solve.restype = C.c_int
        _lib = L
    return _lib


class _Fluid(C.Structure):
    _fields_ = [("nc", C.c_int), ("eos", C.c_int), ("Tc", _dp), ("Pc", _dp),
                ("w", _dp), ("bic", _dp)]


class _FlashOut(C.Structure):
    _fields_ = [("L", C.c_double), ("Zl", C.c_double), ("Zg", C.c_double),
                ("so", C.c_double), ("sg", C.c_double), ("rho_l", C.c_double),
                ("rho_g", C.c_double), ("label", C.c_int), ("iters", C.c_int)]


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


class OracleFluid:
    """Holds the fluid arrays alive while the C struct points into them."""

    def __init__(self, fluid):
        self.nc = int(fluid.nc)
        self._keep = [_c(fluid.Tc), _c(fluid.Pc), _c(fluid.omega),
                      _c(1.0 - np.asarray(fluid.kij, dtype=np.float64)).reshape(-1)]
        self.s = _Fluid(self.nc, int(fluid.eos), *[_p(a) for a in self._keep])

    @property
    def ref(self):
        return C.byref(self.s)


def eos_ab(fluid, T):
    f = OracleFluid(fluid)
    a = np.zeros(f.nc)
    b = np.zeros(f.nc)
    lib().or_eos_ab(f.ref, C.c_double(T), _p(a), _p(b))
    return a, b


def solve_z(a1, a2, a3, gas):
    return lib().or_solve_z(a1, a2, a3, int(gas))


def phase(fluid, T, P, x, gas):
    f = OracleFluid(fluid)
    a, b = eos_ab(fluid, T)
    x = _c(x)
    Z = C.c_double()
    lnphi = np.zeros(f.nc)
    lib().or_phase(f.ref, _p(a), _p(b), C.c_double(T), C.c_double(P), _p(x),
                   int(gas), C.byref(Z), _p(lnphi))
    return Z.value, lnphi


def flash_cell(fluid, T, P, z, sw=0.0):
    f = OracleFluid(fluid)
    z = _c(z)
    x = np.zeros(f.nc)
    y = np.zeros(f.nc)
    K = np.zeros(f.nc)
    o = _FlashOut()
    lib().or_flash(f.ref, C.c_double(T), C.c_double(P), _p(z), C.c_double(sw),
                   _p(x), _p(y), _p(K), C.byref(o))
    return dict(x=x, y=y, K=K, L=o.L, Zl=o.Zl, Zg=o.Zg, so=o.so, sg=o.sg,
                rho_l=o.rho_l, rho_g=o.rho_g, label=o.label, iters=o.iters)


def flash_batch(fluid, T, P, z, sw=0.0, nthreads=0):
    """z: [nc, n]; T: scalar or [n]; P: [n].  Returns a dict of arrays."""
    f = OracleFluid(fluid)
    z = _c(z)
    n = z.shape[1]
    P = _c(P)
    Tarr = None if np.isscalar(T) else _c(T)
    T0 = float(T) if np.isscalar(T) else 0.0
    out = dict(x=np.zeros((f.nc, n)), y=np.zeros((f.nc, n)), K=np.zeros((f.nc, n)),
               L=np.zeros(n), label=np.zeros(n, np.int8), Zl=np.zeros(n),
               Zg=np.zeros(n), iters=np.zeros(n, np.int32))
    lib().or_flash_batch(
        f.ref, C.c_int64(n), _p(Tarr), C.c_double(T0), _p(P), _p(z), C.c_double(sw),
        _p(out["x"]), _p(out["y"]), _p(out["K"]), _p(out["L"]),
        out["label"].ctypes.data_as(C.c_void_p), _p(out["Zl"]), _p(out["Zg"]),
        out["iters"].ctypes.data_as(C.c_void_p), C.c_int(nthreads))
    return out


def psat_batch(fluid, T, P, kind, x, y, nthreads=0):
    """kind[c]: 1 bubble point of liquid x (vapour store y), -1 dew point of y."""
    f = OracleFluid(fluid)
    x = _c(x)
    y = _c(y)
    n = x.shape[1]
    P = _c(P)
    kind = _c(kind, np.int8)
    Tarr = None if np.isscalar(T) else _c(T)
    T0 = float(T) if np.isscalar(T) else 0.0
    psat = np.zeros(n)
    xm = np.zeros((f.nc, n))
    lib().or_psat_batch(f.ref, C.c_int64(n), _p(Tarr), C.c_double(T0), _p(P),
                        kind.ctypes.data_as(C.c_void_p), _p(x), _p(y), _p(psat),
                        _p(xm), C.c_int(nthreads))
    return psat, xm


def lu_solve(A, b):
    A = _c(A).copy()
    n = A.shape[0]
    b = _c(b)
    x = np.zeros(n)
    info = lib().or_lu_solve(n, _p(A), _p(b), _p(x))
    return x, info
