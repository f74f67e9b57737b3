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
                ("rho_g", C.c_double), ("label", C.c_int), ("iters", C.c_int), ("rr_iters", C.c_int)]


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
                rho_l=o.rho_l, rho_g=o.rho_g, label=o.label, iters=o.iters, rr_iters=o.rr_iters)


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
               Zg=np.zeros(n), iters=np.zeros(n, np.int32), rr_iters=np.zeros(n, np.int32))
    lib().or_flash_batch_rr(
        f.ref, C.c_int64(n), _p(Tarr), C.c_double(T0), _p(P), _p(z), C.c_double(sw),
        _p(out["x"]), _p(out["y"]), _p(out["K"]), _p(out["L"]),
        out["label"].ctypes.data_as(C.c_void_p), _p(out["Zl"]), _p(out["Zg"]),
        out["iters"].ctypes.data_as(C.c_void_p), out["rr_iters"].ctypes.data_as(C.c_void_p), C.c_int(nthreads))
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


# ------------------------------------------------------------------ assembly
class _FluidExt(C.Structure):
    _fields_ = [("MW", _dp), ("Vc", _dp), ("Zc", _dp), ("parachor", _dp)]


_i8p = C.POINTER(C.c_int8)
_i32p = C.POINTER(C.c_int32)


class _Wells(C.Structure):
    _fields_ = [("nw", C.c_int32), ("welli", _i32p), ("wellf", _dp), ("condition", _i8p),
                ("sim_time", C.c_double), ("re_f", C.c_double)]


class _Grid(C.Structure):
    _fields_ = [("dx", _dp), ("dy", _dp), ("dz", _dp), ("phi_m", _dp), ("phi_f", _dp),
                ("perm_m", _dp), ("perm_f", _dp), ("tor", _dp), ("swt", _dp), ("sgt", _dp),
                ("refP", C.c_double), ("cpor", C.c_double), ("dcpor", C.c_double),
                ("sigma", C.c_double), ("lf", C.c_double), ("cdiff", C.c_double),
                ("T", C.c_double), ("dt", C.c_double), ("wat_ro", C.c_double),
                ("wat_mu", C.c_double), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("const_diff", C.c_int32), ("nswt", C.c_int32), ("nsgt", C.c_int32),
                ("wells", C.POINTER(_Wells))]


class _State(C.Structure):
    _fields_ = [("label_m", _i8p), ("label_f", _i8p), ("Po_m", _dp), ("Po_f", _dp),
                ("x_m", _dp), ("y_m", _dp), ("x_f", _dp), ("y_f", _dp), ("S_m", _dp),
                ("S_f", _dp), ("rho_prev_m", _dp), ("pre_m", _dp), ("pre_f", _dp),
                ("pcoeff", C.c_double)]


class _Asm:
    """Keeps the numpy buffers alive behind the C structs of one case."""

    def __init__(self, fluid, grid, state):
        self.fluid = OracleFluid(fluid)
        g = grid
        self.keep = []

        def P(a, dt=np.float64):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            self.keep.append(a)
            return a.ctypes.data_as({np.float64: _dp, np.int8: _i8p, np.int32: _i32p}[dt])

        self.ext = _FluidExt(P(fluid.MW), P(fluid.Vc), P(fluid.Zc), P(fluid.parachor))
        self.grid = _Grid(P(g.dx), P(g.dy), P(g.dz), P(g.phi_m), P(g.phi_f), P(g.perm_m),
                          P(g.perm_f), P(g.tor), P(g.swt), P(g.sgt), g.refP, g.cpor, g.dcpor,
                          g.sigma, g.lf, g.cdiff, g.T, g.dt, g.wat_ro, g.wat_mu, g.nx, g.ny,
                          g.nz, int(g.const_diff), len(g.swt), len(g.sgt), None)
        w = getattr(g, "wells", None)
        if w is not None and w.nw > 0:
            self.wells = _Wells(w.nw, P(w.welli, np.int32), P(w.wellf), P(w.condition, np.int8),
                                float(w.sim_time), float(w.re_f))
            self.grid.wells = C.pointer(self.wells)
        s = state
        self.state = _State(P(s.label_m, np.int8), P(s.label_f, np.int8), P(s.Po_m), P(s.Po_f),
                            P(s.x_m), P(s.y_m), P(s.x_f), P(s.y_f), P(s.S_m), P(s.S_f),
                            P(s.rho_prev_m), P(getattr(s, "pre_m", None)),
                            P(getattr(s, "pre_f", None)), s.pcoeff)
        self.n = g.nx * g.ny * g.nz
        self.nc = fluid.nc


def layout(fluid, grid, state):
    a = _Asm(fluid, grid, state)
    bs = np.zeros(2 * a.n, np.int32)
    off = np.zeros(2 * a.n + 1, np.int64)
    lib().or_layout(C.byref(a.grid), C.byref(a.state), a.nc, bs.ctypes.data_as(C.c_void_p),
                    off.ctypes.data_as(C.c_void_p))
    return bs, off


def residual(fluid, grid, state):
    a = _Asm(fluid, grid, state)
    _, off = layout(fluid, grid, state)
    F = np.zeros(int(off[-1]))
    lib().or_residual(a.fluid.ref, C.byref(a.ext), C.byref(a.grid), C.byref(a.state), _p(F))
    return F


def props(fluid, grid, state):
    a = _Asm(fluid, grid, state)
    w = 16 + 2 * a.nc
    out = np.zeros((2, a.n, w))
    lib().or_props(a.fluid.ref, C.byref(a.ext), C.byref(a.grid), C.byref(a.state), _p(out))
    return out


def well_data(fluid, grid, state):
    """The grid's wells: (re [nw] from Calc_Re, inj [nw, 2nc + 6] from
    FlashInjection: x, y, oil/gas fraction, density, viscosity)."""
    a = _Asm(fluid, grid, state)
    nw = grid.wells.nw
    re = np.zeros(nw)
    inj = np.zeros((nw, 2 * a.nc + 6))
    lib().or_well_data(a.fluid.ref, C.byref(a.ext), C.byref(a.grid), _p(re), _p(inj))
    return re, inj


def iftran(fluid, grid, state):
    """IFK_Calc's face coefficients kA/h: [2 continua, 3 directions, n]."""
    a = _Asm(fluid, grid, state)
    out = np.zeros((2, 3, a.n))
    lib().or_iftran(a.fluid.ref, C.byref(a.ext), C.byref(a.grid), _p(out))
    return out


def jacobian(fluid, grid, state):
    """Reference-layout Jacobian as (rowptr, cols, vals)."""
    a = _Asm(fluid, grid, state)
    _, off = layout(fluid, grid, state)
    nrow = int(off[-1])
    rowptr = np.zeros(nrow + 1, np.int64)
    cap = nrow * (8 * (2 * a.nc + 4))
    cols = np.zeros(cap, np.int32)
    vals = np.zeros(cap)
    L = lib()
    L.or_jacobian.restype = C.c_int64
    nnz = L.or_jacobian(a.fluid.ref, C.byref(a.ext), C.byref(a.grid), C.byref(a.state),
                        rowptr.ctypes.data_as(C.c_void_p), cols.ctypes.data_as(C.c_void_p),
                        _p(vals), C.c_int64(cap))
    assert nnz >= 0, "Jacobian capacity too small"
    return rowptr, cols[:nnz].copy(), vals[:nnz].copy()


def jacobian_blocks(fluid, grid, state, blocks):
    """Rows of the listed blocks (k n + c) of the exact Jacobian, consecutively
    in list order: (rowptr, cols, vals) with global column indices."""
    a = _Asm(fluid, grid, state)
    bs, off = layout(fluid, grid, state)
    blocks = np.ascontiguousarray(blocks, dtype=np.int32)
    nrow = int(bs[blocks].sum())
    rowptr = np.zeros(nrow + 1, np.int64)
    cap = nrow * (8 * (2 * a.nc + 4))
    cols = np.zeros(cap, np.int32)
    vals = np.zeros(cap)
    L = lib()
    L.or_jacobian_blocks.restype = C.c_int64
    nnz = L.or_jacobian_blocks(a.fluid.ref, C.byref(a.ext), C.byref(a.grid), C.byref(a.state),
                               blocks.ctypes.data_as(C.c_void_p), C.c_int32(len(blocks)),
                               rowptr.ctypes.data_as(C.c_void_p), cols.ctypes.data_as(C.c_void_p),
                               _p(vals), C.c_int64(cap))
    assert nnz >= 0, nnz
    return rowptr, cols[:nnz].copy(), vals[:nnz].copy()


def jacobian_ref(fluid, grid, state, nthreads=1):
    """The reference's hand-written Jacobian (IteratorJacobian restated) as
    (rowptr, cols, vals) in the reference layout; nthreads OpenMP threads
    (<= 0: all), same output for any count."""
    a = _Asm(fluid, grid, state)
    _, off = layout(fluid, grid, state)
    nrow = int(off[-1])
    rowptr = np.zeros(nrow + 1, np.int64)
    cap = nrow * (8 * (2 * a.nc + 4) + 1)
    cols = np.zeros(cap, np.int32)
    vals = np.zeros(cap)
    L = lib()
    L.or_jacobian_ref_mt.restype = C.c_int64
    nnz = L.or_jacobian_ref_mt(a.fluid.ref, C.byref(a.ext), C.byref(a.grid), C.byref(a.state),
                               rowptr.ctypes.data_as(C.c_void_p), cols.ctypes.data_as(C.c_void_p),
                               _p(vals), C.c_int64(cap), C.c_int(nthreads))
    assert nnz != -1, "Jacobian capacity too small"
    assert nnz != -2, "a row emits a column outside the matrix"
    return rowptr, cols[:nnz].copy(), vals[:nnz].copy()


def reduced_split(nc, label):
    """Local indices of one block's (primary unknowns, secondary unknowns,
    balance rows, local rows) in the reference layout: two-phase primaries
    x_0..x_{nc-3}, P, Sw, So; single-phase x_0..x_{nc-2} (or y), P, Sw."""
    if label == 0:
        prim = list(range(nc - 2)) + [2 * nc, 2 * nc + 1, 2 * nc + 2]
        sec = [nc - 2, nc - 1] + list(range(nc, 2 * nc)) + [2 * nc + 3]
        mrow = list(range(nc, 2 * nc + 1))
        lrow = list(range(nc)) + [2 * nc + 1, 2 * nc + 2, 2 * nc + 3]
    else:
        prim = list(range(nc - 1)) + [nc, nc + 1]
        sec = [nc - 1, nc + 2]
        mrow = list(range(nc + 1))
        lrow = [nc + 1, nc + 2]
    return prim, sec, mrow, lrow


def reduced_from_csr(fluid, grid, state, rp, cols, vals, F):
    """The reduced (nc+1)^2 Newton system as the Schur complement of a
    reference-layout CSR Jacobian (dense per block): for every block,
    S = J_ls^-1 J_lp and t = J_ls^-1 F_l of its local rows, and for every
    pair (a, d) of blocks coupled through a balance row the block
    J_mp,ad - J_ms,ad S_d; rhs_a = -F_m,a + sum_d J_ms,ad t_d.  Returns
    (blocks {(a, d): [nc+1, nc+1]}, rhs [2n, nc+1], S {b: array}, t {b:
    array}, prim {b: global column indices}, sec {b: ...})."""
    import scipy.sparse as sp
    nc, n = fluid.nc, grid.n
    _, off = layout(fluid, grid, state)
    N = int(off[-1])
    J = sp.csr_matrix((vals, cols, rp), shape=(N, N))
    labs = np.concatenate([state.label_m, state.label_f]).astype(int)
    P, Sc, M, Lr = {}, {}, {}, {}
    S, t = {}, {}
    for b in range(2 * n):
        pr, se, mr, lr = reduced_split(nc, labs[b])
        o = int(off[b])
        P[b], Sc[b] = [o + i for i in pr], [o + i for i in se]
        M[b], Lr[b] = [o + i for i in mr], [o + i for i in lr]
        Jl = J[Lr[b]][:, o:o + int(off[b + 1] - off[b])].toarray()
        S[b] = np.linalg.solve(Jl[:, se], Jl[:, pr])
        t[b] = np.linalg.solve(Jl[:, se], F[Lr[b]])
    blocks = {}
    rhs = np.zeros((2 * n, nc + 1))
    blk_of = np.searchsorted(off, np.arange(N), side="right") - 1
    for a in range(2 * n):
        Jm = J[M[a]]
        deps = sorted(set(blk_of[Jm.indices].tolist()))
        cu = np.unique(Jm.indices)
        Jd = Jm[:, cu].toarray()

        def sub(idx):
            out = np.zeros((Jd.shape[0], len(idx)))
            pos = np.searchsorted(cu, idx)
            ok = (pos < len(cu)) & (cu[np.minimum(pos, len(cu) - 1)] == idx)
            out[:, ok] = Jd[:, pos[ok]]
            return out

        r = -F[M[a]]
        for d in deps:
            Js = sub(np.asarray(Sc[d]))
            blocks[(a, d)] = sub(np.asarray(P[d])) - Js @ S[d]
            r = r + Js @ t[d]
        rhs[a] = r
    return blocks, rhs, S, t, P, Sc


def reduced_sampled(fluid, grid, state, sample):
    """The reduced blocks and right-hand sides of the listed row blocks only,
    as reduced_from_csr forms them, from the exact Jacobian rows of those
    blocks and of their dependencies (jacobian_blocks): for grids whose full
    oracle Jacobian is too slow to form.  Returns (blocks {(a, d)}, rhs {a},
    cond {d: 2-norm condition number of d's local-row block J_ls})."""
    import scipy.sparse as sp
    nc, n = fluid.nc, grid.n
    bs, off = layout(fluid, grid, state)
    F = residual(fluid, grid, state)
    labs = np.concatenate([state.label_m, state.label_f]).astype(int)
    nx, ny = grid.nx, grid.ny

    def deps(b):
        k, c = divmod(b, n)
        iz, rem = divmod(c, nx * ny)
        iy, ix = divmod(rem, nx)
        out = [b, (1 - k) * n + c]
        for ok, o in ((ix + 1 < nx, c + 1), (ix > 0, c - 1), (iy + 1 < ny, c + nx), (iy > 0, c - nx),
                      (iz + 1 < grid.nz, c + nx * ny), (iz > 0, c - nx * ny)):
            if ok:
                out.append(k * n + o)
        return out

    sample = [int(b) for b in sample]
    need = sorted(set(sample) | {d for a in sample for d in deps(a)})
    rp, cols, vals = jacobian_blocks(fluid, grid, state, need)
    first = {}
    r = 0
    for b in need:
        first[b] = r
        r += int(bs[b])
    N = int(off[-1])
    J = sp.csr_matrix((vals, cols, rp), shape=(len(rp) - 1, N))
    S, t, cond = {}, {}, {}
    for d in need:
        pr, se, mr, lr = reduced_split(nc, labs[d])
        o = int(off[d])
        Jl = J[[first[d] + i for i in lr]][:, o:o + int(bs[d])].toarray()
        A_ls = Jl[:, se]
        S[d] = np.linalg.solve(A_ls, Jl[:, pr])
        t[d] = np.linalg.solve(A_ls, F[[o + i for i in lr]])
        cond[d] = float(np.linalg.cond(A_ls))
    blocks, rhs = {}, {}
    for a in sample:
        pr, se, mr, lr = reduced_split(nc, labs[a])
        Jm = J[[first[a] + i for i in mr]].toarray()
        rr = -F[[int(off[a]) + i for i in mr]]
        for d in deps(a):
            prd, sed, _, _ = reduced_split(nc, labs[d])
            o = int(off[d])
            Js = Jm[:, [o + i for i in sed]]
            blocks[(a, d)] = Jm[:, [o + i for i in prd]] - Js @ S[d]
            rr = rr + Js @ t[d]
        rhs[a] = rr
    return blocks, rhs, cond


def pack(fluid, grid, state):
    a = _Asm(fluid, grid, state)
    _, off = layout(fluid, grid, state)
    X = np.zeros(int(off[-1]))
    lib().or_pack(C.byref(a.grid), C.byref(a.state), a.nc, _p(X))
    return X


def unpack(fluid, grid, state, X):
    """UpdateAll: returns a new state object with arrays updated from X."""
    import copy
    s2 = copy.copy(state)
    for k in ("Po_m", "Po_f", "x_m", "y_m", "x_f", "y_f", "S_m", "S_f"):
        setattr(s2, k, np.array(getattr(state, k), dtype=np.float64, copy=True))
    a = _Asm(fluid, grid, s2)
    X = _c(X)
    # the mutable view aliases the same buffers the _Asm holds
    lib().or_unpack(C.byref(a.grid), C.byref(a.state), a.nc, _p(X))
    names = ("Po_m", "Po_f", "x_m", "y_m", "x_f", "y_f", "S_m", "S_f")
    ptrs = {k: getattr(a.state, k) for k in names}
    n = a.n
    sizes = {"Po_m": n, "Po_f": n, "x_m": a.nc * n, "y_m": a.nc * n, "x_f": a.nc * n,
             "y_f": a.nc * n, "S_m": 3 * n, "S_f": 3 * n}
    for k in names:
        arr = np.ctypeslib.as_array(ptrs[k], shape=(sizes[k],)).copy()
        setattr(s2, k, arr.reshape(np.shape(getattr(state, k))))
    return s2


def phase_ctrl(fluid, T, label, Po, Pg, x, y, S, nthreads=0):
    """PhaseCtrl over one continuum; returns (label, x, y, S, psat) copies."""
    f = OracleFluid(fluid)
    label = np.array(label, dtype=np.int8, copy=True)
    x = np.array(x, dtype=np.float64, copy=True)
    y = np.array(y, dtype=np.float64, copy=True)
    S = np.array(S, dtype=np.float64, copy=True)
    n = label.shape[0]
    Po = _c(Po)
    Pg = _c(Pg)
    Tarr = None if np.isscalar(T) else _c(T)
    T0 = float(T) if np.isscalar(T) else 0.0
    psat = np.zeros(n)
    lib().or_phase_ctrl(f.ref, C.c_int64(n), _p(Tarr), C.c_double(T0),
                        label.ctypes.data_as(C.c_void_p), _p(Po), _p(Pg), _p(x), _p(y), _p(S),
                        _p(psat), C.c_int(nthreads))
    return label, x, y, S, psat


def residual_scale(fluid, grid, state):
    """Per-row sum of |terms| of the residual (yardstick for parity)."""
    a = _Asm(fluid, grid, state)
    _, off = layout(fluid, grid, state)
    F = np.zeros(int(off[-1]))
    lib().or_residual_scale(a.fluid.ref, C.byref(a.ext), C.byref(a.grid), C.byref(a.state), _p(F))
    return F


def rst_bytes(nx, ny, nz, nc):
    lib().or_rst_bytes.restype = C.c_int64
    return int(lib().or_rst_bytes(nx, ny, nz, nc))


def rst_write(nx, ny, nz, nc, P_m, P_f, state, sim_time, dt, tsm_ctrl, smethq):
    """WriteRST (Do_Time_Step.cpp:179-235) into bytes.  P_m, P_f: [3, n]
    water, oil, gas pressures; state: labels, S, x, y of both continua."""
    L = lib()
    L.or_rst_write.restype = C.c_int64
    L.or_rst_write.argtypes = [C.c_int] * 4 + [C.c_void_p] * 10 + [C.c_double, C.c_double, C.c_int32,
                                                                    C.c_double, C.c_void_p]
    img = np.zeros(rst_bytes(nx, ny, nz, nc), np.uint8)
    a = {k: np.ascontiguousarray(getattr(state, k), np.int8 if k.startswith("label") else np.float64)
         for k in ("S_m", "x_m", "y_m", "label_m", "S_f", "x_f", "y_f", "label_f")}
    P_m, P_f = _c(P_m), _c(P_f)
    w = L.or_rst_write(nx, ny, nz, nc, _p(P_m), _p(P_f), _p(a["S_m"]), _p(a["x_m"]), _p(a["y_m"]),
                       a["label_m"].ctypes.data_as(C.c_void_p), _p(a["S_f"]), _p(a["x_f"]), _p(a["y_f"]),
                       a["label_f"].ctypes.data_as(C.c_void_p), float(sim_time), float(dt), int(tsm_ctrl),
                       float(smethq), img.ctypes.data_as(C.c_void_p))
    assert w == img.size
    return img.tobytes()


def rst_read(nx, ny, nz, nc, data):
    """ReadRST (Do_Time_Step.cpp:237-294) of bytes: dict of P_m, P_f [3, n],
    S_*, x_*, y_*, label_* and the four scalars."""
    L = lib()
    L.or_rst_read.restype = C.c_int64
    L.or_rst_read.argtypes = [C.c_int] * 4 + [C.c_void_p] * 11 + [C.c_void_p] * 4
    n = nx * ny * nz
    img = np.frombuffer(data, np.uint8)
    assert img.size == rst_bytes(nx, ny, nz, nc)
    out = {"P_m": np.zeros((3, n)), "P_f": np.zeros((3, n))}
    for sfx in ("_m", "_f"):
        out["S" + sfx] = np.zeros((3, n))
        out["x" + sfx] = np.zeros((nc, n))
        out["y" + sfx] = np.zeros((nc, n))
        out["label" + sfx] = np.zeros(n, np.int8)
    t, d, q = C.c_double(), C.c_double(), C.c_double()
    ts = C.c_int32()
    r = L.or_rst_read(nx, ny, nz, nc, img.ctypes.data_as(C.c_void_p), _p(out["P_m"]), _p(out["P_f"]),
                      _p(out["S_m"]), _p(out["x_m"]), _p(out["y_m"]), out["label_m"].ctypes.data_as(C.c_void_p),
                      _p(out["S_f"]), _p(out["x_f"]), _p(out["y_f"]), out["label_f"].ctypes.data_as(C.c_void_p),
                      C.byref(t), C.byref(d), C.byref(ts), C.byref(q))
    assert r == img.size
    out.update(sim_time=t.value, dt=d.value, tsm_ctrl=ts.value, smethq=q.value)
    return out
