"""ctypes binding of libosiris_amd.so and torch-tensor front ends.

Mirrors the reference's per-cell entry points by role: ``flash`` is
SuccFalsh/AllFlash (compsim/Init_Calcs.cpp:571-738, 949-962), ``psat`` is
Bubl_Succ/Dew_Succ (compsim/Minor_NR.cpp:491-1249).  All arrays are torch
tensors on the current HIP device, float64, structure-of-arrays
``[component, cell]``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libosiris_amd.so")

FLASH_PLAIN, FLASH_PERSISTENT, FLASH_MIXED = 0, 1, 2
_MODES = {"plain": FLASH_PLAIN, "persistent": FLASH_PERSISTENT, "mixed": FLASH_MIXED}

_dp = C.POINTER(C.c_double)
_lib = None


class OsrError(RuntimeError):
    pass


class _Fluid(C.Structure):
    _fields_ = [("nc", C.c_int32), ("eos", C.c_int32), ("Tc", _dp), ("Pc", _dp),
                ("omega", _dp), ("kij", _dp)]


class _FlashOut(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in
                ("x", "y", "K", "L", "label", "Zl", "Zg", "so", "sg", "rho_l", "rho_g",
                 "iters", "rr_iters")]


def lib():
    """Load libosiris_amd.so; raise loudly if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OsrError(f"{LIB_PATH} is missing: run osiris_amd/build.py (or "
                           "__graft_entry__.build()) first; there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        L.osr_version.restype = C.c_char_p
        L.osr_ctx_create.argtypes = [C.POINTER(_Fluid), C.POINTER(C.c_void_p)]
        L.osr_ctx_destroy.argtypes = [C.c_void_p]
        L.osr_ctx_freeze_workspaces.argtypes = [C.c_void_p, C.c_int]
        L.osr_flash_device.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_double, C.c_void_p,
                                       C.c_void_p, C.c_double, C.POINTER(_FlashOut), C.c_int,
                                       C.c_void_p]
        L.osr_flash_device_noalloc.argtypes = L.osr_flash_device.argtypes
        L.osr_flash_noalloc_max_cells.argtypes = [C.c_int32, C.c_int32]
        L.osr_flash_noalloc_max_cells.restype = C.c_int64
        L.osr_psat_device.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_double, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p]
        _lib = L
    return _lib


def _check(rc, what):
    if rc < 0:
        raise OsrError(f"{what} failed with status {rc}")
    return rc


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream_ptr(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class Context:
    """Owns an osr_ctx (fluid constants uploaded to the current device)."""

    def __init__(self, fluid):
        self.fluid = fluid
        self.nc = int(fluid.nc)
        self._arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
                      (fluid.Tc, fluid.Pc, fluid.omega, np.asarray(fluid.kij).reshape(-1))]
        s = _Fluid(self.nc, int(fluid.eos), *[a.ctypes.data_as(_dp) for a in self._arrs])
        h = C.c_void_p()
        _check(lib().osr_ctx_create(C.byref(s), C.byref(h)), "osr_ctx_create")
        self.h = h

    def freeze(self, owner):
        """Pin the context's workspaces while `owner` (e.g. a captured
        graph) holds their addresses (osr_ctx_freeze_workspaces); they stay
        pinned until every owner has called unfreeze."""
        holders = self.__dict__.setdefault("_ws_holders", set())
        if not holders:
            _check(lib().osr_ctx_freeze_workspaces(self.h, 1), "osr_ctx_freeze_workspaces")
        holders.add(id(owner))

    def unfreeze(self, owner):
        holders = self.__dict__.get("_ws_holders", set())
        holders.discard(id(owner))
        if not holders and getattr(self, "h", None):
            _check(lib().osr_ctx_freeze_workspaces(self.h, 0), "osr_ctx_freeze_workspaces")

    def close(self):
        if getattr(self, "h", None):
            lib().osr_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tune(self, key, value):
        L = lib()
        L.osr_tune_set.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
        _check(L.osr_tune_set(self.h, key.encode(), float(value)), f"osr_tune_set({key})")


FLASH_STATS = ("parked", "newton", "fallback", "newton_ok", "newton_boundary", "newton_nonfinite",
               "newton_singular", "newton_diverged", "newton_maxit", "newton_rejected")


def flash_stats(ctx, stream=None):
    """Stage counts of the last persistent/mixed flash (dict over FLASH_STATS)."""
    L = lib()
    L.osr_flash_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.c_int32, C.c_void_p]
    arr = (C.c_int64 * len(FLASH_STATS))()
    _check(L.osr_flash_stats(ctx.h, arr, len(FLASH_STATS), _stream_ptr(stream)), "osr_flash_stats")
    return dict(zip(FLASH_STATS, list(arr)))


def _as_dev(t, dtype=None):
    import torch
    if not isinstance(t, torch.Tensor):
        t = torch.as_tensor(np.asarray(t))
    t = t.to(device="cuda", dtype=dtype or t.dtype)
    return t.contiguous()


def alloc_flash_out(nc, n, device="cuda", full=True):
    import torch
    f64 = dict(dtype=torch.float64, device=device)
    out = dict(x=torch.empty(nc, n, **f64), y=torch.empty(nc, n, **f64),
               K=torch.empty(nc, n, **f64), L=torch.empty(n, **f64),
               label=torch.empty(n, dtype=torch.int8, device=device),
               Zl=torch.empty(n, **f64), Zg=torch.empty(n, **f64),
               iters=torch.empty(n, dtype=torch.int32, device=device))
    if full:
        out.update(so=torch.empty(n, **f64), sg=torch.empty(n, **f64),
                   rho_l=torch.empty(n, **f64), rho_g=torch.empty(n, **f64),
                   rr_iters=torch.empty(n, dtype=torch.int32, device=device))
    return out


_FLASH_OUT_SHAPE = dict(x="nc", y="nc", K="nc", L="n", label="n", Zl="n", Zg="n", so="n", sg="n",
                        rho_l="n", rho_g="n", iters="n", rr_iters="n")
_FLASH_OUT_DTYPE = dict(label="int8", iters="int32", rr_iters="int32")


def _check_flash_out(out, nc, n):
    """Every given output buffer has the shape, dtype and layout the
    kernels write: [nc, n] or [n], contiguous, on the GPU."""
    import torch
    for k, v in out.items():
        if k not in _FLASH_OUT_SHAPE:
            raise ValueError(f"unknown flash output {k!r}")
        if v is None:
            continue
        shape = (nc, n) if _FLASH_OUT_SHAPE[k] == "nc" else (n,)
        dt = getattr(torch, _FLASH_OUT_DTYPE.get(k, "float64"))
        if tuple(v.shape) != shape or v.dtype != dt or not v.is_cuda or not v.is_contiguous():
            raise ValueError(f"flash output {k}: need a contiguous {dt} GPU tensor of shape {shape}, got "
                             f"{tuple(v.shape)} {v.dtype} on {v.device}")


def flash(ctx, T, P, z, sw=0.0, mode="persistent", out=None, stream=None, noalloc=False):
    """Flash n cells (SuccFalsh).  T: scalar or [n]; P: [n]; z: [nc, n].
    noalloc=True calls osr_flash_device_noalloc (fixed per-device workspace,
    no device allocation inside the library call; n is bounded by
    osr_flash_noalloc_max_cells)."""
    import torch
    z = _as_dev(z, torch.float64)
    nc, n = z.shape
    if nc != ctx.nc:
        raise ValueError(f"z has {nc} components, context has {ctx.nc}")
    P = _as_dev(P, torch.float64)
    if P.shape != (n,):
        raise ValueError("P must be [n]")
    if np.isscalar(T):
        Tt, T0 = None, float(T)
    else:
        Tt, T0 = _as_dev(T, torch.float64), 0.0
        if Tt.shape != (n,):
            raise ValueError("T must be scalar or [n]")
    if out is None:
        out = alloc_flash_out(nc, n)
    _check_flash_out(out, nc, n)
    fo = _FlashOut(*[(_ptr(out.get(k)).value if out.get(k) is not None else None)
                     for k, _ in _FlashOut._fields_])
    fn = lib().osr_flash_device_noalloc if noalloc else lib().osr_flash_device
    m = _MODES[mode] if isinstance(mode, str) else int(mode)
    rc = fn(ctx.h, n, _ptr(Tt), T0, _ptr(P), _ptr(z), float(sw), C.byref(fo), m, _stream_ptr(stream))
    if rc == -4 and noalloc:
        raise OsrError(f"osr_flash_device_noalloc: {n} cells exceed the fixed workspace "
                       f"({lib().osr_flash_noalloc_max_cells(nc, m)} cells at nc={nc})")
    if rc == -4:
        raise OsrError(f"osr_flash_device: {n} cells need a larger park store while the context's "
                       "workspaces are pinned (a captured graph holds them)")
    _check(rc, "osr_flash_device_noalloc" if noalloc else "osr_flash_device")
    return out


def _bind_ad(L):
    if getattr(L, "_ad_bound", False):
        return
    L.osr_flash_tangent_device.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_double] + [C.c_void_p] * 10
    L.osr_flash_adjoint_device.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_double] + [C.c_void_p] * 10
    L._ad_bound = True


def _temp(T, n):
    import torch
    if np.isscalar(T):
        return None, float(T)
    Tt = _as_dev(T, torch.float64)
    if Tt.shape != (n,):
        raise ValueError("T must be scalar or [n]")
    return Tt, 0.0


def flash_tangent(ctx, T, P, z, K, label, dP=None, dz=None, stream=None):
    """Tangent (dx, dy, dL) of the flash at the converged state (K, label)
    for a perturbation (dP [n], dz [nc, n]); osr_flash_tangent_device."""
    import torch
    L = lib()
    _bind_ad(L)
    z = _as_dev(z, torch.float64)
    nc, n = z.shape
    P = _as_dev(P, torch.float64)
    K = _as_dev(K, torch.float64)
    label = _as_dev(label, torch.int8)
    Tt, T0 = _temp(T, n)
    dP = None if dP is None else _as_dev(dP, torch.float64)
    dz = None if dz is None else _as_dev(dz, torch.float64)
    dx = torch.empty(nc, n, dtype=torch.float64, device="cuda")
    dy = torch.empty_like(dx)
    dL = torch.empty(n, dtype=torch.float64, device="cuda")
    _check(L.osr_flash_tangent_device(ctx.h, n, _ptr(Tt), T0, _ptr(P), _ptr(z), _ptr(K), _ptr(label),
                                      _ptr(dP), _ptr(dz), _ptr(dx), _ptr(dy), _ptr(dL), _stream_ptr(stream)),
           "osr_flash_tangent_device")
    return dx, dy, dL


def flash_adjoint(ctx, T, P, z, K, label, xbar=None, ybar=None, Lbar=None, stream=None):
    """Adjoint (Pbar, zbar) of the flash at the converged state for the
    cotangents of x, y [nc, n] and L [n]; osr_flash_adjoint_device."""
    import torch
    L = lib()
    _bind_ad(L)
    z = _as_dev(z, torch.float64)
    nc, n = z.shape
    P = _as_dev(P, torch.float64)
    K = _as_dev(K, torch.float64)
    label = _as_dev(label, torch.int8)
    Tt, T0 = _temp(T, n)
    xbar = None if xbar is None else _as_dev(xbar, torch.float64)
    ybar = None if ybar is None else _as_dev(ybar, torch.float64)
    Lbar = None if Lbar is None else _as_dev(Lbar, torch.float64)
    Pbar = torch.empty(n, dtype=torch.float64, device="cuda")
    zbar = torch.empty(nc, n, dtype=torch.float64, device="cuda")
    _check(L.osr_flash_adjoint_device(ctx.h, n, _ptr(Tt), T0, _ptr(P), _ptr(z), _ptr(K), _ptr(label),
                                      _ptr(xbar), _ptr(ybar), _ptr(Lbar), _ptr(Pbar), _ptr(zbar),
                                      _stream_ptr(stream)), "osr_flash_adjoint_device")
    return Pbar, zbar


def psat(ctx, T, P, kind, x, y, stream=None):
    """Saturation pressure (kind 1 bubble of x, -1 dew of y).  Returns (psat, xm)."""
    import torch
    x = _as_dev(x, torch.float64)
    y = _as_dev(y, torch.float64)
    nc, n = x.shape
    if nc != ctx.nc:
        raise ValueError(f"x has {nc} components, context has {ctx.nc}")
    P = _as_dev(P, torch.float64)
    kind = _as_dev(kind, torch.int8)
    if tuple(y.shape) != (nc, n) or tuple(P.shape) != (n,) or tuple(kind.shape) != (n,):
        raise ValueError(f"psat: x {tuple(x.shape)}, y {tuple(y.shape)}, P {tuple(P.shape)} and kind "
                         f"{tuple(kind.shape)} must agree on (nc, n) = ({nc}, {n})")
    if np.isscalar(T):
        Tt, T0 = None, float(T)
    else:
        Tt, T0 = _as_dev(T, torch.float64), 0.0
        if tuple(Tt.shape) != (n,):
            raise ValueError("T must be scalar or [n]")
    ps = torch.empty(n, dtype=torch.float64, device="cuda")
    xm = torch.empty(nc, n, dtype=torch.float64, device="cuda")
    _check(lib().osr_psat_device(ctx.h, n, _ptr(Tt), T0, _ptr(P), _ptr(kind), _ptr(x), _ptr(y),
                                 _ptr(ps), _ptr(xm), _stream_ptr(stream)), "osr_psat_device")
    return ps, xm


def psat_numpy(ctx):
    """psat front end with numpy in/out (for the synthetic-input generator)."""
    def fn(T, P, kind, x, y):
        ps, xm = psat(ctx, T, P, kind, x, y)
        return ps.cpu().numpy(), xm.cpu().numpy()
    return fn


# ------------------------------------------------------------------ assembly
class _FluidExt(C.Structure):
    _fields_ = [("MW", _dp), ("Vc", _dp), ("Zc", _dp), ("parachor", _dp)]


class _Grid(C.Structure):
    _fields_ = [("dx", _dp), ("dy", _dp), ("dz", _dp), ("phi_m", _dp), ("phi_f", _dp),
                ("perm_m", _dp), ("perm_f", _dp), ("tor", _dp), ("swt", _dp), ("sgt", _dp),
                ("refP", C.c_double), ("cpor", C.c_double), ("dcpor", C.c_double),
                ("sigma", C.c_double), ("lf", C.c_double), ("cdiff", C.c_double),
                ("T", C.c_double), ("dt", C.c_double), ("wat_ro", C.c_double),
                ("wat_mu", C.c_double), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("const_diff", C.c_int32), ("nswt", C.c_int32), ("nsgt", C.c_int32)]


class _Wells(C.Structure):
    _fields_ = [("nw", C.c_int32), ("welli", C.POINTER(C.c_int32)), ("wellf", _dp),
                ("condition", C.POINTER(C.c_int8)), ("sim_time", C.c_double), ("re_f", C.c_double)]


class _StateDev(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in
                ("label_m", "label_f", "Po_m", "Po_f", "x_m", "y_m", "x_f", "y_f", "S_m", "S_f",
                 "rho_prev_m", "pre_m", "pre_f")] + [("pcoeff", C.c_double)]


def _bind_asm(L):
    if getattr(L, "_asm_bound", False):
        return
    L.osr_asm_create.argtypes = [C.c_void_p, C.POINTER(_FluidExt), C.POINTER(_Grid),
                                 C.POINTER(C.c_void_p)]
    L.osr_asm_destroy.argtypes = [C.c_void_p]
    L.osr_asm_layout.argtypes = [C.c_void_p, C.POINTER(_StateDev), C.c_void_p,
                                 C.POINTER(C.c_int64), C.c_void_p]
    L.osr_residual_device.argtypes = [C.c_void_p, C.POINTER(_StateDev), C.c_void_p, C.c_void_p]
    L.osr_jacobian_rowptr_device.argtypes = [C.c_void_p, C.POINTER(_StateDev), C.c_void_p,
                                             C.POINTER(C.c_int64), C.c_void_p]
    L.osr_jacobian_device.argtypes = [C.c_void_p, C.POINTER(_StateDev), C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p]
    L.osr_jacobian_ref_rowptr_device.argtypes = L.osr_jacobian_rowptr_device.argtypes
    L.osr_jacobian_ref_device.argtypes = L.osr_jacobian_device.argtypes
    L.osr_props_device.argtypes = [C.c_void_p, C.POINTER(_StateDev), C.c_void_p, C.c_void_p]
    # osr_state_mut_dev has the field layout of osr_state_dev
    L.osr_unpack_device.argtypes = [C.c_void_p, C.POINTER(_StateDev), C.c_void_p, C.c_void_p]
    L.osr_pack_device.argtypes = [C.c_void_p, C.POINTER(_StateDev), C.c_void_p, C.c_void_p]
    L.osr_residual_perm_vjp_device.argtypes = [C.c_void_p, C.POINTER(_StateDev), C.c_void_p, C.c_void_p,
                                               C.c_void_p, C.c_void_p]
    L.osr_bsr_reduced_rowptr_device.argtypes = [C.c_void_p, C.POINTER(_StateDev), C.c_void_p,
                                                C.POINTER(C.c_int64), C.c_void_p]
    L.osr_jacobian_bsr_reduced_device.argtypes = [C.c_void_p, C.POINTER(_StateDev)] + [C.c_void_p] * 10
    L.osr_bsr_reduced_expand_device.argtypes = [C.c_void_p, C.POINTER(_StateDev)] + [C.c_void_p] * 5
    L.osr_jacobian_bsr_reduced_direct_device.argtypes = [C.c_void_p, C.POINTER(_StateDev)] + [C.c_void_p] * 8
    L.osr_asm_workspace_size.argtypes = [C.c_void_p, C.POINTER(C.c_size_t)]
    L.osr_asm_set_workspace.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    L.osr_asm_set_wells.argtypes = [C.c_void_p, C.POINTER(_Wells), C.c_void_p]
    L.osr_asm_well_data.argtypes = [C.c_void_p, _dp, _dp]
    L._asm_bound = True


class Assembler:
    """Residual/Jacobian assembly for one grid (an osr_asm) -- the role of
    FormFunction/FormJacobian (compsim/PetscAUX.cpp:59-173)."""

    def __init__(self, ctx, fluid, grid):
        L = lib()
        _bind_asm(L)
        self.ctx = ctx
        self.nc = ctx.nc
        self.n = grid.nx * grid.ny * grid.nz
        self.grid = grid
        keep = []

        def P(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return a.ctypes.data_as(_dp)

        fe = _FluidExt(P(fluid.MW), P(fluid.Vc), P(fluid.Zc), P(fluid.parachor))
        g = _Grid(P(grid.dx), P(grid.dy), P(grid.dz), P(grid.phi_m), P(grid.phi_f),
                  P(grid.perm_m), P(grid.perm_f), P(grid.tor), P(grid.swt), P(grid.sgt),
                  grid.refP, grid.cpor, grid.dcpor, grid.sigma, grid.lf, grid.cdiff, grid.T,
                  grid.dt, grid.wat_ro, grid.wat_mu, grid.nx, grid.ny, grid.nz,
                  int(grid.const_diff), len(grid.swt), len(grid.sgt))
        h = C.c_void_p()
        _check(L.osr_asm_create(ctx.h, C.byref(fe), C.byref(g), C.byref(h)), "osr_asm_create")
        self.h = h
        self.nrows = None
        self.nw = 0
        if getattr(grid, "wells", None) is not None and grid.wells.nw > 0:
            self.set_wells(grid.wells)

    def set_wells(self, wells, stream=None):
        """Complete the wells (a synth.Wells, or None for none) in this grid:
        the residual, Jacobian and reduced-system calls add their terms.
        Call again when sim_time or condition change."""
        nw = 0 if wells is None else wells.nw
        if nw == 0:
            w = _Wells(0, None, None, None, 0.0, 0.0)
            _check(lib().osr_asm_set_wells(self.h, C.byref(w), _stream_ptr(stream)), "osr_asm_set_wells")
            self.nw = 0
            return
        wi = np.ascontiguousarray(wells.welli, dtype=np.int32)
        wf = np.ascontiguousarray(wells.wellf, dtype=np.float64)
        if wi.shape != (nw, 5) or wf.shape != (nw, self.nc + 6):
            raise ValueError("wells: welli must be [nw, 5] and wellf [nw, nc + 6]")
        cond = np.zeros(nw, np.int8) if wells.condition is None else \
            np.ascontiguousarray(wells.condition, dtype=np.int8)
        w = _Wells(nw, wi.ctypes.data_as(C.POINTER(C.c_int32)), wf.ctypes.data_as(_dp),
                   cond.ctypes.data_as(C.POINTER(C.c_int8)), float(wells.sim_time), float(wells.re_f))
        _check(lib().osr_asm_set_wells(self.h, C.byref(w), _stream_ptr(stream)), "osr_asm_set_wells")
        self.nw = nw

    def well_data(self):
        """(re [nw], inj [nw, 2nc + 6]) of the last set_wells: Calc_Re and
        the FlashInjection outputs (x, y, oil/gas fraction, density, viscosity)."""
        re = np.zeros(self.nw)
        inj = np.zeros((self.nw, 2 * self.nc + 6))
        if self.nw:
            _check(lib().osr_asm_well_data(self.h, re.ctypes.data_as(_dp), inj.ctypes.data_as(_dp)),
                   "osr_asm_well_data")
        return re, inj

    def close(self):
        if getattr(self, "h", None):
            lib().osr_asm_destroy(self.h)
            self.h = None
        self._ws = None

    def workspace_bytes(self):
        b = C.c_size_t()
        _check(lib().osr_asm_workspace_size(self.h, C.byref(b)), "osr_asm_workspace_size")
        return b.value

    def set_workspace(self, buf):
        """Use a caller-owned device buffer (a uint8 tensor of at least
        workspace_bytes()) as the Jacobian workspace; None returns to a
        library-owned one.  Assemblers sharing a buffer must run on one stream."""
        if buf is None:
            _check(lib().osr_asm_set_workspace(self.h, None, 0), "osr_asm_set_workspace")
        else:
            _check(lib().osr_asm_set_workspace(self.h, C.c_void_p(buf.data_ptr()), buf.numel() * buf.element_size()),
                   "osr_asm_set_workspace")
        self._ws = buf          # keep the buffer alive while in use

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_state(self, st):
        """Upload a host state (synth.State-like) to device tensors."""
        import torch
        d = {}
        for k in ("label_m", "label_f"):
            d[k] = torch.as_tensor(np.ascontiguousarray(getattr(st, k), np.int8)).cuda()
        for k in ("Po_m", "Po_f", "x_m", "y_m", "x_f", "y_f", "S_m", "S_f", "rho_prev_m",
                  "pre_m", "pre_f"):
            v = getattr(st, k, None)
            d[k] = None if v is None else torch.as_tensor(
                np.ascontiguousarray(v, np.float64)).cuda()
        d["pcoeff"] = float(st.pcoeff)
        return d

    def _sd(self, d):
        return _StateDev(*[(d[k].data_ptr() if d.get(k) is not None else None)
                           for k, _ in _StateDev._fields_[:-1]], d["pcoeff"])

    def layout(self, d, stream=None):
        import torch
        off = torch.empty(2 * self.n + 1, dtype=torch.int64, device="cuda")
        nrows = C.c_int64()
        sd = self._sd(d)
        _check(lib().osr_asm_layout(self.h, C.byref(sd), C.c_void_p(off.data_ptr()),
                                    C.byref(nrows), _stream_ptr(stream)), "osr_asm_layout")
        self.nrows = nrows.value
        return off, nrows.value

    def residual(self, d, out=None, stream=None):
        import torch
        if self.nrows is None:
            self.layout(d, stream)
        F = out if out is not None else torch.empty(self.nrows, dtype=torch.float64, device="cuda")
        sd = self._sd(d)
        _check(lib().osr_residual_device(self.h, C.byref(sd), C.c_void_p(F.data_ptr()),
                                         _stream_ptr(stream)), "osr_residual_device")
        return F

    def jacobian(self, d, stream=None, reference=False):
        """Returns (rowptr, cols, vals) as device tensors: the exact
        derivative of the residual (osr_jacobian_device) or, with
        reference=True, the reference's hand-written Jacobian in its own
        pattern (osr_jacobian_ref_device)."""
        import torch
        if self.nrows is None:
            self.layout(d, stream)
        L = lib()
        rp_fn, fn = ((L.osr_jacobian_ref_rowptr_device, L.osr_jacobian_ref_device) if reference else
                     (L.osr_jacobian_rowptr_device, L.osr_jacobian_device))
        rowptr = torch.empty(self.nrows + 1, dtype=torch.int64, device="cuda")
        nnz = C.c_int64()
        sd = self._sd(d)
        _check(rp_fn(self.h, C.byref(sd), C.c_void_p(rowptr.data_ptr()), C.byref(nnz), _stream_ptr(stream)),
               rp_fn.__name__)
        cols = torch.empty(nnz.value, dtype=torch.int32, device="cuda")
        vals = torch.empty(nnz.value, dtype=torch.float64, device="cuda")
        _check(fn(self.h, C.byref(sd), C.c_void_p(rowptr.data_ptr()), C.c_void_p(cols.data_ptr()),
                  C.c_void_p(vals.data_ptr()), _stream_ptr(stream)), fn.__name__)
        return rowptr, cols, vals

    def pack(self, d, stream=None):
        """InitGuess2: state -> unknown vector X (device tensor [nrows])."""
        import torch
        if self.nrows is None:
            self.layout(d, stream)
        X = torch.empty(self.nrows, dtype=torch.float64, device="cuda")
        sd = self._sd(d)
        _check(lib().osr_pack_device(self.h, C.byref(sd), C.c_void_p(X.data_ptr()),
                                     _stream_ptr(stream)), "osr_pack_device")
        return X

    def unpack(self, d, X, stream=None):
        """UpdateAll's X -> state step, in place on d's device arrays."""
        if self.nrows is None:
            self.layout(d, stream)
        if X.numel() != self.nrows:
            raise ValueError(f"X has {X.numel()} entries, layout has {self.nrows} rows")
        sd = self._sd(d)
        _check(lib().osr_unpack_device(self.h, C.byref(sd), C.c_void_p(X.data_ptr()),
                                       _stream_ptr(stream)), "osr_unpack_device")
        return d

    def _reduced_out(self, d, stream):
        import torch
        nb = 2 * self.n
        np1 = self.nc + 1
        browptr = torch.empty(nb + 1, dtype=torch.int64, device="cuda")
        nnzb = C.c_int64()
        sd = self._sd(d)
        _check(lib().osr_bsr_reduced_rowptr_device(self.h, C.byref(sd), C.c_void_p(browptr.data_ptr()),
                                                   C.byref(nnzb), _stream_ptr(stream)),
               "osr_bsr_reduced_rowptr_device")
        bcols = torch.empty(nnzb.value, dtype=torch.int32, device="cuda")
        bvals = torch.empty(nnzb.value, np1, np1, dtype=torch.float64, device="cuda")
        brhs = torch.empty(nb, np1, dtype=torch.float64, device="cuda")
        S = torch.empty(nb, self.nc + 3, np1, dtype=torch.float64, device="cuda")
        t = torch.empty(nb, self.nc + 3, dtype=torch.float64, device="cuda")
        return browptr, bcols, bvals, brhs, S, t

    def reduced(self, d, F, rowptr, vals, stream=None):
        """Reduced (nc+1)^2 BSR of the Newton system from the full CSR:
        (browptr, bcols, bvals [nnzb, nc+1, nc+1], brhs [2n, nc+1], S, t)."""
        out = self._reduced_out(d, stream)
        browptr, bcols, bvals, brhs, S, t = out
        sd = self._sd(d)
        _check(lib().osr_jacobian_bsr_reduced_device(
            self.h, C.byref(sd), C.c_void_p(F.data_ptr()), C.c_void_p(rowptr.data_ptr()),
            C.c_void_p(vals.data_ptr()), C.c_void_p(browptr.data_ptr()), C.c_void_p(bcols.data_ptr()),
            C.c_void_p(bvals.data_ptr()), C.c_void_p(brhs.data_ptr()), C.c_void_p(S.data_ptr()),
            C.c_void_p(t.data_ptr()), _stream_ptr(stream)), "osr_jacobian_bsr_reduced_device")
        return out

    def reduced_direct(self, d, F, out=None, stream=None):
        """The same reduced system assembled straight from the state (no full
        CSR); out: a tuple from an earlier call to refill in place."""
        if out is None:
            out = self._reduced_out(d, stream)
        browptr, bcols, bvals, brhs, S, t = out
        sd = self._sd(d)
        _check(lib().osr_jacobian_bsr_reduced_direct_device(
            self.h, C.byref(sd), C.c_void_p(F.data_ptr()), C.c_void_p(browptr.data_ptr()),
            C.c_void_p(bcols.data_ptr()), C.c_void_p(bvals.data_ptr()), C.c_void_p(brhs.data_ptr()),
            C.c_void_p(S.data_ptr()), C.c_void_p(t.data_ptr()), _stream_ptr(stream)),
            "osr_jacobian_bsr_reduced_direct_device")
        return out

    def reduced_expand(self, d, S, t, dp, stream=None):
        """Full Newton update [nrows] from the reduced solution dp [2n, nc+1]."""
        import torch
        du = torch.empty(self.nrows, dtype=torch.float64, device="cuda")
        sd = self._sd(d)
        _check(lib().osr_bsr_reduced_expand_device(self.h, C.byref(sd), C.c_void_p(S.data_ptr()),
                                                   C.c_void_p(t.data_ptr()), C.c_void_p(dp.data_ptr()),
                                                   C.c_void_p(du.data_ptr()), _stream_ptr(stream)),
               "osr_bsr_reduced_expand_device")
        return du

    def perm_vjp(self, d, Fbar, stream=None):
        """(g_perm_m, g_perm_f) [3, n]: d(Fbar . F)/d ln perm at fixed state."""
        import torch
        if self.nrows is None:
            self.layout(d, stream)
        if Fbar.numel() != self.nrows:
            raise ValueError(f"Fbar has {Fbar.numel()} entries, layout has {self.nrows} rows")
        gm = torch.empty(3, self.n, dtype=torch.float64, device="cuda")
        gf = torch.empty_like(gm)
        sd = self._sd(d)
        _check(lib().osr_residual_perm_vjp_device(self.h, C.byref(sd), C.c_void_p(Fbar.data_ptr()),
                                                  C.c_void_p(gm.data_ptr()), C.c_void_p(gf.data_ptr()),
                                                  _stream_ptr(stream)), "osr_residual_perm_vjp_device")
        return gm, gf

    def props(self, d, stream=None):
        import torch
        nf = 16 + 2 * self.nc
        out = torch.empty(nf, 2 * self.n, dtype=torch.float64, device="cuda")
        sd = self._sd(d)
        _check(lib().osr_props_device(self.h, C.byref(sd), C.c_void_p(out.data_ptr()),
                                      _stream_ptr(stream)), "osr_props_device")
        return out


def phase_ctrl(ctx, T, label, Po, Pg, x, y, S, stream=None):
    """PhaseCtrl (compsim/Do_Cycle.cpp:1558-1722) on one continuum.  label,
    x, y, S are device tensors updated in place; returns psat."""
    import torch
    L = lib()
    if not getattr(L, "_pc_bound", False):
        L.osr_phase_ctrl_device.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_double,
                                            C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L._pc_bound = True
    n = label.shape[0]
    Po = _as_dev(Po, torch.float64)
    Pg = _as_dev(Pg, torch.float64)
    if np.isscalar(T):
        Tt, T0 = None, float(T)
    else:
        Tt, T0 = _as_dev(T, torch.float64), 0.0
    ps = torch.empty(n, dtype=torch.float64, device="cuda")
    _check(L.osr_phase_ctrl_device(ctx.h, n, _ptr(Tt), T0, _ptr(label), _ptr(Po), _ptr(Pg),
                                   _ptr(x), _ptr(y), _ptr(S), _ptr(ps), _stream_ptr(stream)),
           "osr_phase_ctrl_device")
    return ps


# ------------------------------------------------------------ linear solve
class _KspOpts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("dtol", C.c_double),
                ("maxit", C.c_int32), ("x0_nonzero", C.c_int32)]


class _KspResult(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("reason", C.c_int32), ("rnorm", C.c_double),
                ("bnorm", C.c_double), ("r0norm", C.c_double)]


# KSPConvergedReason codes returned in the result
KSP_REASONS = {2: "converged_rtol", 3: "converged_atol", -3: "diverged_its", -4: "diverged_dtol",
               -5: "diverged_breakdown", -9: "diverged_nanorinf"}
# the reference's KSP tolerances when the deck sets none (Data_Input.cpp:232-266)
KSP_REFERENCE_DEFAULTS = dict(rtol=1e-18, atol=1e-50, dtol=1e3, maxit=5000)


class Ksp:
    """BiCGStab + block-Jacobi on a block-CSR system (osr_ksp_*): the linear
    solve of the reference's SNES (KSPBCGS + PCVPBJACOBI, PetscAUX.cpp:628-646)
    on the reduced (nc+1)^2 system."""

    def __init__(self, bs, nbrows):
        L = lib()
        if not getattr(L, "_ksp_bound", False):
            L.osr_ksp_create.argtypes = [C.c_int32, C.c_int64, C.POINTER(C.c_void_p)]
            L.osr_ksp_destroy.argtypes = [C.c_void_p]
            L.osr_ksp_setup_device.argtypes = [C.c_void_p] + [C.c_void_p] * 4
            L.osr_ksp_solve_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(_KspOpts),
                                               C.POINTER(_KspResult), C.c_void_p]
            L._ksp_bound = True
        h = C.c_void_p()
        _check(L.osr_ksp_create(int(bs), int(nbrows), C.byref(h)), "osr_ksp_create")
        self.h, self.bs, self.nbrows = h, int(bs), int(nbrows)
        self._keep = None

    def close(self):
        if getattr(self, "h", None):
            lib().osr_ksp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setup(self, browptr, bcols, bvals, stream=None):
        """Keep the matrix (device tensors) and invert its diagonal blocks;
        returns the number of singular diagonal blocks."""
        self._keep = (browptr, bcols, bvals)
        return _check(lib().osr_ksp_setup_device(self.h, _ptr(browptr), _ptr(bcols), _ptr(bvals),
                                                 _stream_ptr(stream)), "osr_ksp_setup_device")

    def solve(self, b, x=None, rtol=1e-10, atol=1e-50, dtol=1e3, maxit=1000, stream=None):
        """Solve A x = b; x (device, optional) is the initial guess.  Returns
        (x, {iterations, reason, rnorm, bnorm, r0norm})."""
        import torch
        n = self.bs * self.nbrows
        if b.numel() != n:
            raise ValueError(f"b has {b.numel()} entries, the system has {n}")
        x0 = x is not None
        if x is None:
            x = torch.empty(n, dtype=torch.float64, device="cuda")
        o = _KspOpts(float(rtol), float(atol), float(dtol), int(maxit), int(x0))
        r = _KspResult()
        _check(lib().osr_ksp_solve_device(self.h, _ptr(b), _ptr(x), C.byref(o), C.byref(r),
                                          _stream_ptr(stream)), "osr_ksp_solve_device")
        return x, {"iterations": r.iterations, "reason": KSP_REASONS.get(r.reason, r.reason),
                   "rnorm": r.rnorm, "bnorm": r.bnorm, "r0norm": r.r0norm}


# ------------------------------------------------------------ Newton step
class _NewtonOpts(C.Structure):
    _fields_ = [("ksp", _KspOpts), ("tolmain", C.c_double)]


class _NewtonResult(C.Structure):
    _fields_ = [("accepted", C.c_int32), ("snes_reason", C.c_int32), ("singular_blocks", C.c_int32),
                ("reserved", C.c_int32), ("ksp", _KspResult), ("vnorm", C.c_double), ("nrows", C.c_int64),
                ("nrows_next", C.c_int64)]


class _StateRw(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("label_m", "label_f", "Po_m", "Po_f", "x_m", "y_m", "x_f", "y_f",
                                          "S_m", "S_f", "rho_prev_m")] + [("pcoeff", C.c_double)]


# SNESConvergedReason codes of a SNESKSPONLY step
SNES_REASONS = {5: "converged_its", -3: "diverged_linear_solve"}
# TOLMAIN (Functions.h:241): largest accepted mean |update| per unknown
TOLMAIN = 1e-2


class NewtonStep:
    """One time step's nonlinear solve of the reference (SolveNonLinear,
    PetscAUX.cpp:566-730: SNESKSPONLY -- one Jacobian, one BiCGStab solve,
    X1 = X0 - Y -- the vnorm <= TOLMAIN acceptance, and for an accepted step
    UpdateAll(X, true) with PhaseCtrl, CalcPCoeff and PJac) on the device
    (osr_newton_*), on the reduced (nc+1)^2 system."""

    def __init__(self, asm):
        L = lib()
        if not getattr(L, "_newton_bound", False):
            L.osr_newton_create.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
            L.osr_newton_destroy.argtypes = [C.c_void_p]
            L.osr_newton_step_device.argtypes = [C.c_void_p, C.POINTER(_StateRw), C.POINTER(_NewtonOpts),
                                                 C.c_void_p, C.POINTER(_NewtonResult), C.c_void_p]
            L.osr_newton_solution_update.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_int64), C.c_void_p]
            L._newton_bound = True
        h = C.c_void_p()
        _check(L.osr_newton_create(asm.h, C.byref(h)), "osr_newton_create")
        self.h, self.asm = h, asm

    def close(self):
        if getattr(self, "h", None):
            lib().osr_newton_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, d, tolmain=TOLMAIN, rtol=KSP_REFERENCE_DEFAULTS["rtol"], atol=KSP_REFERENCE_DEFAULTS["atol"],
             dtol=KSP_REFERENCE_DEFAULTS["dtol"], maxit=KSP_REFERENCE_DEFAULTS["maxit"], x0_nonzero=False,
             stream=None):
        """Run one step on the device state d (Assembler.device_state), in
        place when accepted: arrays, d["pcoeff"] and d["rho_prev_m"] (the new
        state's matrix densities, for the next step).  d["pre_m"]/["pre_f"]
        must be None.  Returns the result as a dict."""
        import torch
        if d.get("pre_m") is not None or d.get("pre_f") is not None:
            raise ValueError("the step recomputes preProp per sweep: pre_m/pre_f must be None")
        s = _StateRw(*[None if d.get(k) is None else d[k].data_ptr()
                       for k in ("label_m", "label_f", "Po_m", "Po_f", "x_m", "y_m", "x_f", "y_f", "S_m", "S_f",
                                 "rho_prev_m")], float(d["pcoeff"]))
        rho_next = d.get("rho_prev_m")
        if rho_next is None:
            rho_next = torch.empty(2 * self.asm.n, dtype=torch.float64, device="cuda")
        o = _NewtonOpts(_KspOpts(float(rtol), float(atol), float(dtol), int(maxit), int(bool(x0_nonzero))),
                        float(tolmain))
        r = _NewtonResult()
        _check(lib().osr_newton_step_device(self.h, C.byref(s), C.byref(o), _ptr(rho_next), C.byref(r),
                                            _stream_ptr(stream)), "osr_newton_step_device")
        if r.accepted:
            d["pcoeff"] = s.pcoeff
            d["rho_prev_m"] = rho_next
        self.asm.nrows = int(r.nrows_next)
        return {"accepted": bool(r.accepted), "snes_reason": SNES_REASONS.get(r.snes_reason, r.snes_reason),
                "singular_blocks": r.singular_blocks, "vnorm": r.vnorm, "nrows": r.nrows,
                "nrows_next": r.nrows_next,
                "ksp": {"iterations": r.ksp.iterations, "reason": KSP_REASONS.get(r.ksp.reason, r.ksp.reason),
                        "rnorm": r.ksp.rnorm, "bnorm": r.ksp.bnorm, "r0norm": r.ksp.r0norm}}

    def solution_update(self, stream=None):
        """A copy of the last step's update du (device tensor, X1 = X0 + du)."""
        import torch
        n = C.c_int64()
        L = lib()
        _check(L.osr_newton_solution_update(self.h, None, C.byref(n), _stream_ptr(stream)),
               "osr_newton_solution_update")
        du = torch.empty(int(n.value), dtype=torch.float64, device="cuda")
        _check(L.osr_newton_solution_update(self.h, _ptr(du), C.byref(n), _stream_ptr(stream)),
               "osr_newton_solution_update")
        return du


# ------------------------------------------------------------ restart file
class _RstScalars(C.Structure):
    _fields_ = [("sim_time", C.c_double), ("dt", C.c_double), ("tsm_ctrl", C.c_int32), ("reserved", C.c_int32),
                ("smethq", C.c_double)]


def _bind_rst(L):
    if getattr(L, "_rst_bound", False):
        return
    L.osr_rst_bytes.argtypes = [C.c_int32] * 4 + [C.POINTER(C.c_int64)]
    L.osr_rst_write_device.argtypes = [C.c_void_p, C.POINTER(_StateDev), C.c_void_p, C.POINTER(_RstScalars),
                                       C.c_void_p, C.c_void_p]
    L.osr_rst_read_device.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(_StateRw), C.c_void_p,
                                      C.POINTER(_RstScalars), C.c_void_p]
    L._rst_bound = True


def rst_bytes(nx, ny, nz, nc):
    """Size of the reference's restart file (WriteRST) for this grid."""
    L = lib()
    _bind_rst(L)
    b = C.c_int64()
    _check(L.osr_rst_bytes(nx, ny, nz, nc, C.byref(b)), "osr_rst_bytes")
    return int(b.value)


def rst_write(asm, d, props=None, sim_time=0.0, dt=0.0, tsm_ctrl=0, smethq=0.0, path=None, stream=None):
    """WriteRST (compsim/Do_Time_Step.cpp:179-235): the restart image of the
    device state d as a uint8 device tensor (and, with path, the file).
    props: the property table of d's sweep (default: a fresh sweep)."""
    import torch
    L = lib()
    _bind_rst(L)
    g = asm.grid
    if props is None:
        if asm.nrows is None:
            asm.layout(d, stream)
        props = asm.props(d, stream)
    nb = rst_bytes(g.nx, g.ny, g.nz, asm.nc)
    img = torch.empty((nb + 7) // 8 * 8, dtype=torch.uint8, device="cuda")
    sc = _RstScalars(float(sim_time), float(dt), int(tsm_ctrl), 0, float(smethq))
    sd = asm._sd(d)
    _check(L.osr_rst_write_device(asm.h, C.byref(sd), _ptr(props), C.byref(sc), _ptr(img), _stream_ptr(stream)),
           "osr_rst_write_device")
    img = img[:nb]
    if path is not None:
        (torch.cuda.current_stream() if stream is None else stream).synchronize()
        img.cpu().numpy().tofile(path)
    return img


def rst_read(asm, d, src, stream=None):
    """ReadRST (compsim/Do_Time_Step.cpp:237-294): fill the device state d
    (labels, Po, x, y, S) from a restart image (uint8 device tensor) or file
    path.  Returns (scalars dict, phase pressures [4, n]: Pw_m, Pg_m, Pw_f,
    Pg_f)."""
    import torch
    L = lib()
    _bind_rst(L)
    g = asm.grid
    nb = rst_bytes(g.nx, g.ny, g.nz, asm.nc)
    if isinstance(src, (str, os.PathLike)):
        host = np.fromfile(src, dtype=np.uint8)
        if host.size != nb:
            raise ValueError(f"{src}: {host.size} bytes, a restart of this grid has {nb}")
        img = torch.empty((nb + 7) // 8 * 8, dtype=torch.uint8, device="cuda")
        img[:nb].copy_(torch.from_numpy(host))
    else:
        if src.numel() != nb:
            raise ValueError(f"image has {src.numel()} bytes, a restart of this grid has {nb}")
        img = src
        if img.data_ptr() % 8:
            img = img.clone()
    s = _StateRw(*[d[k].data_ptr() for k in ("label_m", "label_f", "Po_m", "Po_f", "x_m", "y_m", "x_f", "y_f",
                                              "S_m", "S_f")], None, float(d["pcoeff"]))
    pp = torch.empty(4, asm.n, dtype=torch.float64, device="cuda")
    sc = _RstScalars()
    _check(L.osr_rst_read_device(asm.h, _ptr(img), C.byref(s), _ptr(pp), C.byref(sc), _stream_ptr(stream)),
           "osr_rst_read_device")
    return {"sim_time": sc.sim_time, "dt": sc.dt, "tsm_ctrl": sc.tsm_ctrl, "smethq": sc.smethq}, pp
