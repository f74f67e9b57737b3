"""The allocation-free flash entry point (osr_flash_device_noalloc) against
the allocating one (osr_flash_device), called through the C ABI.

Both run the same kernels; the allocation-free form keeps the parked-cell
store in a fixed per-device workspace instead of a buffer the library grows
with hipMalloc.  Bar: labels, pass counts and RR iteration counts identical;
every floating-point output within 1e-12 relative (the kernels are the same,
so the results are expected to agree bit for bit), with identical NaN/inf
patterns.  Two calls in a row per mode, the second on fewer cells, so the
second call runs on a workspace that still holds the first call's contents.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from osiris_amd import synth

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "K", "L", "label", "Zl", "Zg", "so", "sg", "rho_l", "rho_g", "iters", "rr_iters")
MODES = {"persistent": 1, "mixed": 2}


def _call(gpu, fn, ctx, T, P, z, sw, mode):
    nc, n = z.shape
    out = gpu.alloc_flash_out(nc, n)
    for v in out.values():
        v.fill_(-7)
    fo = gpu._FlashOut(*[(out[k].data_ptr() if k in out else None) for k, _ in gpu._FlashOut._fields_])
    Tp, T0 = (None, float(T)) if np.isscalar(T) else (C.c_void_p(T.data_ptr()), 0.0)
    rc = fn(ctx.h, n, Tp, T0, C.c_void_p(P.data_ptr()), C.c_void_p(z.data_ptr()), sw, C.byref(fo),
            MODES[mode], None)
    torch.cuda.synchronize()
    return rc, out


def _same(a, b, what):
    for k in FIELDS:
        g, r = a[k].cpu().numpy(), b[k].cpu().numpy()
        if g.dtype.kind in "iu":
            assert np.array_equal(g, r), f"{what} {k}: integer outputs differ"
        else:
            assert np.array_equal(np.isfinite(g), np.isfinite(r)), f"{what} {k}: NaN/inf pattern differs"
            f = np.isfinite(r)
            e = np.abs(g[f] - r[f]) / np.maximum(np.abs(r[f]), 1e-300)
            assert e.size == 0 or e.max() <= 1e-12, f"{what} {k}: max rel diff {e.max():.3e}"


@pytest.mark.parametrize("mode", ["persistent", "mixed"])
def test_noalloc_flash_matches_allocating_entry_point(gpu, mode):
    L = gpu.lib()
    fl = synth.fluid_pseudo(8, synth.PR, 1)
    ctx = gpu.Context(fl)
    for n, seed in ((20000, 11), (5000, 12)):
        # cells near their saturation pressure, so that many iterate long
        # enough to be parked
        cells = synth.flash_cells(fl, n, seed, gpu.psat_numpy(ctx))
        T = torch.tensor(cells.T, dtype=torch.float64, device="cuda")
        P = torch.tensor(cells.P, dtype=torch.float64, device="cuda")
        z = torch.tensor(cells.z, dtype=torch.float64, device="cuda")
        print(f"noalloc flash: mode {mode}, n {n}, seed {seed}", flush=True)
        rc0, ref = _call(gpu, L.osr_flash_device, ctx, T, P, z, 0.2, mode)
        rc1, out = _call(gpu, L.osr_flash_device_noalloc, ctx, T, P, z, 0.2, mode)
        assert rc0 >= 0 and rc1 == rc0, (rc0, rc1)
        assert (ref["iters"].cpu().numpy() > 40).any(), "no cell reached the park threshold"
        _same(out, ref, f"{mode} n={n}")
    ctx.close()


def test_noalloc_flash_rejects_oversize_input(gpu):
    L = gpu.lib()
    fl = synth.fluid_c1_nc4_nc10()
    ctx = gpu.Context(fl)
    cap = L.osr_flash_noalloc_max_cells(ctx.nc, MODES["persistent"])
    n = cap + 1
    # buffers of the rejected size, so that nothing could go out of bounds
    P = torch.full((n,), 1e7, dtype=torch.float64, device="cuda")
    z = torch.full((ctx.nc, n), 1.0 / ctx.nc, dtype=torch.float64, device="cuda")
    print(f"noalloc flash: oversize n {n}", flush=True)
    rc, out = _call(gpu, L.osr_flash_device_noalloc, ctx, 360.0, P, z, 0.0, "persistent")
    assert rc == -4
    assert bool((out["L"] == -7).all()), "a rejected call wrote its outputs"
    ctx.close()
