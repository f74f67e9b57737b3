"""CPU checks of the C ABI: the library loads, exports every entry point the
headers declare, and rejects bad arguments before touching a device."""
import ctypes as C
import glob
import os
import re

import numpy as np

from osiris_amd import runtime as rt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names.update(re.findall(r"\b(osr_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    lib = rt.lib()
    names = declared_symbols()
    assert len(names) >= 6
    missing = [s for s in names if not hasattr(lib, s)]
    assert not missing, missing


def test_version_and_supported_nc():
    lib = rt.lib()
    assert b"gfx950" in lib.osr_version()
    for nc in range(2, 13):
        assert lib.osr_supported_nc(nc) == 1
    for nc in (0, 1, 13, 16, 17):
        assert lib.osr_supported_nc(nc) == 0


def test_ctx_create_rejects_bad_arguments():
    lib = rt.lib()
    h = C.c_void_p()
    assert lib.osr_ctx_create(None, C.byref(h)) == -1
    arrs = [np.ones(13) * 300, np.ones(13) * 4e6, np.zeros(13), np.zeros(169)]
    ptrs = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]
    assert lib.osr_ctx_create(C.byref(rt._Fluid(0, 0, *ptrs)), C.byref(h)) == -1
    assert lib.osr_ctx_create(C.byref(rt._Fluid(13, 0, *ptrs)), C.byref(h)) == -3
    assert lib.osr_ctx_create(C.byref(rt._Fluid(13, 5, *ptrs)), C.byref(h)) == -1
    assert lib.osr_flash_device(None, 1, None, 300.0, None, None, 0.0, None, 1, None) == -1


def test_noalloc_flash_limits():
    # 8 Mi doubles of park workspace: (nc + 5) words per cell persistent,
    # (2 nc + 9) mixed, no limit for the plain form.
    lib = rt.lib()
    for nc in range(2, 13):
        assert lib.osr_flash_noalloc_max_cells(nc, 1) == (8 << 20) // (nc + 5)
        assert lib.osr_flash_noalloc_max_cells(nc, 2) == (8 << 20) // (2 * nc + 9)
        assert lib.osr_flash_noalloc_max_cells(nc, 0) == 2 ** 63 - 1
    assert lib.osr_flash_noalloc_max_cells(13, 1) == -3
    assert lib.osr_flash_noalloc_max_cells(6, 3) == -1
    assert lib.osr_flash_device_noalloc(None, 1, None, 300.0, None, None, 0.0, None, 1, None) == -1


def test_freeze_workspaces_rejects_bad_arguments():
    lib = rt.lib()
    assert lib.osr_ctx_freeze_workspaces(None, 1) == -1
