"""The reference's restart file (WriteRST / ReadRST,
compsim/Do_Time_Step.cpp:179-294) as a byte image.

CPU: the oracle's restatement of the two loops is pinned against an
independent construction of the same bytes with numpy -- the pressure arrays
as C-ordered [Nx+2][Ny+2][Nz+2][3] blocks with a zero halo, the cells as a
packed structured dtype (sat f8[3], comp f8[nc][2], phaseStat i1; itemsize
24 + 16 nc + 1) in Ix-outer/Iz-inner order, the scalars by struct.pack --
and against the size the reference's example deck implies (test.dat: GRID
10 10 5, NC 5 -> 153,412 bytes).  Reading inverts writing exactly.

GPU: the device image (osr_rst_write_device) equals the oracle's image of
the same state and pressures byte for byte, for grids whose image ends in a
partial 4-byte word and for nc = 3 and 10; reading it back
(osr_rst_read_device) restores every state array bitwise, the scalars, and
the water and gas pressures; the same through a file on disk.
"""
import struct

import numpy as np
import pytest

from osiris_amd import synth
from tests.asm_cases import c0_case, oracle_flash_fn


def _numpy_image(nx, ny, nz, nc, P_m, P_f, st, sc):
    n = nx * ny * nz

    def press(P):
        a = np.zeros((nx + 2, ny + 2, nz + 2, 3))
        a[1:-1, 1:-1, 1:-1, :] = np.asarray(P).reshape(3, nz, ny, nx).transpose(3, 2, 1, 0)
        return a.tobytes()

    rec = np.dtype([("S", "<f8", (3,)), ("c", "<f8", (nc, 2)), ("lab", "i1")])
    assert rec.itemsize == 24 + 16 * nc + 1

    def file_order(a, k):
        return np.asarray(a).reshape(k, nz, ny, nx).transpose(3, 2, 1, 0).reshape(n, k)

    def cells(sfx):
        r = np.zeros(n, rec)
        r["S"] = file_order(getattr(st, "S" + sfx), 3)
        r["c"][:, :, 0] = file_order(getattr(st, "x" + sfx), nc)
        r["c"][:, :, 1] = file_order(getattr(st, "y" + sfx), nc)
        r["lab"] = file_order(getattr(st, "label" + sfx), 1)[:, 0]
        return r.tobytes()

    tail = struct.pack("<ddid", sc["sim_time"], sc["dt"], sc["tsm_ctrl"], sc["smethq"])
    return press(P_m) + press(P_f) + cells("_m") + cells("_f") + tail


def _random_state(nx, ny, nz, nc, seed):
    g = synth.rng(seed)
    n = nx * ny * nz

    class S:
        pass

    st = S()
    for sfx in ("_m", "_f"):
        setattr(st, "S" + sfx, g.dirichlet([1, 1, 1], n).T.copy())
        setattr(st, "x" + sfx, g.dirichlet(np.ones(nc), n).T.copy())
        setattr(st, "y" + sfx, g.dirichlet(np.ones(nc), n).T.copy())
        setattr(st, "label" + sfx, g.integers(-1, 2, n).astype(np.int8))
    P_m = g.uniform(5e6, 3e7, (3, n))
    P_f = g.uniform(5e6, 3e7, (3, n))
    return st, P_m, P_f


SC = {"sim_time": 86400.0 * 3.5, "dt": 3600.0, "tsm_ctrl": 7, "smethq": -12.25}


@pytest.mark.parametrize("dims, nc", [((3, 5, 3), 3), ((4, 2, 1), 5), ((2, 3, 4), 10)])
def test_oracle_image_is_the_reference_layout(oracle, dims, nc):
    nx, ny, nz = dims
    st, P_m, P_f = _random_state(nx, ny, nz, nc, seed=nx * 100 + nc)
    img = oracle.rst_write(nx, ny, nz, nc, P_m, P_f, st, **SC)
    assert img == _numpy_image(nx, ny, nz, nc, P_m, P_f, st, SC)
    back = oracle.rst_read(nx, ny, nz, nc, img)
    for k in ("S_m", "x_m", "y_m", "label_m", "S_f", "x_f", "y_f", "label_f"):
        assert np.array_equal(back[k].reshape(-1), np.asarray(getattr(st, k)).reshape(-1)), k
    assert np.array_equal(back["P_m"], P_m) and np.array_equal(back["P_f"], P_f)
    for k, v in SC.items():
        assert back[k] == v, k


def test_restart_size_of_the_example_deck(oracle):
    # test.dat: GRID 10 10 5, NC 5; 2 x 12*12*7 padded cells x 3 doubles,
    # 2 x 500 records of 3 + 2*5 doubles and one byte, 8 + 8 + 4 + 8 bytes
    assert oracle.rst_bytes(10, 10, 5, 5) == 2 * 12 * 12 * 7 * 24 + 2 * 500 * (24 + 80 + 1) + 28 == 153412


def _gpu_case(gpu, oracle, fl, grid, st):
    import torch
    ctx = gpu.Context(fl)
    asm = gpu.Assembler(ctx, fl, grid)
    d = asm.device_state(st)
    asm.layout(d)
    props = asm.props(d)
    torch.cuda.synchronize()
    n = grid.n
    pr = props.cpu().numpy()
    P_m = np.stack([pr[1, :n], d["Po_m"].cpu().numpy(), pr[3, :n]])
    P_f = np.stack([pr[1, n:], d["Po_f"].cpu().numpy(), pr[3, n:]])
    return ctx, asm, d, props, P_m, P_f


@pytest.mark.gpu
def test_device_image_equals_the_oracle(gpu, oracle, tmp_path):
    import torch
    fl10 = synth.fluid_pseudo(10, synth.PR, 5)
    cases = [c0_case(oracle, nx=3, ny=5, nz=3, seed=2),                      # 45 cells: a partial last word
             (fl10,) + synth.dual_porosity_case(fl10, 5, 4, 3, 6, oracle_flash_fn(oracle, fl10),
                                                p_range=(8e6, 20e6))]
    for fl, grid, st in cases:
        name = f"nc{fl.nc} {grid.nx}x{grid.ny}x{grid.nz}"
        print("restart", name, flush=True)
        ctx, asm, d, props, P_m, P_f = _gpu_case(gpu, oracle, fl, grid, st)
        img = gpu.rst_write(asm, d, props, **SC)
        torch.cuda.synchronize()
        ref = oracle.rst_write(grid.nx, grid.ny, grid.nz, fl.nc, P_m, P_f, st, **SC)
        got = img.cpu().numpy().tobytes()
        assert len(got) == len(ref) == gpu.rst_bytes(grid.nx, grid.ny, grid.nz, fl.nc), name
        assert got == ref, (name, int(np.nonzero(np.frombuffer(got, np.uint8) != np.frombuffer(ref, np.uint8))[0][0]))

        # read back into a scrambled state: everything restored bitwise
        d2 = asm.device_state(st)
        for k in ("Po_m", "Po_f", "x_m", "y_m", "x_f", "y_f", "S_m", "S_f"):
            d2[k].fill_(-1.0)
        d2["label_m"].fill_(5)
        d2["label_f"].fill_(5)
        sc, pp = gpu.rst_read(asm, d2, img)
        torch.cuda.synchronize()
        for k in ("label_m", "label_f", "Po_m", "Po_f", "x_m", "y_m", "x_f", "y_f", "S_m", "S_f"):
            assert torch.equal(d2[k], d[k]), (name, k)
        assert sc == SC, name
        pp = pp.cpu().numpy()
        assert np.array_equal(pp, np.stack([P_m[0], P_m[2], P_f[0], P_f[2]])), name

        # through a file
        path = tmp_path / "test.rst"
        gpu.rst_write(asm, d, props, path=str(path), **SC)
        assert path.read_bytes() == ref, name
        d3 = asm.device_state(st)
        d3["Po_m"].fill_(0.0)
        sc3, _ = gpu.rst_read(asm, d3, str(path))
        assert torch.equal(d3["Po_m"], d["Po_m"]) and sc3 == SC, name
        asm.close()
        ctx.close()
