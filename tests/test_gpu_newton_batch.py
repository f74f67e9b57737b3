"""One Newton-iteration evaluation of several realizations of different grid
sizes (batched flash, property sweep + residual, PhaseCtrl on scratch copies,
Jacobian in either form) captured in one hipGraph.

Bars: every replay -- with the outputs clobbered and eager work between
replays -- reproduces the eager run bitwise (the kernels are deterministic
and atomic-free); the batch's residual and reduced system equal, bitwise,
those of each grid assembled on its own (the shared workspace changes
nothing); the batched flash equals the flash of each grid's cells on its
own; the scratch PhaseCtrl leaves the state untouched and its labels equal
PhaseCtrl run on copies outside the batch."""
import numpy as np
import pytest
import torch

from osiris_amd import synth
from tests.asm_cases import oracle_flash_fn

pytestmark = pytest.mark.gpu


def _cases(oracle, fl):
    shapes = [(6, 5, 3), (9, 4, 2), (5, 7, 4), (3, 3, 1)]
    return [synth.dual_porosity_case(fl, nx, ny, nz, 20 + i, oracle_flash_fn(oracle, fl), p_range=(8e6, 20e6))
            for i, (nx, ny, nz) in enumerate(shapes)]


def _snapshot(batch):
    snap = {"flash": {k: v.clone() for k, v in batch.fout.items()}}
    for i, it in enumerate(batch.items):
        snap[i] = {"F": it["F"].clone(), "psat": it["psat"].clone(),
                   "labels": [s["label"].clone() for s in it["scratch"]]}
        if batch.form == "csr":
            snap[i]["vals"] = it["vals"].clone()
            snap[i]["cols"] = it["cols"].clone()
        else:
            snap[i]["red"] = [t.clone() for t in it["red"]]
    return snap


def _equal(a, b):
    return torch.equal(a.nan_to_num(7.0), b.nan_to_num(7.0)) if a.is_floating_point() else torch.equal(a, b)


def _same(s0, s1):
    for k in s0["flash"]:
        assert _equal(s0["flash"][k], s1["flash"][k]), k
    for i in s0:
        if i == "flash":
            continue
        for key, v in s0[i].items():
            if isinstance(v, list):
                for a, b in zip(v, s1[i][key]):
                    assert _equal(a, b), (i, key)
            else:
                assert _equal(v, s1[i][key]), (i, key)


@pytest.mark.parametrize("form", ["reduced", "csr"])
def test_newton_iteration_graph(gpu, oracle, form):
    from osiris_amd.batch import NewtonBatch
    fl = synth.fluid_pseudo(6, synth.PR, 3)
    cases = _cases(oracle, fl)
    ctx = gpu.Context(fl)
    batch = NewtonBatch(ctx, fl, cases, form=form, flash_mode="mixed")
    state0 = [{k: v.clone() for k, v in it["d"].items() if torch.is_tensor(v)} for it in batch.items]
    batch.run_eager()
    torch.cuda.synchronize()
    eager = _snapshot(batch)
    # the eager run is finite where the reference's arithmetic is: every
    # residual row, Jacobian entry, reduced block and right-hand side of
    # these states, every flash x and L, y of the two-phase cells
    for i in range(len(batch.items)):
        assert torch.isfinite(eager[i]["F"]).all(), i
        if form == "csr":
            assert torch.isfinite(eager[i]["vals"]).all(), i
        else:
            assert torch.isfinite(eager[i]["red"][2]).all() and torch.isfinite(eager[i]["red"][3]).all(), i
    fo = eager["flash"]
    assert torch.isfinite(fo["x"]).all() and torch.isfinite(fo["L"]).all()
    assert torch.isfinite(fo["y"][:, fo["label"] == 0]).all()
    batch.capture()
    for rep in range(3):
        for it in batch.items:            # clobber the outputs between replays
            it["F"].fill_(np.nan)
            it["psat"].fill_(np.nan)
            if form == "csr":
                it["vals"].fill_(np.nan)
            else:
                it["red"][2].fill_(np.nan)
                it["red"][3].fill_(np.nan)
        batch.fout["x"].fill_(np.nan)
        torch.cuda.synchronize()
        batch.replay()
        torch.cuda.synchronize()
        _same(eager, _snapshot(batch))
        _ = torch.ones(1000, device="cuda").sum().item()     # eager work between replays
    # the state being linearised is untouched by the scratch PhaseCtrl
    for it, s0 in zip(batch.items, state0):
        for k, v in s0.items():
            assert _equal(it["d"][k], v), k
    # each grid on its own, with its own workspace
    for i, ((grid, st), it) in enumerate(zip(cases, batch.items)):
        asm = gpu.Assembler(ctx, fl, grid)
        d = asm.device_state(st)
        asm.layout(d)
        F = asm.residual(d)
        assert _equal(F, eager[i]["F"]), i
        if form == "reduced":
            red = asm.reduced_direct(d, F)
            for a, b in zip(red, eager[i]["red"]):
                assert _equal(a, b), i
        else:
            rp, cols, vals = asm.jacobian(d)
            assert torch.equal(cols, eager[i]["cols"]) and _equal(vals, eager[i]["vals"]), i
        # PhaseCtrl on copies, outside the batch (Pg from the property table)
        P = asm.props(d)
        for k, sfx in enumerate(("_m", "_f")):
            lab, x, y, S = (d["label" + sfx].clone(), d["x" + sfx].clone(), d["y" + sfx].clone(),
                            d["S" + sfx].clone())
            Pg = P[3, k * grid.n:(k + 1) * grid.n].contiguous()
            ps = gpu.phase_ctrl(ctx, grid.T, lab, d["Po" + sfx], Pg, x, y, S)
            assert torch.equal(lab, eager[i]["labels"][k]), (i, k)
            assert _equal(ps, eager[i]["psat"][k]), (i, k)
        asm.close()
    # batched flash = flash of each grid's cells on their own
    off = 0
    for grid, st in cases:
        for z, P in ((st.z_m, st.Po_m), (st.z_f, st.Po_f)):
            n = grid.n
            o = gpu.flash(ctx, grid.T, torch.as_tensor(P).cuda(), torch.as_tensor(z).cuda(), mode="mixed")
            for k in ("x", "y", "L", "label"):
                ref = o[k]
                got = eager["flash"][k][..., off:off + n]
                assert _equal(got.contiguous(), ref), k
            off += n


def test_captured_graph_pins_the_context_workspaces(gpu, oracle):
    """While a NewtonBatch graph exists, a call that would grow the
    context's flash park store refuses (OSR_E_SIZE) instead of freeing the
    buffer the graph addresses; closing the batch lifts it."""
    from osiris_amd.batch import NewtonBatch
    fl = synth.fluid_pseudo(6, synth.PR, 3)
    cases = _cases(oracle, fl)[:2]
    ctx = gpu.Context(fl)
    batch = NewtonBatch(ctx, fl, cases, form="reduced", flash_mode="persistent")
    batch.capture()
    batch.replay()
    torch.cuda.synchronize()
    n = 4 * batch.ncells
    g = synth.rng(3)
    P = torch.as_tensor(g.uniform(8e6, 20e6, n)).cuda()
    z = torch.as_tensor(g.dirichlet(np.full(fl.nc, 2.0), n).T.copy()).cuda()
    print("frozen workspaces: flash of", n, "cells after a", batch.ncells, "-cell capture", flush=True)
    with pytest.raises(gpu.OsrError, match="pinned"):
        gpu.flash(ctx, cases[0][0].T, P, z, mode="persistent")
    batch.replay()                      # the graph still runs on its workspaces
    torch.cuda.synchronize()
    batch.close()
    out = gpu.flash(ctx, cases[0][0].T, P, z, mode="persistent")
    torch.cuda.synchronize()
    assert torch.isfinite(out["L"]).all()
    ctx.close()
