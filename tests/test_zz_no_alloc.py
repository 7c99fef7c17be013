"""The allocation-free entry point fd3_eval_celllist_noalloc.  GPU: its
results equal those of fd3_eval_celllist with a caller workspace, on two
calls in a row, the second with a smaller box (the static scratch keeps the
first call's contents, which must not leak into the second).  CPU: a box
beyond the static scratch is refused before any device work."""
import ctypes as C
import sys
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from fourierd3_amd import FourierD3ASECalculator
from fourierd3_amd import _lib as L
from fourierd3_amd import workloads as W



def _eval(calc, box, noalloc):
    model = calc._model
    dev = torch.device("cuda:0")
    n = box.n_atoms
    pos = torch.tensor(box.positions, dtype=torch.float64, device=dev)
    spec = torch.tensor([{int(z): i for i, z in enumerate(model.species_unique)}[int(v)] for v in box.numbers],
                        dtype=torch.int32, device=dev)
    f64 = dict(dtype=torch.float64, device=dev)
    out = {"energy": torch.full((1,), np.nan, **f64), "grad_pos": torch.full((n, 3), np.nan, **f64),
           "virial": torch.full((3, 3), np.nan, **f64)}
    o = L.Outputs()
    o.energy, o.grad_pos, o.virial = L.ptr(out["energy"]), L.ptr(out["grad_pos"]), L.ptr(out["virial"])
    g = model._geom(box.cell)
    m = model._model_struct()
    stream = torch.cuda.current_stream(dev).cuda_stream
    if noalloc:
        L.check(L.lib().fd3_eval_celllist_noalloc(C.byref(m), C.byref(g), n, L.ptr(pos), L.FD3_DTYPE_F64,
                                                  L.ptr(spec), L.FD3_NEED_GRAD, C.byref(o), stream))
    else:
        nb = C.c_size_t()
        L.check(L.lib().fd3_eval_celllist_workspace(C.byref(m), C.byref(g), n, L.FD3_NEED_GRAD, C.byref(nb)))
        ws = torch.empty(nb.value, dtype=torch.uint8, device=dev)
        L.check(L.lib().fd3_eval_celllist(C.byref(m), C.byref(g), n, L.ptr(pos), L.FD3_DTYPE_F64, L.ptr(spec),
                                          L.FD3_NEED_GRAD, L.ptr(ws), ws.numel(), C.byref(o), stream))
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


@pytest.mark.gpu
def test_noalloc_entry_point_matches_workspace_entry_point():
    boxes = [W.strained(W.hcno_box(1500, seed=61), 0.1, seed=61),   # the larger box first
             W.strained(W.hcno_box(300, seed=62), 0.1, seed=62)]
    for k, box in enumerate(boxes):
        print(f"[case] noalloc call {k}: {box.n_atoms} atoms", flush=True)
        sys.stdout.flush()
        calc = FourierD3ASECalculator(device=torch.device("cuda:0"))
        calc._build_model(SimpleNamespace(numbers=box.numbers, cell=box.cell, pbc=np.array([True] * 3)))
        ref = _eval(calc, box, noalloc=False)
        got = _eval(calc, box, noalloc=True)
        de = abs(got["energy"][0] - ref["energy"][0]) * 27.21138505 / box.n_atoms
        df = np.abs(got["grad_pos"] - ref["grad_pos"]).max() * 27.21138505
        dv = np.abs(got["virial"] - ref["virial"]).max() * 27.21138505
        print(f"[errors] noalloc call {k}: E {de:.3e} eV/atom  F {df:.3e} eV/A  virial {dv:.3e} eV", flush=True)
        assert np.isfinite(got["grad_pos"]).all()
        # the spreading's float adds run in a varying order: agreement to float
        # rounding, far inside the stated 1e-6 eV/atom and 1e-5 eV/A
        assert de < 1e-9 and df < 1e-6 and dv < 1e-7 * max(1.0, np.abs(ref["virial"]).max() * 27.21138505)


def test_noalloc_rejects_a_box_beyond_its_scratch():
    """Host-only check (runs without a GPU): a box whose workspace exceeds the
    static scratch (the 1M-atom 512^3 mesh) is refused before any device work."""
    import ctypes as CT
    dummy = torch.zeros(64)
    m = L.Model()
    m.n_species, m.n_rank = 2, 3
    for f in ("rcov", "cnref", "vq", "eigs", "qq"):
        setattr(m, f, dummy.data_ptr())
    m.s6, m.s8, m.a1, m.a2 = 1.0, 0.7875, 0.4289, 4.4407
    m.r_cut = 6.0
    g = L.make_geom(np.eye(3) * 247.0, [512] * 3, 5, L.FD3_INTERP_P3M)
    need = CT.c_size_t()
    L.check(L.lib().fd3_eval_celllist_workspace(CT.byref(m), CT.byref(g), 1000, L.FD3_NEED_GRAD, CT.byref(need)))
    assert need.value > L.FD3_NOALLOC_SCRATCH_BYTES
    o = L.Outputs()
    rc = L.lib().fd3_eval_celllist_noalloc(CT.byref(m), CT.byref(g), 1000, None, L.FD3_DTYPE_F64, None,
                                           L.FD3_NEED_GRAD, CT.byref(o), None)
    assert rc == L.FD3_ERR_WORKSPACE
    assert "static scratch" in L.lib().fd3_last_error().decode()
