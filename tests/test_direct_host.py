"""Argument checks of the real-space entry point (fd3_eval_direct) and the
module options around it; host only (the checks fail before any device call)."""
import ctypes as C

import numpy as np
import pytest

from fourierd3_amd import _lib as L


def _model():
    m = L.Model()
    m.n_species, m.n_rank = 2, 3
    for f in ("rcov", "cnref", "vq", "eigs", "qq"):
        setattr(m, f, 8)  # never dereferenced by the workspace query
    m.r_cut = 6.0
    return m


def test_direct_workspace_and_argument_checks():
    h = L.lib()
    m = _model()
    nb = C.c_size_t()
    cell = L.cell_array(np.eye(3) * 10.0)
    L.check(h.fd3_eval_direct_workspace(C.byref(m), cell, 20.0, 100, 1, C.byref(nb)))
    small = nb.value
    L.check(h.fd3_eval_direct_workspace(C.byref(m), cell, 20.0, 10_000, 1, C.byref(nb)))
    assert nb.value > small
    for bad in (0.0, -1.0, float("nan"), float("inf")):
        with pytest.raises(ValueError, match="r_max"):
            L.check(h.fd3_eval_direct_workspace(C.byref(m), cell, bad, 100, 1, C.byref(nb)))
    # a tiny cell with a large radius needs too many images
    with pytest.raises(NotImplementedError, match="lattice images"):
        L.check(h.fd3_eval_direct_workspace(C.byref(m), L.cell_array(np.eye(3) * 2.0), 200.0, 10, 1,
                                            C.byref(nb)))
    with pytest.raises(ValueError, match="singular"):
        L.check(h.fd3_eval_direct_workspace(C.byref(m), L.cell_array(np.zeros((3, 3))), 20.0, 10, 1,
                                            C.byref(nb)))
    with pytest.raises(ValueError, match="negative"):
        L.check(h.fd3_eval_direct_workspace(C.byref(m), cell, 20.0, -1, 1, C.byref(nb)))
    # evaluation without a workspace
    with pytest.raises(RuntimeError, match="workspace"):
        L.check(h.fd3_eval_direct(C.byref(m), cell, 20.0, 10, None, 1, None, 1, None, 0, None, None))


def test_direct_method_options():
    import torch

    from fourierd3_amd import FourierD3
    cell = np.eye(3) * 10.0
    with pytest.raises(NotImplementedError):
        FourierD3([1, 8], cell, method="direct", cnfunc="d4", device="cpu", verbose=False)
    with pytest.raises(NotImplementedError):
        FourierD3([1, 8], cell, method="direct", npt_reach=1.1, device="cpu", verbose=False)
    m = FourierD3([1, 8], cell, method="direct", r_direct=15.0, device="cpu", verbose=False)
    assert m.r_direct == 15.0
    with pytest.raises(ValueError, match="unknown method"):
        FourierD3([1, 8], cell, method="directsum", device="cpu", verbose=False)
    del torch
