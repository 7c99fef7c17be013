"""CPU tests of the host side: the C ABI library loads and exports every symbol
include/fd3.h declares, status codes map to the reference's exception types,
and the module/calculator surface mirrors the reference's constructor checks.
No GPU kernel is launched here."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest
import torch

from fourierd3_amd import _lib as L
from fourierd3_amd import tables as T
from fourierd3_amd import workloads as W

HEADER = Path(__file__).resolve().parents[1] / "include" / "fd3.h"


def test_library_exports_every_declared_symbol():
    decl = set(re.findall(r"\b(fd3_\w+)\s*\(", HEADER.read_text()))
    assert decl, "no declarations parsed"
    assert decl == set(L.EXPORTS)
    h = L.lib()
    for name in decl:
        assert getattr(h, name) is not None
    assert h.fd3_version() >= 1


def test_status_codes_map_to_reference_exceptions():
    h = L.lib()
    g = L.make_geom(np.eye(3) * 10.0, [16, 12, 16], 5, L.FD3_INTERP_P3M)  # 12 is not a power of two
    nbytes = C.c_size_t()
    with pytest.raises(NotImplementedError, match="power of two"):
        L.check(h.fd3_fft_workspace(g.ns, 1, C.byref(nbytes)))
    m = L.Model()
    m.n_species = 0
    with pytest.raises(ValueError):
        L.check(h.fd3_eval_edges_workspace(C.byref(m), C.byref(g), 10, 10, 1, C.byref(nbytes)))
    sing = L.make_geom(np.zeros((3, 3)), [16, 16, 16], 5, 0)
    with pytest.raises(ValueError, match="singular"):
        L.check(h.fd3_spread(C.byref(sing), 1, 1, 0, None, 0, None, None, None, None))
    arr = L.cell_array(np.eye(3) * 5.0)
    with pytest.raises(ValueError, match="r_cut"):
        L.check(h.fd3_neighbours_workspace(10, arr, -1.0, C.byref(nbytes)))
    # NPT entry point: reach below 1, no device cell, Ewald (mesh methods only)
    m = L.Model()
    m.n_species, m.n_rank = 1, 1
    for f in ("rcov", "cnref", "vq", "eigs", "qq"):
        setattr(m, f, 8)  # never dereferenced: the checks fail first
    m.r_cut = 6.0
    g16 = L.make_geom(np.eye(3) * 20.0, [16] * 3, 5, L.FD3_INTERP_P3M)
    with pytest.raises(ValueError, match="reach"):
        L.check(h.fd3_eval_celllist_npt_workspace(C.byref(m), C.byref(g16), 0.9, 10, 1, C.byref(nbytes)))
    L.check(h.fd3_eval_celllist_npt_workspace(C.byref(m), C.byref(g16), 1.1, 10, 1, C.byref(nbytes)))
    with pytest.raises(ValueError, match="device cell"):
        L.check(h.fd3_eval_celllist_npt(C.byref(m), C.byref(g16), 1.1, None, None, 10, None, 1, None, 1,
                                        None, 0, None, None))
    gew = L.make_geom(np.eye(3) * 20.0, [4] * 3, 5, L.FD3_INTERP_EWALD)
    with pytest.raises(NotImplementedError, match="mesh methods"):
        L.check(h.fd3_eval_celllist_npt(C.byref(m), C.byref(gew), 1.1, 8, None, 10, None, 1, None, 1,
                                        None, 0, None, None))
    with pytest.raises(ValueError, match="reach"):
        L.check(h.fd3_eval_edges_npt_workspace(C.byref(m), C.byref(g16), 0.5, 10, 20, 1, C.byref(nbytes)))
    L.check(h.fd3_eval_edges_npt_workspace(C.byref(m), C.byref(g16), 1.1, 10, 20, 1, C.byref(nbytes)))
    with pytest.raises(ValueError, match="device cell"):
        L.check(h.fd3_eval_edges_npt(C.byref(m), C.byref(g16), 1.1, None, None, 10, None, 1, None, 20, None,
                                     None, None, 1, 1, None, 0, None, None))
    with pytest.raises(NotImplementedError, match="mesh methods"):
        L.check(h.fd3_eval_edges_npt(C.byref(m), C.byref(gew), 1.1, 8, None, 10, None, 1, None, 20, None,
                                     None, None, 1, 1, None, 0, None, None))
    d4 = L.CnD4()
    with pytest.raises(ValueError, match="reach"):
        L.check(h.fd3_eval_d4_npt_workspace(C.byref(m), C.byref(d4), C.byref(g16), 0.5, 10, 1, C.byref(nbytes)))
    with pytest.raises(NotImplementedError, match="mesh methods"):
        L.check(h.fd3_eval_d4_npt(C.byref(m), C.byref(d4), C.byref(gew), 1.1, 8, None, 10, None, 1, None, 1,
                                  None, 0, None, None))


def test_workspace_sizes_scale_with_mesh_and_atoms():
    h = L.lib()
    dummy = torch.zeros(4)
    m = L.Model()
    m.n_species, m.n_rank = 4, 2
    for f in ("rcov", "cnref", "vq", "eigs", "qq"):
        setattr(m, f, dummy.data_ptr())
    m.s6, m.s8, m.a1, m.a2 = T.PBE_PARAMS
    m.r_cut = 6.0
    sizes = []
    for n, ns in [(100, 16), (100, 32), (1000, 32)]:
        g = L.make_geom(np.eye(3) * 20.0, [ns] * 3, 5, 0)
        nb = C.c_size_t()
        L.check(h.fd3_eval_celllist_workspace(C.byref(m), C.byref(g), n, 1, C.byref(nb)))
        sizes.append(nb.value)
    assert sizes[0] < sizes[1] < sizes[2]
    # two complex meshes of C = 8 channels dominate: 2 * 8 * ns^2 (ns/2+1) * 8 bytes
    assert sizes[1] > 2 * 8 * 32 * 32 * 17 * 8


def test_module_constructor_checks():
    from fourierd3_amd import FourierD3
    box = W.sio2_box(30, seed=1)
    with pytest.raises(ValueError):
        FourierD3(box.numbers, box.cell, device="cpu", method="p3m", verbose=False)
    with pytest.raises(ValueError):
        FourierD3(box.numbers, box.cell, device="cpu", cnfunc="d3", verbose=False)
    md4 = FourierD3(box.numbers, box.cell, device="cpu", cnfunc="d4", verbose=False)
    d4 = T.cn_d4_tables(md4.species_unique)
    # f_XX(0) = delta_XX / 2 erfc(-k0) ~ delta_XX (pair_pot.py:324-327); r0 symmetric, in Bohr
    assert np.allclose(d4["selfcont"], np.diagonal(d4["delta"]), rtol=1e-12)
    assert np.allclose(d4["r0"], d4["r0"].T) and float(md4.cn_selfcont[0]) == pytest.approx(d4["selfcont"][0])
    # Ewald: the k grid is ceil(k_cutoff |a_i| / 2 pi) with the cell in Bohr (core.py:228-230)
    me = FourierD3(box.numbers, box.cell, device="cpu", method="ewald", k_cutoff=4.0, verbose=False)
    la = np.linalg.norm(box.cell * T.ANGSTROM_TO_BOHR, axis=1)
    assert list(me.ns_mesh) == [int(np.ceil(4.0 * v / (2 * np.pi))) for v in la]
    with pytest.raises(NotImplementedError):
        FourierD3(box.numbers, box.cell, device="cpu", xcfunc="b3lyp", verbose=False)
    with pytest.raises(AssertionError):
        FourierD3(box.numbers, box.cell, device="cpu", pbc=torch.tensor([True, True, False]), verbose=False)
    m = FourierD3(box.numbers, box.cell, device="cpu", verbose=False)
    assert list(m.ns_mesh) == list(T.get_ns_mesh(box.cell * T.ANGSTROM_TO_BOHR, 1.2 * T.ANGSTROM_TO_BOHR))
    assert m.n_species == 2 and m.v_q_reshaped.shape == (2, 7, m.n_rank)
    vol = abs(np.linalg.det(box.cell)) * T.ANGSTROM_TO_BOHR ** 3
    assert float(m.volume) == pytest.approx(vol, rel=1e-6)
    with pytest.raises(ValueError):
        m._update_cell(torch.eye(2))


def test_workload_generators():
    b = W.hcno_box(2000, seed=3)
    counts = {int(z): int(c) for z, c in zip(*np.unique(b.numbers, return_counts=True))}
    assert counts == {1: 1000, 6: 500, 7: 150, 8: 350}
    L_A = b.cell[0, 0]
    d = b.positions[:, None, :] - b.positions[None, :, :]
    d -= L_A * np.round(d / L_A)
    r = np.sqrt((d ** 2).sum(-1)) + np.eye(len(d)) * 10
    assert r.min() >= 1.0
    t = W.strained(b, 0.1, seed=1)
    f0 = b.positions @ np.linalg.inv(b.cell)
    f1 = t.positions @ np.linalg.inv(t.cell)
    assert np.allclose(f0, f1)


def test_float64_dtype_is_refused_not_silently_float32():
    """The kernels compute in float32 with float64 reductions; the reference's
    all-float64 mode (dtype=torch.float64, src/fourierd3/core.py:60) is not
    built and must raise rather than return float32-level answers."""
    from fourierd3_amd import FourierD3, FourierD3ASECalculator
    box = W.sio2_box(30, seed=1)
    with pytest.raises(NotImplementedError, match="float64"):
        FourierD3(box.numbers, box.cell, device="cpu", dtype=torch.float64, verbose=False)
    with pytest.raises(NotImplementedError, match="float64"):
        FourierD3ASECalculator(device="cpu", dtype=torch.float64)
    with pytest.raises(ValueError):
        FourierD3(box.numbers, box.cell, device="cpu", dtype=torch.float16, verbose=False)


def test_backward_is_first_order_only():
    """The custom backward returns gradients the adjoint kernels computed, so
    it has no graph: a second derivative (create_graph=True, e.g. a force
    loss) must raise instead of silently dropping the Hessian terms.  The HIP
    call is replaced by fixed gradients, so no GPU is needed."""
    from fourierd3_amd import FourierD3
    box = W.sio2_box(30, seed=1)
    n = box.n_atoms
    for cnfunc in ("smooth_cut", "d4"):
        m = FourierD3(box.numbers, box.cell, device="cpu", verbose=False, cnfunc=cnfunc)
        f64 = dict(dtype=torch.float64)
        fake = {"energy": torch.tensor([-1.5], **f64), "grad_pos": torch.full((n, 3), 0.25, **f64),
                "grad_shifts": torch.zeros((4, 3), **f64), "grad_cell": torch.zeros((3, 3), **f64),
                "virial": torch.zeros((3, 3), **f64), "grad_params": torch.zeros(4, **f64)}
        m._run_edges = lambda *a, **k: fake
        m._run_d4 = lambda *a, **k: fake
        pos = torch.tensor(box.positions, requires_grad=True)
        if cnfunc == "d4":
            e = m(pos)
        else:
            e = m(pos, torch.zeros((2, 4), dtype=torch.long), torch.zeros((4, 3), dtype=torch.float64))
        (g,) = torch.autograd.grad(e, pos)                      # first order: the kernels' gradient
        assert torch.allclose(g, fake["grad_pos"])
        e2 = m(pos) if cnfunc == "d4" else m(pos, torch.zeros((2, 4), dtype=torch.long),
                                              torch.zeros((4, 3), dtype=torch.float64))
        with pytest.raises(RuntimeError, match="second derivatives"):
            torch.autograd.grad(e2, pos, create_graph=True)


def test_eval_flags_follow_torch_deterministic_mode():
    """FD3_DETERMINISTIC is added exactly while torch's deterministic mode is
    on, and the header and the binding agree on the flag values."""
    import re

    import torch

    text = HEADER.read_text()
    for name in ("FD3_NEED_GRAD", "FD3_NEED_PARAMS", "FD3_DETERMINISTIC"):
        m = re.search(rf"#define {name} (\d+)", text)
        assert m and int(m.group(1)) == getattr(L, name), name
    prev = torch.are_deterministic_algorithms_enabled()
    try:
        torch.use_deterministic_algorithms(False)
        assert L.eval_flags(True, True) == L.FD3_NEED_GRAD | L.FD3_NEED_PARAMS
        assert L.eval_flags(False) == 0
        torch.use_deterministic_algorithms(True)
        assert L.eval_flags(True, False) == L.FD3_NEED_GRAD | L.FD3_DETERMINISTIC
        assert L.eval_flags(False, True) == L.FD3_NEED_PARAMS | L.FD3_DETERMINISTIC
    finally:
        torch.use_deterministic_algorithms(prev)
