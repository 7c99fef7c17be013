"""``FourierD3ASECalculator``: energy, forces and stress of a periodic box.

Mirrors ``src/fourierd3/calculator.py:10-192`` (constructor arguments,
``implemented_properties``, lazy ``_build_model(atoms)``, ``results`` in eV,
eV/Angstrom and a 3x3 stress in eV/Angstrom^3).  Where the reference builds a
matscipy neighbour list on the CPU every call and differentiates through a
strain tensor with autograd, this calculator makes one call into the HIP
library (``fd3_eval_celllist``): the neighbour pairs are enumerated on the GPU
from a cell list, and forces and the strain derivative come from the adjoint
kernels, so
    forces = -dE/dr,  stress = (dE/d strain) / volume
exactly as ``src/fourierd3/calculator.py:180-188`` defines them.

ASE is used as the base class when it is importable; otherwise a minimal
stand-in with the same ``calculate``/``get_*`` protocol is used, and any
object with ``numbers``, ``positions``, ``cell`` and ``pbc`` works as atoms.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from . import tables as T
from .core import FourierD3, check_dtype

try:  # pragma: no cover - ASE is optional
    from ase.calculators.calculator import Calculator as _Base, all_changes
except Exception:  # ASE absent
    all_changes = ["positions", "numbers", "cell", "pbc"]

    class _Base:
        """Minimal stand-in for ``ase.calculators.calculator.Calculator``."""

        implemented_properties: list = []

        def __init__(self, **kwargs):
            self.results = {}
            self.atoms = None
            self.parameters = dict(kwargs)

        def calculate(self, atoms=None, properties=None, system_changes=all_changes):
            if atoms is not None:
                self.atoms = atoms
            self.results = {}

        def get_property(self, name, atoms=None):
            atoms = atoms if atoms is not None else self.atoms
            self.calculate(atoms, [name], all_changes)
            return self.results[name]

        def get_potential_energy(self, atoms=None, force_consistent=False):
            return self.get_property("energy", atoms)

        def get_forces(self, atoms=None):
            return self.get_property("forces", atoms)

        def get_stress(self, atoms=None):
            return self.get_property("stress", atoms)


def _cell_array(atoms) -> np.ndarray:
    cell = getattr(atoms, "cell")
    return np.asarray(getattr(cell, "array", cell), dtype=np.float64).reshape(3, 3)


class FourierD3ASECalculator(_Base):
    """Fourier-D3 dispersion energy, forces and stress (eV, eV/A, eV/A^3)."""

    implemented_properties = ["energy", "forces", "stress"]

    def __init__(self, r_cut=6.0, method="spme",
                 device=torch.device("cuda" if torch.cuda.is_available() else "cpu"), verbose=False,
                 params=None, c6tol: float = 1, xcfunc: str = "pbe", cnfunc="smooth_cut",
                 k_cutoff: float = 10.0, mesh_spacing: float = 1.2, interpolation_nodes: int = 5,
                 dtype=torch.float32, r_direct: float = 20.0, **kwargs):
        super().__init__(**kwargs)
        check_dtype(dtype)  # float32 kernels; dtype=torch.float64 raises NotImplementedError
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda" if torch.cuda.is_available() else "cpu")
        self.r_cut = float(r_cut)
        self.method = method
        self.verbose = verbose
        self.params = params
        self.angstrom_to_bohr = T.ANGSTROM_TO_BOHR
        self.HARTREE_TO_EV = T.HARTREE_TO_EV
        self.k_cutoff = k_cutoff
        self.xcfunc = xcfunc
        self.cnfunc = cnfunc
        self.c6tol = c6tol
        self.mesh_spacing = mesh_spacing
        self.interpolation_nodes = interpolation_nodes
        self.dtype = dtype
        self.r_direct = float(r_direct)  # method='direct' (extension): real-space image radius, Angstrom
        self._model = None
        self._ws = None
        self._numbers = None

    def _update_cell(self, cell):
        self._model._update_cell(cell=cell)

    def _build_graph(self, atoms, rcut=None):
        """Directed neighbour pairs within ``rcut`` (Angstrom, default r_cut)
        as ``edge_index`` (2, E) long and integer ``unit_shifts`` (E, 3) in
        this calculator's dtype, the matscipy ``neighbour_list("ijS")``
        semantics of ``src/fourierd3/calculator.py:94-123``, built on the GPU
        (``fd3_neighbours_count/fill``; pairs sorted by (i, j, S)).  The
        evaluation paths do not need it: they enumerate pairs from a cell list."""
        if rcut is None:
            rcut = self.r_cut
        if not bool(np.asarray(atoms.pbc).all()):
            raise AssertionError("particle-mesh only supports 3d pbc")
        dev = self.device
        pos = torch.tensor(np.asarray(atoms.positions, dtype=np.float64), device=dev)
        n = pos.shape[0]
        arr = L.cell_array(_cell_array(atoms))
        h = L.lib()
        nb = C.c_size_t()
        L.check(h.fd3_neighbours_workspace(n, arr, float(rcut), C.byref(nb)))
        ws = torch.empty(max(nb.value, 256), dtype=torch.uint8, device=dev)
        cnt = torch.zeros(1, dtype=torch.int64, device=dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        L.check(h.fd3_neighbours_count(n, L.ptr(pos), L.FD3_DTYPE_F64, arr, float(rcut), L.ptr(ws), ws.numel(),
                                       L.ptr(cnt), stream))
        e = int(cnt.item())
        src = torch.empty(max(e, 1), dtype=torch.int64, device=dev)
        dst = torch.empty_like(src)
        sh = torch.empty((max(e, 1), 3), dtype=torch.int32, device=dev)
        L.check(h.fd3_neighbours_fill(n, L.ptr(pos), L.FD3_DTYPE_F64, arr, float(rcut), L.ptr(ws), ws.numel(), e,
                                      L.ptr(src), L.ptr(dst), L.ptr(sh), stream))
        edge_index = torch.stack((src[:e], dst[:e]))
        return edge_index, sh[:e].to(self.dtype)

    def _build_model(self, atoms):
        """Instantiate the model for the species and cell of ``atoms``
        (``src/fourierd3/calculator.py:74-92``)."""
        if self.cnfunc not in ("smooth_cut", "d4"):
            raise ValueError(f"Unknown cnfunc '{self.cnfunc}'. Expected 'smooth_cut' or 'd4'.")
        self._model = FourierD3(
            species=np.asarray(atoms.numbers), cell=_cell_array(atoms),
            pbc=torch.as_tensor(np.asarray(atoms.pbc)), mesh_spacing=self.mesh_spacing, params=self.params,
            c6tol=self.c6tol, xcfunc=self.xcfunc, cnfunc=self.cnfunc, device=self.device, method=self.method,
            interpolation_nodes=self.interpolation_nodes, k_cutoff=self.k_cutoff, verbose=self.verbose,
            r_cut=self.r_cut, dtype=self.dtype, r_direct=self.r_direct)
        self._numbers = np.asarray(atoms.numbers).copy()

    # ------------------------------------------------------------------
    def evaluate(self, positions: torch.Tensor, cell_A: np.ndarray, need_grad: bool = True,
                 need_params: bool = False, out: dict | None = None) -> dict:
        """One E+F+stress evaluation of device-resident positions (N, 3) in
        Angstrom (float32/float64) for the host cell ``cell_A``.  Returns
        device tensors: energy (1,) Hartree, grad_pos (N, 3) Hartree/A, virial
        (3, 3) Hartree (dE/d strain) and, if asked, grad_params (4,).  No host
        synchronisation happens here; ``out`` may hold preallocated outputs."""
        model = self._model
        if self.cnfunc == "d4":
            cell_t = torch.as_tensor(np.asarray(cell_A, dtype=np.float64))
            return model._run_d4(positions, cell_t, need_grad or need_params, need_params)
        if self.method == "direct":
            return model._run_direct(positions, np.asarray(cell_A, dtype=np.float64), need_grad, need_params, out)
        n = positions.shape[0]
        g = model._geom(np.asarray(cell_A, dtype=np.float64))
        m = model._model_struct()
        flags = L.eval_flags(need_grad, need_params)
        nbytes = C.c_size_t()
        L.check(L.lib().fd3_eval_celllist_workspace(C.byref(m), C.byref(g), n, flags, C.byref(nbytes)))
        if self._ws is None or self._ws.numel() < nbytes.value:
            self._ws = torch.empty(nbytes.value, dtype=torch.uint8, device=self.device)
        f64 = dict(dtype=torch.float64, device=self.device)
        if out is None:
            out = {"energy": torch.empty(1, **f64), "grad_pos": torch.empty((n, 3), **f64),
                   "virial": torch.empty((3, 3), **f64), "grad_params": torch.empty(4, **f64)}
        o = L.Outputs()
        o.energy = L.ptr(out["energy"])
        if need_grad:
            o.grad_pos = L.ptr(out["grad_pos"])
            o.virial = L.ptr(out["virial"])
        if need_params:
            o.grad_params = L.ptr(out["grad_params"])
        stream = torch.cuda.current_stream(self.device).cuda_stream
        L.check(L.lib().fd3_eval_celllist(
            C.byref(m), C.byref(g), n, L.ptr(positions), L.dtype_code(positions),
            L.ptr(model._species_i32), flags, L.ptr(self._ws), self._ws.numel(), C.byref(o), stream))
        return out

    def evaluate_npt(self, positions: torch.Tensor, cell: torch.Tensor, ref_cell_A: np.ndarray,
                     reach: float = 1.1, need_grad: bool = True, need_params: bool = False,
                     out: dict | None = None) -> dict:
        """``evaluate`` for NPT steps (``fd3_eval_celllist_npt``, or
        ``fd3_eval_d4_npt`` with cnfunc='d4'), after the
        reference's ``_update_cell`` (src/fourierd3/core.py:244-267): the cell
        is the device tensor ``cell`` (3, 3) float64 Angstrom, read on the
        stream, so a barostat can change it in place between replays of a
        captured graph.  ``ref_cell_A`` (host) fixes the mesh, the cell list
        and the |k| tables, which cover cells whose plane spacings shrink, or
        largest |k| grows, by up to ``reach``.  ``out["status"]`` (int32, on
        the device) reads 0, or a bit mask when the cell left that range; the
        energy and virial are then NaN."""
        model = self._model
        d4 = self.cnfunc == "d4"
        if cell.dtype != torch.float64 or cell.shape != (3, 3) or cell.device != positions.device:
            raise ValueError("cell must be a (3, 3) float64 tensor on the positions' device")
        cell = cell.contiguous()
        n = positions.shape[0]
        g = model._geom(np.asarray(ref_cell_A, dtype=np.float64))
        m = model._model_struct()
        flags = L.eval_flags(need_grad, need_params)
        nbytes = C.c_size_t()
        if d4:
            d4s = model._d4_struct()
            L.check(L.lib().fd3_eval_d4_npt_workspace(C.byref(m), C.byref(d4s), C.byref(g), reach, n, flags,
                                                      C.byref(nbytes)))
        else:
            L.check(L.lib().fd3_eval_celllist_npt_workspace(C.byref(m), C.byref(g), reach, n, flags,
                                                            C.byref(nbytes)))
        if self._ws is None or self._ws.numel() < nbytes.value:
            self._ws = torch.empty(nbytes.value, dtype=torch.uint8, device=self.device)
        f64 = dict(dtype=torch.float64, device=self.device)
        if out is None:
            out = {"energy": torch.empty(1, **f64), "grad_pos": torch.empty((n, 3), **f64),
                   "virial": torch.empty((3, 3), **f64), "grad_params": torch.empty(4, **f64)}
        if "status" not in out:
            out["status"] = torch.zeros(1, dtype=torch.int32, device=self.device)
        o = L.Outputs()
        o.energy = L.ptr(out["energy"])
        if need_grad:
            o.grad_pos = L.ptr(out["grad_pos"])
            o.virial = L.ptr(out["virial"])
        if need_params:
            o.grad_params = L.ptr(out["grad_params"])
        stream = torch.cuda.current_stream(self.device).cuda_stream
        args = (reach, L.ptr(cell), L.ptr(out["status"]), n, L.ptr(positions), L.dtype_code(positions),
                L.ptr(model._species_i32), flags, L.ptr(self._ws), self._ws.numel(), C.byref(o), stream)
        if d4:
            L.check(L.lib().fd3_eval_d4_npt(C.byref(m), C.byref(d4s), C.byref(g), *args))
        else:
            L.check(L.lib().fd3_eval_celllist_npt(C.byref(m), C.byref(g), *args))
        return out

    def evaluate_batch(self, positions: torch.Tensor, species: torch.Tensor, offsets, cells,
                       need_grad: bool = True, need_params: bool = False) -> dict:
        """E+F+stress of many boxes of this model's species set, one box after
        another (``fd3_eval_celllist_batch``, the serial form; ``plan_batch``
        merges the boxes' launches instead).

        positions (N_total, 3) Angstrom and species (N_total,) int32 compact
        ids (positions in the model's ascending atomic-number order) are
        device tensors packed box after box; ``offsets`` (B + 1) are the atom
        offsets and ``cells`` the B host (3, 3) cells in Angstrom.  Each box
        gets its own mesh by the reference's rule for its cell.  Returns
        device tensors energy (B,), grad_pos (N_total, 3), virial (B, 3, 3)
        and grad_params (B, 4) in Hartree units, with no host sync."""
        model = self._model
        offs = np.asarray(offsets, dtype=np.int64)
        nb = len(offs) - 1
        geoms = (L.Geom * max(nb, 1))()
        a2b = self.angstrom_to_bohr
        for b in range(nb):
            cell = np.asarray(cells[b], dtype=np.float64).reshape(3, 3)
            ns = T.get_ns_mesh(cell * a2b, self.mesh_spacing * a2b)
            geoms[b] = L.make_geom(cell, ns, model.interpolation_nodes, model._interp)
        m = model._model_struct()
        flags = L.eval_flags(need_grad, need_params)
        offs_c = (C.c_int64 * len(offs))(*[int(v) for v in offs])
        nbytes = C.c_size_t()
        L.check(L.lib().fd3_eval_celllist_batch_workspace(C.byref(m), nb, geoms, offs_c, flags, C.byref(nbytes)))
        if self._ws is None or self._ws.numel() < nbytes.value:
            self._ws = torch.empty(nbytes.value, dtype=torch.uint8, device=self.device)
        f64 = dict(dtype=torch.float64, device=self.device)
        out = {"energy": torch.empty(nb, **f64), "grad_pos": torch.empty((int(offs[-1]), 3), **f64),
               "virial": torch.empty((nb, 3, 3), **f64), "grad_params": torch.empty((nb, 4), **f64)}
        outs = (L.Outputs * max(nb, 1))()
        for b in range(nb):
            o = outs[b]
            o.energy = L.ptr(out["energy"]) + 8 * b
            if need_grad:
                o.grad_pos = L.ptr(out["grad_pos"]) + 24 * int(offs[b])
                o.virial = L.ptr(out["virial"]) + 72 * b
            if need_params:
                o.grad_params = L.ptr(out["grad_params"]) + 32 * b
        stream = torch.cuda.current_stream(self.device).cuda_stream
        L.check(L.lib().fd3_eval_celllist_batch(
            C.byref(m), nb, geoms, offs_c, L.ptr(positions), L.dtype_code(positions), L.ptr(species), flags,
            L.ptr(self._ws), self._ws.numel(), outs, stream))
        return out

    def plan_batch(self, atoms_list, need_params: bool = False):
        """A ``BoxBatch`` of the boxes in ``atoms_list`` with this calculator's
        settings: any mix of species sets (one model and decomposition per
        set), each box on the mesh the reference's rule gives its own cell,
        evaluated with one merged launch sequence (``fd3_batch_*``)."""
        from .batch import BoxBatch
        for at in atoms_list:
            if not bool(np.asarray(at.pbc).all()):
                raise AssertionError("particle-mesh only supports 3d pbc")
        return BoxBatch([np.asarray(at.numbers) for at in atoms_list], [_cell_array(at) for at in atoms_list],
                        device=self.device, need_params=need_params, mesh_spacing=self.mesh_spacing,
                        r_cut=self.r_cut, c6tol=self.c6tol, method=self.method, params=self.params,
                        xcfunc=self.xcfunc, interpolation_nodes=self.interpolation_nodes, dtype=self.dtype)

    def calculate_batch(self, atoms_list) -> list:
        """Energy (eV), forces (eV/A) and 3x3 stress (eV/A^3) of every ``atoms``
        in ``atoms_list`` in one batched evaluation (``plan_batch``); the
        result of box b equals ``calculate`` on a fresh calculator for it."""
        if self.cnfunc != "smooth_cut":
            raise NotImplementedError("the batched form evaluates cnfunc='smooth_cut'")
        batch = self.plan_batch(atoms_list)
        pos = torch.tensor(np.concatenate([np.asarray(at.positions, dtype=np.float64) for at in atoms_list]),
                           device=self.device)
        batch.evaluate(pos)
        return batch.results()

    def calculate(self, atoms=None, properties=None, system_changes=all_changes):
        """Energy (eV), forces (eV/A) and 3x3 stress (eV/A^3) of ``atoms``
        (``src/fourierd3/calculator.py:125-192``)."""
        super().calculate(atoms, properties, system_changes)
        atoms = atoms if atoms is not None else self.atoms
        if self.cnfunc not in ("smooth_cut", "d4"):
            raise ValueError(f"Unknown cnfunc '{self.cnfunc}'. Expected 'smooth_cut' or 'd4'.")
        if self._model is None:
            self._build_model(atoms)
        cell = _cell_array(atoms)
        self._model._update_cell(torch.tensor(cell, dtype=torch.float64))
        pos = torch.tensor(np.asarray(atoms.positions, dtype=np.float64), device=self.device)
        if self.cnfunc == "d4":   # calculator.py:169-171: positions only, CN in k-space
            res = self._model._run_d4(pos, torch.tensor(cell, dtype=torch.float64), True, False)
        else:
            res = self.evaluate(pos, cell)
        e_ha = res["energy"].item()
        volume_b = abs(np.linalg.det(cell)) * self.angstrom_to_bohr ** 3
        self.results["energy"] = e_ha * self.HARTREE_TO_EV
        self.results["forces"] = (-res["grad_pos"] * self.HARTREE_TO_EV).cpu().numpy()
        self.results["stress"] = (res["virial"] * self.HARTREE_TO_EV / volume_b
                                  * self.angstrom_to_bohr ** 3).cpu().numpy()
