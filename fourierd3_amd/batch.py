"""Many periodic boxes in one merged launch sequence (``fd3_batch_*``).

The reference evaluates one box per ``FourierD3ASECalculator.calculate`` call
(``src/fourierd3/calculator.py:125-192``) and its scaling scripts loop over
boxes on the host (``papercode/get_scaling.py:52-105``).  ``BoxBatch`` plans
a whole set of boxes once -- any number of species sets, each with its own
``FourierD3`` model and decomposition, every box with its own cell and the
mesh the reference's rule gives that cell -- and then evaluates energy,
forces and stress of all of them with one launch per pipeline stage
(``fd3_batch_run``), positions and results staying on the device.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np
import torch

from . import _lib as L
from . import tables as T
from .core import FourierD3


class BoxBatch:
    """A planned batch of boxes.

    numbers_list: per box, the atomic numbers (N_b,); cells: per box the
    (3, 3) cell in Angstrom (rows = lattice vectors).  The remaining keyword
    arguments are those of ``FourierD3`` (r_cut, c6tol, method ('spme' or
    'pme'), mesh_spacing, interpolation_nodes, params, xcfunc, cncorr).
    Positions passed to ``evaluate`` are packed box after box, in the order
    of ``numbers_list``.
    """

    def __init__(self, numbers_list: Sequence, cells: Sequence, device=None, need_grad: bool = True,
                 need_params: bool = False, mesh_spacing: float = 1.2, **model_kwargs):
        self.device = torch.device(device if device is not None else "cuda")
        method = model_kwargs.get("method", "spme")
        if method not in ("spme", "pme"):
            raise NotImplementedError("the batched form supports method='spme' and 'pme'")
        self.n_boxes = len(numbers_list)
        if self.n_boxes == 0 or len(cells) != self.n_boxes:
            raise ValueError("need one cell per box and at least one box")
        a2b = T.ANGSTROM_TO_BOHR
        # one model per species set (ascending atomic numbers, as FourierD3 compacts them)
        self.models: list = []
        key_of: dict = {}
        model_of_box = []
        species = []
        offs = [0]
        geoms = (L.Geom * self.n_boxes)()
        for b, (z, cell) in enumerate(zip(numbers_list, cells)):
            z = np.asarray(z, dtype=np.int64)
            cell = np.asarray(cell, dtype=np.float64).reshape(3, 3)
            key = tuple(int(v) for v in np.unique(z))
            if key not in key_of:
                key_of[key] = len(self.models)
                self.models.append(FourierD3(z, cell, device=self.device, verbose=False, mesh_spacing=mesh_spacing,
                                             **model_kwargs))
            mi = key_of[key]
            model = self.models[mi]
            model_of_box.append(mi)
            smap = {int(s): i for i, s in enumerate(model.species_unique)}
            species.append(np.array([smap[int(v)] for v in z], dtype=np.int32))
            ns = T.get_ns_mesh(cell * a2b, mesh_spacing * a2b)     # the reference's rule, per box
            geoms[b] = L.make_geom(cell, ns, model.interpolation_nodes, model._interp)
            offs.append(offs[-1] + len(z))
        self.geoms = geoms
        self.offsets = np.asarray(offs, dtype=np.int64)
        self.n_atoms = int(self.offsets[-1])
        self.cells = [np.asarray(c, dtype=np.float64).reshape(3, 3) for c in cells]
        self.species = torch.tensor(np.concatenate(species), dtype=torch.int32, device=self.device)
        self.model_of_box = np.asarray(model_of_box, dtype=np.int32)
        self._mstructs = (L.Model * len(self.models))(*[m._model_struct() for m in self.models])
        self.flags = L.eval_flags(need_grad, need_params)
        self._plan = C.c_void_p()
        mob = (C.c_int32 * self.n_boxes)(*[int(v) for v in model_of_box])
        offs_c = (C.c_int64 * (self.n_boxes + 1))(*[int(v) for v in self.offsets])
        L.check(L.lib().fd3_batch_create(len(self.models), self._mstructs, self.n_boxes, mob, geoms, offs_c,
                                         self.flags, C.byref(self._plan)))
        nbytes = C.c_size_t()
        L.check(L.lib().fd3_batch_workspace(self._plan, C.byref(nbytes)))
        self.workspace = torch.empty(max(nbytes.value, 256), dtype=torch.uint8, device=self.device)
        f64 = dict(dtype=torch.float64, device=self.device)
        self.out = {"energy": torch.empty(self.n_boxes, **f64), "grad_pos": torch.empty((self.n_atoms, 3), **f64),
                    "virial": torch.empty((self.n_boxes, 3, 3), **f64),
                    "grad_params": torch.empty((self.n_boxes, 4), **f64)}
        self._outs = (L.Outputs * self.n_boxes)()
        for b in range(self.n_boxes):
            o = self._outs[b]
            o.energy = L.ptr(self.out["energy"]) + 8 * b
            if need_grad:
                o.grad_pos = L.ptr(self.out["grad_pos"]) + 24 * int(self.offsets[b])
                o.virial = L.ptr(self.out["virial"]) + 72 * b
            if need_params:
                o.grad_params = L.ptr(self.out["grad_params"]) + 32 * b

    def info(self) -> dict:
        """Launches per run: merged, and what one call per box would make."""
        merged, per_box = C.c_int32(), C.c_int32()
        L.check(L.lib().fd3_batch_info(self._plan, C.byref(merged), C.byref(per_box)))
        return {"merged_launches": merged.value, "box_launches": per_box.value,
                "n_models": len(self.models), "workspace_bytes": self.workspace.numel()}

    def evaluate(self, positions: torch.Tensor) -> dict:
        """E+F+stress of every box: positions (N_total, 3) Angstrom on the
        device (float32/float64), packed box after box.  Returns the device
        tensors energy (B,) Hartree, grad_pos (N_total, 3) Hartree/A, virial
        (B, 3, 3) Hartree (dE/d strain) and grad_params (B, 4); no host sync."""
        if positions.device != self.device or positions.shape != (self.n_atoms, 3):
            raise ValueError(f"positions must be ({self.n_atoms}, 3) on {self.device}")
        positions = positions.contiguous()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        L.check(L.lib().fd3_batch_run(self._plan, L.ptr(positions), L.dtype_code(positions), L.ptr(self.species),
                                      L.ptr(self.workspace), self.workspace.numel(), self._outs, stream))
        return self.out

    def results(self) -> list:
        """The last evaluation as the reference calculator reports it: per box
        energy (eV), forces (eV/A) and 3x3 stress (eV/A^3)."""
        ev = T.HARTREE_TO_EV
        e = (self.out["energy"] * ev).cpu().numpy()
        f = (-self.out["grad_pos"] * ev).cpu().numpy()
        v = (self.out["virial"] * ev).cpu().numpy()
        a2b3 = T.ANGSTROM_TO_BOHR ** 3
        res = []
        for b, cell in enumerate(self.cells):
            vol_b = abs(np.linalg.det(cell)) * a2b3
            o0, o1 = int(self.offsets[b]), int(self.offsets[b + 1])
            res.append({"energy": float(e[b]), "forces": f[o0:o1], "stress": v[b] / vol_b * a2b3})
        return res

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            try:
                L.lib().fd3_batch_destroy(plan)
            except Exception:
                pass
            self._plan = C.c_void_p()
