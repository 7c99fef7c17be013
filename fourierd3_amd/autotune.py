"""Tile autotuner for the mesh passes.

The library picks the tiles of its mesh kernels from fixed rules (lines per
block from the register budget, ranks per k-space tile from the spectra a
thread holds, k-space tiles per block from the block count).  This module
measures the alternatives on the GPU for given workloads and writes the
winners to ``fourierd3_amd/data/tuning.json``, which the package registers
with the library (``fd3_tune_set``) when it loads it.  Results do not depend
on the tiles; only the time does.

Each tile moves one group of kernels, so the search is coordinate-wise:

  tiles[0]  y lines per z-pass block   k_spread_r2c_z + k_c2r_z_gather
  tiles[1]  x lines per y-pass block   k_c2c_y_fwd + k_c2c_y_inv
  tiles[2]  ranks per k-space tile     k_kspace            (searched jointly
  tiles[3]  k-space tiles per block    k_kspace             with tiles[2];
                                                            any count up to 64
                                                            for the split x kernel)

A candidate the library would not take (``fd3_plan_tiles`` reports a
different tile) is skipped.  A key keeps its tuned tiles only if a full
evaluation is at least 1 % faster than with the library's own choice.

    python -m fourierd3_amd.autotune [--workloads headline,rank4,rank6,sio2,sio2_1m]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time
from types import SimpleNamespace

import numpy as np

from . import _lib as L

GROUPS = {
    0: ("k_spread_r2c_z", "k_c2r_z_gather"),
    1: ("k_c2c_y_fwd", "k_c2c_y_inv"),
    2: ("k_kspace",),
}


def plan_tiles(ns, n_species: int, n_rank: int) -> list[int]:
    t = (C.c_int32 * 4)()
    L.check(L.lib().fd3_plan_tiles((C.c_int32 * 3)(*ns), n_species, n_rank, t))
    return list(t)


def set_tiles(ns, n_species: int, n_rank: int, tiles) -> None:
    L.check(L.lib().fd3_tune_set((C.c_int32 * 3)(*ns), n_species, n_rank, (C.c_int32 * 4)(*tiles)))


def _workload(name: str):
    """(label, box, c6tol) of a named workload."""
    from . import workloads as W
    if name == "headline":
        return "H/C/N/O 100k atoms, c6tol 1 %", W.hcno_box(100_000, seed=0), 1.0
    if name == "rank4":
        return "H/C/N/O 100k atoms, c6tol 2 %", W.hcno_box(100_000, seed=0), 0.02
    if name == "rank6":
        return "H/C/N/O 100k atoms, c6tol 1e-4", W.hcno_box(100_000, seed=0), 1e-4
    if name == "sio2":
        return "SiO2 648 atoms", W.sio2_box(), 1.0
    if name == "sio2_1m":
        return "SiO2 999,999 atoms", W.sio2_box(999_999, seed=0), 1.0
    raise ValueError(f"unknown workload {name!r}")


class _Runner:
    """One box on the GPU: full evaluations, per-kernel times from the
    library's HIP-event profile."""

    def __init__(self, box, c6tol: float):
        import torch

        from .calculator import FourierD3ASECalculator
        dev = torch.device("cuda:0")
        self.torch = torch
        self.box = box
        self.calc = FourierD3ASECalculator(device=dev, c6tol=c6tol)
        self.calc._build_model(SimpleNamespace(numbers=box.numbers, cell=box.cell, pbc=np.array([True] * 3)))
        m = self.calc._model
        self.key = ([int(v) for v in m.ns_mesh], int(m.n_species), int(m.n_rank))
        self.pos = torch.tensor(box.positions, dtype=torch.float64, device=dev)
        f64 = dict(dtype=torch.float64, device=dev)
        self.out = {"energy": torch.empty(1, **f64), "grad_pos": torch.empty((box.n_atoms, 3), **f64),
                    "virial": torch.empty((3, 3), **f64), "grad_params": torch.empty(4, **f64)}

    def step(self, n: int = 1) -> None:
        for _ in range(n):
            self.calc.evaluate(self.pos, self.box.cell, out=self.out)

    def kernel_ms(self, names, reps: int) -> float:
        self.step(1)
        self.torch.cuda.synchronize()
        L.profile_only(None)
        L.profile_enable(True)
        self.step(reps)
        rep = L.profile_report()
        L.profile_enable(False)
        return sum(rep[k][0] / rep[k][1] for k in names if k in rep)

    def step_ms(self, reps: int) -> float:
        self.step(2)
        self.torch.cuda.synchronize()
        best = float("inf")
        for _ in range(3):
            t0 = time.perf_counter()
            self.step(reps)
            self.torch.cuda.synchronize()
            best = min(best, (time.perf_counter() - t0) / reps * 1e3)
        return best


def _pow2_upto(n: int):
    v = 1
    while v <= n:
        yield v
        v <<= 1


def tune_box(box, c6tol: float, reps: int = 5, log=print) -> dict:
    """Coordinate search of one box's key; returns its table entry."""
    run = _Runner(box, c6tol)
    ns, nS, r = run.key
    L.check(L.lib().fd3_tune_clear())
    base = plan_tiles(ns, nS, r)
    run.step(2)
    default_ms = run.step_ms(reps)
    best = list(base)
    trials = []

    def measure(tiles, group):
        set_tiles(ns, nS, r, tiles)
        got = plan_tiles(ns, nS, r)
        if [g for g, t in zip(got, tiles) if t] != [t for t in tiles if t]:
            return None  # the library would not take this tile
        ms = run.kernel_ms(GROUPS[group], reps)
        trials.append({"tiles": got, "group": group, "ms": ms})
        return ms

    for g, idx in ((0, (0,)), (1, (1,)), (2, (2, 3))):
        cands = []
        if g == 0:
            cands = [[v, best[1], best[2], best[3]] for v in _pow2_upto(ns[1])]
        elif g == 1:
            cands = [[best[0], v, best[2], best[3]] for v in _pow2_upto(ns[0])]
        else:
            cands = [[best[0], best[1], rt, kt] for rt in range(1, min(r, 8) + 1) for kt in (1, 2, 3, 4, 5, 6, 8)]
        scored = []
        for c in cands:
            ms = measure(c, g)
            if ms is not None:
                scored.append((ms, c))
        if scored:
            ms, c = min(scored, key=lambda x: x[0])
            for i in idx:
                best[i] = c[i]
            log(f"  group {g}: " + " ".join(f"{[x[1][i] for i in idx]}={x[0]:.4f}" for x in scored)
                + f" -> best {[c[i] for i in idx]}")
    set_tiles(ns, nS, r, best)
    tuned_ms = run.step_ms(reps)
    keep = best != base and tuned_ms < 0.99 * default_ms
    L.check(L.lib().fd3_tune_clear())
    entry = {"ns": ns, "n_species": nS, "n_rank": r, "tiles": best if keep else base, "default_tiles": base,
             "ms_default": default_ms, "ms_tuned": tuned_ms, "kept": keep, "trials": trials}
    del run
    return entry


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--workloads", default="headline,rank4,rank6,sio2,sio2_1m")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=str(L.TUNING_PATH))
    ap.add_argument("--dry-run", action="store_true", help="print the table instead of writing it")
    args = ap.parse_args(argv)
    import torch
    if not torch.cuda.is_available():
        print("autotune needs a GPU", file=sys.stderr)
        return 2
    entries = []
    for name in args.workloads.split(","):
        label, box, tol = _workload(name)
        print(f"[autotune] {name}: {label}", flush=True)
        e = tune_box(box, tol, args.reps, log=lambda s: print(s, flush=True))
        e["workload"] = label
        print(f"[autotune] {name}: key {e['ns']} nS {e['n_species']} r {e['n_rank']}: default {e['default_tiles']} "
              f"{e['ms_default']:.4f} ms, tuned {e['tiles']} {e['ms_tuned']:.4f} ms, kept {e['kept']}", flush=True)
        entries.append(e)
        torch.cuda.empty_cache()
    table = {"device": torch.cuda.get_device_name(0), "tiles": ["z lines", "y lines", "ranks per k tile",
                                                                  "k tiles per block"],
             "entries": [{k: v for k, v in e.items() if k != "trials"} for e in entries if e["kept"]],
             "measured": entries}
    text = json.dumps(table, indent=1)
    if args.dry_run:
        print(text)
    else:
        with open(args.out, "w") as fh:
            fh.write(text + "\n")
        print(f"[autotune] wrote {args.out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
