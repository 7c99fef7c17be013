"""Seeded synthetic periodic boxes for the benchmark and test configurations.

No structures or datasets ship with this package; every box is generated from
a seed:

* ``sio2_box``   amorphous SiO2-like cubic box (Si:O = 1:2) at a mass density,
  random sequential placement with a minimum separation.
* ``hcno_box``   H/C/N/O in ratio 2 : 1 : 0.3 : 0.7 at a mass density, likewise.
* ``strained``   applies a random triclinic strain cell <- cell (I + U(-s, s)).
* ``mixed_z_box`` elements drawn uniformly from Z = 1..zmax at a number density.
* ``batch_boxes`` many boxes with log-uniform atom counts, mixed compositions.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_AMU = {1: 1.008, 6: 12.011, 7: 14.007, 8: 15.999, 14: 28.085}
_NA = 6.02214076e23


@dataclass
class Box:
    numbers: np.ndarray      # (N,) atomic numbers
    positions: np.ndarray    # (N, 3) Angstrom
    cell: np.ndarray         # (3, 3) Angstrom, rows are lattice vectors

    @property
    def n_atoms(self) -> int:
        return int(self.numbers.shape[0])


def _rsa_cubic(n: int, length: float, dmin: float, rng: np.random.Generator, batch: int = 4096) -> np.ndarray:
    """Random sequential addition of n points in a cubic periodic box with a
    minimum (minimum-image) separation dmin, using a grid of cells of edge
    >= dmin / sqrt(3) that each hold at most one point."""
    nc = max(1, int(math.floor(length / (dmin / math.sqrt(3.0)))))
    cs = length / nc
    reach = int(math.ceil(dmin / cs))
    grid = -np.ones((nc, nc, nc), dtype=np.int64)
    pts = np.empty((n, 3))
    count = 0
    offs = np.array([(a, b, c) for a in range(-reach, reach + 1) for b in range(-reach, reach + 1)
                     for c in range(-reach, reach + 1)], dtype=np.int64)
    tries = 0
    while count < n:
        tries += 1
        if tries > 100000:
            raise RuntimeError("random sequential addition did not converge; density too high")
        bsz = min(batch, 2 * (n - count) + 16)
        cand = rng.random((bsz, 3)) * length
        cell_idx = np.minimum((cand / cs).astype(np.int64), nc - 1)
        # conflicts with accepted points
        nb = (cell_idx[:, None, :] + offs[None, :, :]) % nc                  # (B, K, 3)
        occ = grid[nb[..., 0], nb[..., 1], nb[..., 2]]                         # (B, K)
        ok = np.ones(bsz, dtype=bool)
        has = occ >= 0
        if has.any():
            d = pts[np.where(has, occ, 0)] - cand[:, None, :]
            d -= length * np.round(d / length)
            close = (np.einsum("bkx,bkx->bk", d, d) < dmin * dmin) & has
            ok &= ~close.any(axis=1)
        # a cell can hold one point only
        ok &= grid[cell_idx[:, 0], cell_idx[:, 1], cell_idx[:, 2]] < 0
        idx = np.nonzero(ok)[0]
        # conflicts between survivors of the same batch: one point per cell,
        # and of each close pair only the earlier candidate is kept
        lin = (cell_idx[idx, 0] * nc + cell_idx[idx, 1]) * nc + cell_idx[idx, 2]
        _, first = np.unique(lin, return_index=True)
        idx = idx[np.sort(first)]
        if len(idx) > 1:
            from scipy.spatial import cKDTree
            pairs = cKDTree(np.mod(cand[idx], length), boxsize=length).query_pairs(dmin, output_type="ndarray")
            if len(pairs):
                drop = np.zeros(len(idx), dtype=bool)
                drop[np.maximum(pairs[:, 0], pairs[:, 1])] = True
                idx = idx[~drop]
        kept = list(idx[: n - count])
        for t in kept:
            pts[count] = cand[t]
            ci = cell_idx[t]
            grid[ci[0], ci[1], ci[2]] = count
            count += 1
    return pts


def _box_from_composition(counts: dict, density_gcc: float, dmin: float, seed: int) -> Box:
    rng = np.random.default_rng(seed)
    numbers = np.concatenate([np.full(c, z, dtype=np.int64) for z, c in counts.items()])
    mass = sum(_AMU[z] * c for z, c in counts.items())
    vol = mass / _NA / density_gcc * 1e24                  # Angstrom^3
    length = vol ** (1.0 / 3.0)
    pos = _rsa_cubic(len(numbers), length, dmin, rng)
    numbers = numbers[rng.permutation(len(numbers))]
    return Box(numbers=numbers, positions=pos, cell=np.eye(3) * length)


def sio2_box(n_atoms: int = 648, density_gcc: float = 2.2, dmin: float = 1.5, seed: int = 0) -> Box:
    n_si = n_atoms // 3
    return _box_from_composition({14: n_si, 8: n_atoms - n_si}, density_gcc, dmin, seed)


def hcno_box(n_atoms: int = 100_000, density_gcc: float = 1.0, dmin: float = 1.0, seed: int = 0) -> Box:
    # H : C : N : O = 2 : 1 : 0.3 : 0.7
    nh = int(round(n_atoms * 0.5))
    nc = int(round(n_atoms * 0.25))
    nn = int(round(n_atoms * 0.075))
    no = n_atoms - nh - nc - nn
    return _box_from_composition({1: nh, 6: nc, 7: nn, 8: no}, density_gcc, dmin, seed)


def strained(box: Box, amplitude: float = 0.1, seed: int = 0) -> Box:
    """Random triclinic strain: cell <- cell (I + U(-a, a)^{3x3}), positions
    mapped with the same linear map (fractional coordinates unchanged)."""
    rng = np.random.default_rng(seed)
    m = np.eye(3) + rng.uniform(-amplitude, amplitude, size=(3, 3))
    return Box(numbers=box.numbers.copy(), positions=box.positions @ m, cell=box.cell @ m)


def mixed_z_box(n_atoms: int = 10_000, zmax: int = 36, number_density: float = 0.0819,
                dmin: float = 1.0, seed: int = 0) -> Box:
    rng = np.random.default_rng(seed)
    length = (n_atoms / number_density) ** (1.0 / 3.0)
    pos = _rsa_cubic(n_atoms, length, dmin, rng)
    numbers = rng.integers(1, zmax + 1, size=n_atoms)
    return Box(numbers=numbers.astype(np.int64), positions=pos, cell=np.eye(3) * length)


def batch_boxes(n_boxes: int = 256, n_min: int = 50, n_max: int = 5000, amplitude: float = 0.1,
                seed: int = 0):
    """Boxes with log-uniform atom counts, half SiO2-like and half H/C/N/O,
    each with a random triclinic strain."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(n_boxes):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        if b % 2 == 0:
            box = sio2_box(max(3, n - n % 3), seed=seed * 100003 + b)
        else:
            box = hcno_box(max(8, n), seed=seed * 100003 + b)
        out.append(strained(box, amplitude, seed=seed * 100003 + b))
    return out
