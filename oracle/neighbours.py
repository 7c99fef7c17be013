"""Brute-force periodic neighbour list (oracle; test infrastructure only).

Restates the semantics FourierD3 takes from matscipy ``neighbour_list("ijS",
pbc, cell, positions, cutoff)`` (called at ``src/fourierd3/calculator.py:103-109``):
every directed pair (i, j, S) with |r_j + S.cell - r_i| < cutoff, S integer,
excluding i == j with S == 0; both directions are listed, and in boxes thinner
than the cutoff an atom also pairs with its own images.  Positions need not be
inside the cell.

The search wraps fractional coordinates into [0, 1) and tries every image
shift S_w in [-n_a, n_a]^3 with n_a = ceil(cutoff / d_a), d_a the distance
between the lattice planes spanned by the other two vectors; that range holds
every image within the cutoff.  Shifts are reported for the unwrapped
positions: S = S_w + floor(f_i) - floor(f_j).
"""
from __future__ import annotations

import itertools
import math

import numpy as np


def plane_distances(cell: np.ndarray) -> np.ndarray:
    cell = np.asarray(cell, dtype=np.float64)
    inv_t = np.linalg.inv(cell).T          # rows: reciprocal vectors / 2pi
    return 1.0 / np.linalg.norm(inv_t, axis=1)


def neighbour_list_kdtree(positions: np.ndarray, cell: np.ndarray, cutoff: float):
    """Same pairs as ``neighbour_list`` for orthorhombic cells with every edge
    longer than 2 * cutoff (then each pair has at most one image in range),
    using scipy's periodic k-d tree; used to time the CPU baseline."""
    from scipy.spatial import cKDTree

    cell = np.asarray(cell, dtype=np.float64)
    lengths = np.diag(cell)
    if np.abs(cell - np.diag(lengths)).max() > 0 or (lengths <= 2 * cutoff).any():
        return neighbour_list(positions, cell, cutoff)
    pos = np.asarray(positions, dtype=np.float64)
    wrapped = np.mod(pos, lengths)
    wrapped = np.where(wrapped >= lengths, 0.0, wrapped)
    pairs = cKDTree(wrapped, boxsize=lengths).query_pairs(cutoff, output_type="ndarray")
    i = np.concatenate([pairs[:, 0], pairs[:, 1]])
    j = np.concatenate([pairs[:, 1], pairs[:, 0]])
    d = pos[j] - pos[i]
    S = -np.round(d / lengths).astype(np.int64)
    order = np.lexsort((S[:, 2], S[:, 1], S[:, 0], j, i))
    return i[order].astype(np.int64), j[order].astype(np.int64), np.ascontiguousarray(S[order])


def neighbour_list(positions: np.ndarray, cell: np.ndarray, cutoff: float, chunk: int = 512):
    """Return (i, j, S) as int64 arrays of shape (E,), (E,), (E, 3), sorted by (i, j, S)."""
    pos = np.asarray(positions, dtype=np.float64)
    cell = np.asarray(cell, dtype=np.float64)
    n = pos.shape[0]
    frac = pos @ np.linalg.inv(cell)
    fl = np.floor(frac)
    pw = (frac - fl) @ cell
    nimg = [int(math.ceil(cutoff / d)) for d in plane_distances(cell)]
    shifts = np.array(list(itertools.product(*[range(-k, k + 1) for k in nimg])), dtype=np.int64)
    shift_vec = shifts.astype(np.float64) @ cell
    cut2 = cutoff * cutoff
    out_i, out_j, out_s = [], [], []
    for s_idx in range(len(shifts)):
        sv = shift_vec[s_idx]
        zero = not shifts[s_idx].any()
        for i0 in range(0, n, chunk):
            i1 = min(n, i0 + chunk)
            d = pw[None, :, :] + sv[None, None, :] - pw[i0:i1, None, :]      # (ci, n, 3)
            r2 = np.einsum("abk,abk->ab", d, d)
            mask = r2 < cut2
            if zero:
                ii = np.arange(i0, i1)
                mask[ii - i0, ii] = False
            a, b = np.nonzero(mask)
            out_i.append(a + i0)
            out_j.append(b)
            out_s.append(np.broadcast_to(shifts[s_idx], (len(a), 3)))
    i = np.concatenate(out_i) if out_i else np.zeros(0, np.int64)
    j = np.concatenate(out_j) if out_j else np.zeros(0, np.int64)
    s = np.concatenate(out_s) if out_s else np.zeros((0, 3), np.int64)
    s = s + (fl[i] - fl[j]).astype(np.int64)
    order = np.lexsort((s[:, 2], s[:, 1], s[:, 0], j, i))
    return i[order].astype(np.int64), j[order].astype(np.int64), np.ascontiguousarray(s[order])
