"""Real-space lattice sum of the BJ-damped D3 energy (oracle; test infrastructure only).

The k-space energy of ``src/fourierd3/core.py:553-581`` is the Fourier
representation of
    E = -1/2 sum_{i,j,n}' C6_ij [ s6 / (r^6 + R0^6) + 3 s8 QQ_ij / (r^8 + R0^8) ]
with C6_ij = sum_r lambda_r c_ir c_jr (the low-rank C6 of the decomposition)
and R0_ij, QQ_ij of the pair's species (``src/fourierd3/pair_pot.py:13-21,64``);
the prime drops i == j at n = 0.  Summing it directly over lattice images up
to a radius R_c, plus the continuum tail beyond R_c, gives an independent
value that SPME converges to as the mesh is refined (and Ewald as k_cutoff
grows).  Used to pin the sign, the 1/(2 Omega) factor, the k = 0 term and the
self term of the reciprocal-space energy.
"""
from __future__ import annotations

import itertools
import math

import numpy as np
from scipy.integrate import quad


def direct_energy(pos_B: np.ndarray, cell_B: np.ndarray, spec: np.ndarray, c: np.ndarray,
                  eigs: np.ndarray, R0: np.ndarray, QQ: np.ndarray, params, r_max: float) -> float:
    """Direct lattice sum in Hartree (all lengths Bohr)."""
    s6, s8 = float(params[0]), float(params[1])
    pos = np.asarray(pos_B, dtype=np.float64)
    cell = np.asarray(cell_B, dtype=np.float64)
    c6 = (c * eigs) @ c.T                                          # (N, N)
    r0 = R0[spec][:, spec]
    qq = QQ[spec][:, spec]
    inv_t = np.linalg.inv(cell).T
    d = 1.0 / np.linalg.norm(inv_t, axis=1)
    nimg = [int(math.ceil(r_max / dd)) + 1 for dd in d]
    total = 0.0
    diff = pos[None, :, :] - pos[:, None, :]                        # r_j - r_i
    n = len(pos)
    for sh in itertools.product(*[range(-k, k + 1) for k in nimg]):
        v = diff + np.asarray(sh, dtype=np.float64) @ cell
        r2 = np.einsum("ijk,ijk->ij", v, v)
        mask = r2 < r_max * r_max
        if not any(sh):
            mask[np.arange(n), np.arange(n)] = False
        if not mask.any():
            continue
        r6 = r2 ** 3
        r8 = r6 * r2
        term = c6 * (s6 / (r6 + r0 ** 6) + 3 * s8 * qq / (r8 + r0 ** 8))
        total += float(np.sum(term[mask]))
    # continuum tail beyond r_max: each pair sees a uniform density 1/Omega
    omega = abs(np.linalg.det(cell))
    tail = 0.0
    for a in range(n):
        for b in range(n):
            f = lambda r, a=a, b=b: 4 * math.pi * r * r * c6[a, b] * (
                s6 / (r ** 6 + r0[a, b] ** 6) + 3 * s8 * qq[a, b] / (r ** 8 + r0[a, b] ** 8))
            tail += quad(f, r_max, np.inf, epsabs=0, epsrel=1e-12, limit=200)[0] / omega
    return -0.5 * (total + tail)
