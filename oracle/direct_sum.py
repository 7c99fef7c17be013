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


def tail_integrals_gl(R, r_max: float, n_nodes: int = 64):
    """I6 = int_rmax^inf r^2/(r^6+R^6) dr and I8 = int_rmax^inf r^2/(r^8+R^8) dr
    (torch, differentiable in R) by n-point Gauss-Legendre after r = r_max/u:
    I6 = int_0^1 rmax^3 u^2/(rmax^6 + R^6 u^6) du, I8 likewise with u^4 and
    the 8th powers.  Nodes from ``numpy.polynomial.legendre.leggauss``."""
    import torch
    x, w = np.polynomial.legendre.leggauss(n_nodes)
    u = torch.tensor(0.5 * (x + 1.0), dtype=torch.float64)
    w = torch.tensor(0.5 * w, dtype=torch.float64)
    Rx = R.unsqueeze(-1)
    rm3 = r_max ** 3
    i6 = (w * rm3 * u ** 2 / (r_max ** 6 + Rx ** 6 * u ** 6)).sum(-1)
    i8 = (w * rm3 * u ** 4 / (r_max ** 8 + Rx ** 8 * u ** 8)).sum(-1)
    return i6, i8


def direct_energy_torch(pos_B, cell_B, spec, c, eigs, QQ, params, r_max: float):
    """The lattice sum of ``direct_energy`` in torch FP64, differentiable in
    positions, cell, charges and the parameters (s6, s8, a1, a2): R0 is rebuilt
    from ``params`` as in ``src/fourierd3/pair_pot.py:64``; the continuum tail
    uses ``tail_integrals_gl``.  Which images fall inside r_max is decided on
    detached values, so gradients are those of the fixed set of terms."""
    import torch
    s6, s8, a1, a2 = params[0], params[1], params[2], params[3]
    R0 = a1 * torch.sqrt(3 * QQ) + a2
    c6 = (c * eigs) @ c.T
    r0 = R0[spec][:, spec]
    qq = QQ[spec][:, spec]
    cell_np = cell_B.detach().numpy()
    d = 1.0 / np.linalg.norm(np.linalg.inv(cell_np).T, axis=1)
    nimg = [int(math.ceil(r_max / dd)) + 1 for dd in d]
    n = pos_B.shape[0]
    diff = pos_B.unsqueeze(0) - pos_B.unsqueeze(1)                   # r_j - r_i
    total = pos_B.new_zeros(())
    for sh in itertools.product(*[range(-k, k + 1) for k in nimg]):
        shv = torch.tensor(sh, dtype=torch.float64) @ cell_B
        with torch.no_grad():
            v0 = diff + shv
            mask = (v0 * v0).sum(-1) < r_max * r_max
            if not any(sh):
                mask[torch.arange(n), torch.arange(n)] = False
        ii, jj = torch.nonzero(mask, as_tuple=True)
        if len(ii) == 0:
            continue
        v = diff[ii, jj] + shv
        r2 = (v * v).sum(-1)
        r6 = r2 ** 3
        r8 = r6 * r2
        total = total + (c6[ii, jj] * (s6 / (r6 + r0[ii, jj] ** 6)
                                       + 3 * s8 * qq[ii, jj] / (r8 + r0[ii, jj] ** 8))).sum()
    i6, i8 = tail_integrals_gl(R0, r_max)
    omega = torch.abs(torch.linalg.det(cell_B))
    tail = (c6 * 4 * math.pi * (s6 * i6 + 3 * s8 * QQ * i8)[spec][:, spec]).sum() / omega
    return -0.5 * (total + tail)


def direct_calculate(o, positions_A, cell_A, r_max_A: float, params=None):
    """E (eV), forces (eV/A), stress (eV/A^3) and dE/d(s6, s8, a1, a2) (eV) of
    the real-space sum for an ``OracleFourierD3`` model ``o`` (smooth-cut CN
    from the brute-force neighbour list, charges from ``o.charges``), by
    autograd through a strain tensor as ``OracleFourierD3.calculate`` does."""
    import torch

    from .fd3_oracle import ANGSTROM_TO_BOHR, HARTREE_TO_EV
    from .neighbours import neighbour_list
    positions_A = np.asarray(positions_A, dtype=np.float64)
    cell_np = np.asarray(cell_A, dtype=np.float64)
    i, j, S = neighbour_list(positions_A, cell_np, o.r_cut / ANGSTROM_TO_BOHR)
    cell = torch.tensor(cell_np)
    strain = torch.zeros(3, 3, dtype=torch.float64, requires_grad=True)
    scell = cell + torch.einsum("ab,Ab->Aa", strain, cell)
    pos = torch.tensor(positions_A, requires_grad=True)
    spos = pos + torch.einsum("ab,ib->ia", strain, pos)
    prm = (o.params if params is None else torch.as_tensor(params, dtype=torch.float64)).clone().requires_grad_(True)
    ei = torch.tensor(np.stack([i, j]), dtype=torch.long)
    sh = torch.tensor(S, dtype=torch.float64) @ scell
    cn = o.cn_smooth(spos * ANGSTROM_TO_BOHR, ei, sh * ANGSTROM_TO_BOHR)
    c = o.charges(cn)
    e = direct_energy_torch(spos * ANGSTROM_TO_BOHR, scell * ANGSTROM_TO_BOHR, o.species, c, o.eigs, o.QQ, prm,
                            r_max_A * ANGSTROM_TO_BOHR) * HARTREE_TO_EV
    e.backward()
    vol = abs(np.linalg.det(cell_np))
    return (float(e.detach()), (-pos.grad).numpy().copy(), strain.grad.numpy().copy() / vol,
            prm.grad.numpy().copy())
