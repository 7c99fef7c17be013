"""The differentiable real-space lattice-sum oracle (``oracle.direct_sum``):
its torch form against the numpy form with scipy's adaptive quadrature, its
Gauss-Legendre tail against closed forms, and its autograd forces, strain
derivative and parameter gradient against central finite differences.  CPU,
FP64."""
import itertools
import math

import numpy as np
import pytest
import torch

from oracle import direct_sum as D
from oracle.fd3_oracle import ANGSTROM_TO_BOHR, HARTREE_TO_EV, OracleFourierD3
from oracle.neighbours import neighbour_list


def _tiny_box(seed, n, L_A, tric=True):
    rng = np.random.default_rng(seed)
    cell = np.eye(3) * L_A
    if tric:
        cell = cell @ (np.eye(3) + rng.uniform(-0.1, 0.1, (3, 3)))
    for _ in range(1000):
        pos = rng.uniform(0, 1, (n, 3)) @ cell
        i, j, S = neighbour_list(pos, cell, 1.2)
        if len(i) == 0:
            break
    numbers = np.array([1, 8, 6, 8, 1, 7, 14, 8][:n])
    return numbers, pos, cell


def _gap_to_cutoff(pos, cell, r_max):
    """Smallest |d - r_max| over every pair-image distance d."""
    d = 1.0 / np.linalg.norm(np.linalg.inv(cell).T, axis=1)
    nimg = [int(math.ceil((r_max + 1) / dd)) + 1 for dd in d]
    diff = pos[None, :, :] - pos[:, None, :]
    best = np.inf
    for sh in itertools.product(*[range(-k, k + 1) for k in nimg]):
        r = np.linalg.norm(diff + np.asarray(sh) @ cell, axis=-1)
        best = min(best, float(np.abs(r - r_max).min()))
    return best


def test_tail_quadrature_matches_closed_forms():
    """I6 = atan(R^3 / rmax^3) / (3 R^3); I8 by its series in (R / rmax)^8,
    int_T^inf t^2/(t^8+1) dt = sum_k (-1)^k T^-(8k+5) / (8k+5), T = rmax / R."""
    R = torch.tensor([4.0, 6.5, 9.5], dtype=torch.float64)
    for rm in (20.0, 37.8, 60.0):
        i6, i8 = D.tail_integrals_gl(R, rm)
        r = R.numpy()
        ref6 = np.arctan(r ** 3 / rm ** 3) / (3 * r ** 3)
        T = rm / r
        ref8 = sum((-1) ** k * T ** -(8 * k + 5) / (8 * k + 5) for k in range(12)) / r ** 5
        assert np.allclose(i6.numpy(), ref6, rtol=1e-13, atol=0)
        assert np.allclose(i8.numpy(), ref8, rtol=1e-13, atol=0)


@pytest.mark.parametrize("seed,n", [(3, 2), (4, 4), (5, 8)])
def test_torch_form_matches_numpy_form(seed, n):
    numbers, pos, cell = _tiny_box(seed, n, 7.0)
    o = OracleFourierD3(numbers, cell)
    i, j, S = neighbour_list(pos, cell, 6.0)
    ei = torch.tensor(np.stack([i, j]))
    sh = torch.tensor(S, dtype=torch.float64) @ torch.tensor(cell)
    c = o.charges(o.cn_smooth(torch.tensor(pos) * ANGSTROM_TO_BOHR, ei, sh * ANGSTROM_TO_BOHR)).detach().numpy()
    r_max = 12.0
    e_np = D.direct_energy(pos * ANGSTROM_TO_BOHR, cell * ANGSTROM_TO_BOHR, o.species.numpy(), c, o.eigs.numpy(),
                           o.R0.numpy(), o.QQ.numpy(), o.params.numpy(), r_max * ANGSTROM_TO_BOHR)
    e_t = D.direct_calculate(o, pos, cell, r_max)[0]
    # scipy's adaptive quadrature of the r^-8 tail is good to ~1e-5 of that
    # term (about 1e-12 of the energy)
    assert e_t == pytest.approx(e_np * HARTREE_TO_EV, rel=1e-10)


def test_autograd_matches_finite_differences():
    numbers, pos, cell = _tiny_box(6, 5, 6.5)
    o = OracleFourierD3(numbers, cell)
    r_max = 11.0
    # a pair image crossing r_max under the displacement would make the FD jump
    while _gap_to_cutoff(pos, cell, r_max) < 1e-3:
        r_max += 0.0137
    e0, f, s, gp = D.direct_calculate(o, pos, cell, r_max)
    h = 1e-4
    f_fd = np.zeros_like(pos)
    for a in range(len(pos)):
        for d in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[a, d] += h
            p2[a, d] -= h
            f_fd[a, d] = -(D.direct_calculate(o, p1, cell, r_max)[0] - D.direct_calculate(o, p2, cell, r_max)[0]) / (2 * h)
    assert np.abs(f - f_fd).max() < 1e-7
    # strain: r' = r (1 + eps)^T, cell' = cell (1 + eps)^T
    vol = abs(np.linalg.det(cell))
    s_fd = np.zeros((3, 3))
    for a in range(3):
        for b in range(3):
            eps = np.zeros((3, 3))
            eps[a, b] = h
            e1 = D.direct_calculate(o, pos @ (np.eye(3) + eps).T, cell @ (np.eye(3) + eps).T, r_max)[0]
            e2 = D.direct_calculate(o, pos @ (np.eye(3) - eps).T, cell @ (np.eye(3) - eps).T, r_max)[0]
            s_fd[a, b] = (e1 - e2) / (2 * h) / vol
    assert np.abs(s - s_fd).max() < 1e-8
    g_fd = np.zeros(4)
    for k in range(4):
        p1, p2 = o.params.clone(), o.params.clone()
        p1[k] += h
        p2[k] -= h
        g_fd[k] = (D.direct_calculate(o, pos, cell, r_max, params=p1)[0]
                   - D.direct_calculate(o, pos, cell, r_max, params=p2)[0]) / (2 * h)
    assert np.abs(gp - g_fd).max() < 1e-8
    # no net force
    assert np.abs(f.sum(0)).max() < 1e-12
