"""Per-element D3 tables, the C6 reference table and its low-rank decomposition.

Host-side set-up code, run once per species set (FourierD3 ``__init__``).

* ``load_*``     - per-element tables indexed by atomic number, converted from
  FourierD3's ``data/*.pt`` (``src/fourierd3/utils.py:76-127``) into
  ``data/tables.npz``.
* ``c6ref``      - the (104, 104, 7, 7) reference-C6 table.  FourierD3 reads it
  from ``data/reference-c6.pt`` (``src/fourierd3/utils.py:9-22``), a file that is
  not distributed with the upstream snapshot.  A real table is used when the
  environment variable ``FD3_C6REF`` names a ``.pt``/``.npy`` file of that
  shape; otherwise a deterministic synthetic table in the same layout is built
  (``synthetic_c6ref``), so every code path runs and is testable.
* ``decomp``     - the rank search of ``src/fourierd3/utils.py:35-73``: ARPACK
  ``eigsh`` for k = 1, 2, ... with a fixed start vector until the maximum
  relative error of V diag(lambda) V^T drops below ``c6tol`` percent.
* ``get_ns_mesh`` - torch-pme's mesh rule, used by FourierD3
  ``src/fourierd3/core.py:186``: n_i = 2**ceil(log2(2*|a_i|/h + 1)).
"""
from __future__ import annotations

import math
import os
from functools import lru_cache
from pathlib import Path
from typing import Sequence, Tuple

import numpy as np

ANGSTROM_TO_BOHR = 1.0 / 0.52917726          # src/fourierd3/core.py:120
HARTREE_TO_EV = 27.21138505                  # src/fourierd3/calculator.py:54
PBE_PARAMS = (1.0, 0.7875, 0.4289, 4.4407)   # s6, s8, a1, a2; src/fourierd3/core.py:167-168
N_REF = 7                                    # reference environments per element

_DATA = Path(__file__).parent / "data" / "tables.npz"


@lru_cache(maxsize=1)
def _tables():
    with np.load(_DATA, allow_pickle=False) as z:
        return {k: z[k].copy() for k in z.files}


def load_cnref() -> np.ndarray:
    """(104, 7) float32 reference CNs, -1 = unused (``utils.py:98-106``)."""
    return _tables()["cnref"]


def load_rcov() -> np.ndarray:
    """(119,) float32 covalent radii in Bohr (``utils.py:87-95``)."""
    return _tables()["rcov"]


def load_sqrtQz() -> np.ndarray:
    """(119,) float32 sqrt(Q_Z) (D3 r2r4) (``utils.py:76-84``)."""
    return _tables()["sqrtQz"]


def load_en() -> np.ndarray:
    return _tables()["en"]


def load_rcov_cn() -> np.ndarray:
    return _tables()["rcov_cn"]


# Static dipole polarisabilities (atomic units) of the free atoms H..Kr, used
# only to give the synthetic C6 table physically sensible magnitudes.
_ALPHA_FREE = np.array([
    0.0, 4.50, 1.38, 164.0, 37.7, 20.5, 11.3, 7.4, 5.3, 3.74, 2.66,
    163.0, 71.2, 57.8, 37.3, 25.0, 19.4, 14.6, 11.1,
    290.0, 161.0, 97.0, 100.0, 87.0, 83.0, 68.0, 62.0, 55.0, 49.0, 47.0, 38.7,
    50.0, 40.0, 30.0, 29.0, 21.0, 16.8,
])


def _alpha_free(z: int) -> float:
    if z < len(_ALPHA_FREE):
        return float(_ALPHA_FREE[z])
    # heavier elements: repeat the 4th-period trend one period down, scaled
    base = 19 + (z - 37) % 18
    return float(_ALPHA_FREE[base]) * (1.25 + 0.02 * ((z - 37) // 18))


@lru_cache(maxsize=1)
def synthetic_c6ref() -> np.ndarray:
    """Deterministic synthetic stand-in for ``reference-c6.pt``, (104,104,7,7) f64.

    Casimir-Polder integrals of model dynamic polarisabilities with L = 12
    poles per reference, at frequencies w_{Z,l} = g_l (4.5 / alpha_free(Z))^(1/3),
    g_l = geomspace(0.1, 4.0) Hartree:
        alpha_{Z,p}(iw) = sum_l a_{Z,p,l} / (1 + (w / w_{Z,l})^2),
        C6[Za, Zb, p, q] = sum_{l,m} a_{Za,p,l} a_{Zb,q,m} 3/2 w_{Za,l} w_{Zb,m} / (w_{Za,l} + w_{Zb,m}).
    The static polarisability sum_l a_{Z,p,l} = alpha_free(Z) exp(-0.12 CN_p)
    is split over the poles by 0.7 d_Z + 0.3 d_{Z,p}, where d_Z and d_{Z,p} are
    Dirichlet(0.5) draws from ``default_rng(7)`` per element and per reference.
    Entries whose reference is unused (cnref = -1) are zero.
    1/(w + w') = int_0^inf exp(-(w + w') t) dt is a Gram kernel, so every
    species block of ``decomp`` is symmetric positive semi-definite, as the real
    Casimir-Polder table is.  With c6tol = 1 % the H/C/N/O block has rank 2 and O/Si rank 3.
    """
    cnref = load_cnref().astype(np.float64)
    nz = cnref.shape[0]
    valid = cnref >= 0
    rng = np.random.default_rng(7)
    afree = np.array([_alpha_free(z) if z > 0 else 4.5 for z in range(nz)])
    g = np.geomspace(0.1, 4.0, 12)
    w = g[None, :] * np.cbrt(4.5 / afree)[:, None]                          # (nz, L)
    base = rng.dirichlet(np.full(12, 0.5), size=nz)
    per = rng.dirichlet(np.full(12, 0.5), size=(nz, N_REF))
    prof = 0.7 * base[:, None, :] + 0.3 * per
    cn = np.where(valid, cnref, 0.0)
    a = afree[:, None, None] * np.exp(-0.12 * cn)[:, :, None] * prof       # (nz, 7, L)
    a = np.where(valid[:, :, None], a, 0.0)
    a[0] = 0.0
    wa = w[:, None, :, None]                                                # (nz, 1, L, 1)
    wb = w[None, :, None, :]                                                # (1, nz, 1, L)
    kern = 1.5 * wa * wb / (wa + wb)                                        # (nz, nz, L, L)
    c6 = np.einsum("apl,ablm,bqm->abpq", a, kern, a, optimize=True)
    return np.ascontiguousarray(c6)


@lru_cache(maxsize=1)
def c6ref() -> np.ndarray:
    """The (104, 104, 7, 7) C6 reference table in float64 (see module doc)."""
    path = os.environ.get("FD3_C6REF")
    if path:
        if path.endswith(".npy"):
            t = np.load(path, allow_pickle=False)
        else:
            import torch
            t = torch.load(path, weights_only=True).numpy()
        t = np.asarray(t, dtype=np.float64)
        if t.shape != (104, 104, 7, 7):
            raise ValueError(f"FD3_C6REF table has shape {t.shape}, expected (104, 104, 7, 7)")
        return t
    return synthetic_c6ref()


def c6_block(types: Sequence[int]) -> np.ndarray:
    """M[(X,p),(Y,q)] = C6ref[Z_X, Z_Y, p, q], shape (7 nS, 7 nS).

    Same block layout as ``src/fourierd3/utils.py:22``: species-major rows,
    reference index minor.
    """
    t = np.asarray(types, dtype=np.int64)
    blk = c6ref()[t][:, t]                                   # (nS, nS, 7, 7)
    return np.ascontiguousarray(blk.transpose(0, 2, 1, 3).reshape(len(t) * N_REF, len(t) * N_REF))


def maxrel_err(ref: np.ndarray, approx: np.ndarray) -> float:
    """Max |ref - approx| / |ref| over the nonzero entries of ``ref``
    (``src/fourierd3/utils.py:25-32``)."""
    nz = ref != 0
    if not nz.any():
        return 0.0
    return float(np.max(np.abs(ref[nz] - approx[nz]) / np.abs(ref[nz])))


@lru_cache(maxsize=64)
def _decomp_cached(types: Tuple[int, ...], c6tol: float, seed: int):
    from scipy.sparse.linalg import eigsh

    m = c6_block(types)
    n = m.shape[0]
    max_rank = n - 1
    v0 = np.random.default_rng(seed).standard_normal(n).astype(np.float64)
    k = 1
    while True:
        eigs, vecs = eigsh(m, k=k, v0=v0)
        err = maxrel_err(m, (vecs * eigs) @ vecs.T)
        if err * 100 < c6tol or k >= max_rank:
            break
        k += 1
    return eigs.copy(), vecs.copy(), k, err


def decomp(types: Sequence[int], c6tol: float, verbose: bool = False, seed: int = 42):
    """Low-rank decomposition M ~= V diag(lambda) V^T of the C6 block.

    Returns (eigs (r,), eigvecs (7 nS, r)) in float64, exactly the rank search
    of ``src/fourierd3/utils.py:35-73`` (ARPACK eigsh with start vector
    ``default_rng(seed).standard_normal(n)``, k = 1.. until 100*maxrel < c6tol
    or k = n - 1).
    """
    eigs, vecs, k, err = _decomp_cached(tuple(int(t) for t in types), float(c6tol), int(seed))
    if verbose:
        print(f"Using {k}-rank decomposition with maximum relative error: {err * 100} %")
        if k == len(types) * N_REF - 1 and err * 100 >= c6tol:
            print(f"Warning: reached maximum rank {k} without meeting tolerance {c6tol} %")
    return eigs.copy(), vecs.copy()


def get_ns_mesh(cell_bohr: np.ndarray, mesh_spacing_bohr: float) -> np.ndarray:
    """Mesh points per axis: 2**ceil(log2(2*|a_i|/h + 1)) (torch-pme ``get_ns_mesh``,
    called at ``src/fourierd3/core.py:186``)."""
    norms = np.linalg.norm(np.asarray(cell_bohr, dtype=np.float64), axis=1)
    approx = 2.0 * norms / mesh_spacing_bohr + 1.0
    return np.array([1 << int(math.ceil(math.log2(a))) for a in approx], dtype=np.int64)


# D4 CN pair-function constants (src/fourierd3/pair_pot.py:275-279)
CN_D4_K0, CN_D4_K1, CN_D4_K2, CN_D4_K3 = 7.5, 4.10451, 19.08857, 2 * 11.28174 ** 2


def cn_d4_tables(types: Sequence[int]):
    """Pair constants of ``CNPotential`` (``src/fourierd3/pair_pot.py:296-322``)
    in float64: r0 = rcov_cn_X + rcov_cn_Y (Bohr), delta = k1 exp(-(|EN_X -
    EN_Y| + k2)^2 / k3), and the self term f_XX(0) = delta_XX / 2 erfc(-k0)."""
    import math
    t = np.asarray(types, dtype=np.int64)
    rc = load_rcov_cn()[t].astype(np.float64)
    en = load_en()[t].astype(np.float64)
    r0 = rc[:, None] + rc[None, :]
    delta = CN_D4_K1 * np.exp(-(np.abs(en[:, None] - en[None, :]) + CN_D4_K2) ** 2 / CN_D4_K3)
    selfcont = np.diagonal(delta) / 2.0 * math.erfc(-CN_D4_K0)
    return {"r0": r0, "delta": delta, "selfcont": selfcont.copy(), "k0": CN_D4_K0}


def species_tables(types: Sequence[int], params=PBE_PARAMS):
    """Per-species constants of ``D3Potential`` (``src/fourierd3/pair_pot.py:49-105``)
    and of the CN model (``src/fourierd3/core.py:152-161``), in float64.

    Returns a dict with rcov (nS,), cnref (nS,7), sqrtQ (nS,), QQ (nS,nS),
    R0 (nS,nS) = a1 sqrt(3 QQ) + a2, selfcont (nS,) = s6/R0^6 + 3 s8 QQ/R0^8.
    """
    t = np.asarray(types, dtype=np.int64)
    s6, s8, a1, a2 = (float(p) for p in params)
    sq = load_sqrtQz()[t].astype(np.float64)
    qq = np.outer(sq, sq)
    r0 = a1 * np.sqrt(3.0 * qq) + a2
    selfc = s6 / np.diag(r0) ** 6 + 3.0 * s8 * np.diag(qq) / np.diag(r0) ** 8
    return {
        "rcov": load_rcov()[t].astype(np.float64),
        "cnref": load_cnref()[t].astype(np.float64),
        "sqrtQ": sq,
        "QQ": qq,
        "R0": r0,
        "selfcont": selfc,
    }
