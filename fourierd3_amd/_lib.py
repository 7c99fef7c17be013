"""ctypes binding of ``libfd3hip.so`` (the C ABI declared in ``include/fd3.h``).

The library is built in-tree by ``python -m fourierd3_amd.build``.  Loading
fails loudly when it is missing: there is no fallback implementation.
Status codes map to the exception types the reference raises for the same
conditions (ValueError for bad arguments, NotImplementedError for options
that are not built, RuntimeError for HIP failures).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libfd3hip.so"

FD3_OK = 0
FD3_ERR_ARG = -1
FD3_ERR_UNSUPPORTED = -2
FD3_ERR_HIP = -3
FD3_ERR_WORKSPACE = -4

FD3_INTERP_P3M = 0
FD3_INTERP_LAGRANGE = 1
FD3_INTERP_EWALD = 2
FD3_DTYPE_F32 = 0
FD3_DTYPE_F64 = 1
FD3_NEED_GRAD = 1
FD3_NEED_PARAMS = 2
FD3_DETERMINISTIC = 4
FD3_NOALLOC_SCRATCH_BYTES = 256 << 20

# Every symbol include/fd3.h declares (checked by the CPU test suite).
EXPORTS = (
    "fd3_last_error", "fd3_version", "fd3_profile_enable", "fd3_profile_only", "fd3_profile_report",
    "fd3_eval_edges_workspace", "fd3_eval_edges",
    "fd3_eval_celllist_workspace", "fd3_eval_celllist",
    "fd3_eval_celllist_npt_workspace", "fd3_eval_celllist_npt",
    "fd3_eval_d4_npt_workspace", "fd3_eval_d4_npt",
    "fd3_eval_edges_npt_workspace", "fd3_eval_edges_npt",
    "fd3_eval_d4_workspace", "fd3_eval_d4",
    "fd3_eval_celllist_batch_workspace", "fd3_eval_celllist_batch",
    "fd3_eval_celllist_noalloc",
    "fd3_batch_create", "fd3_batch_workspace", "fd3_batch_run", "fd3_batch_info", "fd3_batch_destroy",
    "fd3_tune_set", "fd3_tune_clear", "fd3_plan_tiles",
    "fd3_eval_direct_workspace", "fd3_eval_direct",
    "fd3_neighbours_workspace", "fd3_neighbours_count", "fd3_neighbours_fill",
    "fd3_spread", "fd3_gather",
    "fd3_fft_workspace", "fd3_fft_r2c", "fd3_fft_c2r",
    "fd3_green_host", "fd3_green_table_host", "fd3_interp_weights_host",
)


def eval_flags(need_grad: bool, need_params: bool = False) -> int:
    """Flags of an fd3_eval_* call.  Under ``torch.use_deterministic_algorithms(True)``
    FD3_DETERMINISTIC is added: columns ordered by atom and the spread in that
    order, so repeated calls return bit-identical results."""
    import torch

    f = (FD3_NEED_GRAD if need_grad else 0) | (FD3_NEED_PARAMS if need_params else 0)
    if torch.are_deterministic_algorithms_enabled():
        f |= FD3_DETERMINISTIC
    return f


class Model(C.Structure):
    _fields_ = [
        ("n_species", C.c_int32), ("n_rank", C.c_int32),
        ("rcov", C.c_void_p), ("cnref", C.c_void_p), ("vq", C.c_void_p), ("eigs", C.c_void_p),
        ("qq", C.c_void_p), ("cncorr", C.c_void_p),
        ("s6", C.c_double), ("s8", C.c_double), ("a1", C.c_double), ("a2", C.c_double),
        ("r_cut", C.c_double),
    ]


class Geom(C.Structure):
    _fields_ = [
        ("cell", C.c_double * 9), ("ns", C.c_int32 * 3), ("order", C.c_int32), ("interp", C.c_int32),
    ]


class CnD4(C.Structure):
    _fields_ = [("r0", C.c_void_p), ("delta", C.c_void_p), ("selfcont", C.c_void_p), ("k0", C.c_double)]


class Outputs(C.Structure):
    _fields_ = [
        ("energy", C.c_void_p), ("grad_pos", C.c_void_p), ("grad_shifts", C.c_void_p),
        ("grad_cell", C.c_void_p), ("virial", C.c_void_p), ("grad_params", C.c_void_p),
        ("cn", C.c_void_p),
    ]


_lib = None

_P = C.c_void_p
_SIG = {
    "fd3_last_error": (C.c_char_p, []),
    "fd3_version": (C.c_int, []),
    "fd3_profile_enable": (C.c_int, [C.c_int]),
    "fd3_profile_only": (C.c_int, [C.c_char_p]),
    "fd3_profile_report": (C.c_int, [C.c_int, C.c_char_p, _P, _P, C.POINTER(C.c_int)]),
    "fd3_eval_edges_workspace": (C.c_int, [C.POINTER(Model), C.POINTER(Geom), C.c_int64, C.c_int64, C.c_int,
                                           C.POINTER(C.c_size_t)]),
    "fd3_eval_edges": (C.c_int, [C.POINTER(Model), C.POINTER(Geom), C.c_int64, _P, C.c_int, _P, C.c_int64, _P, _P,
                                 _P, C.c_int, C.c_int, _P, C.c_size_t, C.POINTER(Outputs), _P]),
    "fd3_eval_celllist_workspace": (C.c_int, [C.POINTER(Model), C.POINTER(Geom), C.c_int64, C.c_int,
                                              C.POINTER(C.c_size_t)]),
    "fd3_eval_celllist": (C.c_int, [C.POINTER(Model), C.POINTER(Geom), C.c_int64, _P, C.c_int, _P, C.c_int, _P,
                                    C.c_size_t, C.POINTER(Outputs), _P]),
    "fd3_eval_d4_workspace": (C.c_int, [C.POINTER(Model), C.POINTER(CnD4), C.POINTER(Geom), C.c_int64, C.c_int,
                                        C.POINTER(C.c_size_t)]),
    "fd3_eval_d4": (C.c_int, [C.POINTER(Model), C.POINTER(CnD4), C.POINTER(Geom), C.c_int64, _P, C.c_int, _P,
                              C.c_int, _P, C.c_size_t, C.POINTER(Outputs), _P]),
    "fd3_eval_celllist_npt_workspace": (C.c_int, [C.POINTER(Model), C.POINTER(Geom), C.c_double, C.c_int64,
                                                  C.c_int, C.POINTER(C.c_size_t)]),
    "fd3_eval_celllist_npt": (C.c_int, [C.POINTER(Model), C.POINTER(Geom), C.c_double, _P, _P, C.c_int64, _P,
                                        C.c_int, _P, C.c_int, _P, C.c_size_t, C.POINTER(Outputs), _P]),
    "fd3_eval_edges_npt_workspace": (C.c_int, [C.POINTER(Model), C.POINTER(Geom), C.c_double, C.c_int64,
                                               C.c_int64, C.c_int, C.POINTER(C.c_size_t)]),
    "fd3_eval_edges_npt": (C.c_int, [C.POINTER(Model), C.POINTER(Geom), C.c_double, _P, _P, C.c_int64, _P,
                                     C.c_int, _P, C.c_int64, _P, _P, _P, C.c_int, C.c_int, _P, C.c_size_t,
                                     C.POINTER(Outputs), _P]),
    "fd3_eval_d4_npt_workspace": (C.c_int, [C.POINTER(Model), C.POINTER(CnD4), C.POINTER(Geom), C.c_double,
                                            C.c_int64, C.c_int, C.POINTER(C.c_size_t)]),
    "fd3_eval_d4_npt": (C.c_int, [C.POINTER(Model), C.POINTER(CnD4), C.POINTER(Geom), C.c_double, _P, _P,
                                  C.c_int64, _P, C.c_int, _P, C.c_int, _P, C.c_size_t, C.POINTER(Outputs), _P]),
    "fd3_eval_celllist_batch_workspace": (C.c_int, [C.POINTER(Model), C.c_int32, C.POINTER(Geom),
                                                    C.POINTER(C.c_int64), C.c_int, C.POINTER(C.c_size_t)]),
    "fd3_eval_celllist_batch": (C.c_int, [C.POINTER(Model), C.c_int32, C.POINTER(Geom), C.POINTER(C.c_int64), _P,
                                          C.c_int, _P, C.c_int, _P, C.c_size_t, C.POINTER(Outputs), _P]),
    "fd3_eval_celllist_noalloc": (C.c_int, [C.POINTER(Model), C.POINTER(Geom), C.c_int64, _P, C.c_int, _P, C.c_int,
                                            C.POINTER(Outputs), _P]),
    "fd3_batch_create": (C.c_int, [C.c_int32, C.POINTER(Model), C.c_int32, C.POINTER(C.c_int32), C.POINTER(Geom),
                                   C.POINTER(C.c_int64), C.c_int, C.POINTER(C.c_void_p)]),
    "fd3_batch_workspace": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    "fd3_batch_run": (C.c_int, [C.c_void_p, _P, C.c_int, _P, _P, C.c_size_t, C.POINTER(Outputs), _P]),
    "fd3_batch_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "fd3_batch_destroy": (None, [C.c_void_p]),
    "fd3_eval_direct_workspace": (C.c_int, [C.POINTER(Model), _P, C.c_double, C.c_int64, C.c_int,
                                            C.POINTER(C.c_size_t)]),
    "fd3_eval_direct": (C.c_int, [C.POINTER(Model), _P, C.c_double, C.c_int64, _P, C.c_int, _P, C.c_int, _P,
                                  C.c_size_t, C.POINTER(Outputs), _P]),
    "fd3_tune_set": (C.c_int, [C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.POINTER(C.c_int32)]),
    "fd3_tune_clear": (C.c_int, []),
    "fd3_plan_tiles": (C.c_int, [C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.POINTER(C.c_int32)]),
    "fd3_neighbours_workspace": (C.c_int, [C.c_int64, C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_size_t)]),
    "fd3_neighbours_count": (C.c_int, [C.c_int64, _P, C.c_int, C.POINTER(C.c_double), C.c_double, _P, C.c_size_t,
                                       _P, _P]),
    "fd3_neighbours_fill": (C.c_int, [C.c_int64, _P, C.c_int, C.POINTER(C.c_double), C.c_double, _P, C.c_size_t,
                                      C.c_int64, _P, _P, _P, _P]),
    "fd3_spread": (C.c_int, [C.POINTER(Geom), C.c_int32, C.c_int32, C.c_int64, _P, C.c_int, _P, _P, _P, _P]),
    "fd3_gather": (C.c_int, [C.POINTER(Geom), C.c_int32, C.c_int32, C.c_int64, _P, C.c_int, _P, _P, _P, _P]),
    "fd3_fft_workspace": (C.c_int, [C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_size_t)]),
    "fd3_fft_r2c": (C.c_int, [C.POINTER(C.c_int32), C.c_int32, _P, _P, _P, C.c_size_t, _P]),
    "fd3_fft_c2r": (C.c_int, [C.POINTER(C.c_int32), C.c_int32, _P, _P, _P, C.c_size_t, _P]),
    "fd3_green_host": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int64, _P, _P, _P]),
    "fd3_green_table_host": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int64,
                                       _P, _P, _P]),
    "fd3_interp_weights_host": (C.c_int, [C.c_int32, C.c_int32, C.c_int64, _P, _P, _P]),
}


def lib():
    """Load (once) and return the ctypes handle; raises if the build is missing."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m fourierd3_amd.build` "
                "(there is no non-HIP fallback)")
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIG.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
        _load_tuning(h)
    return _lib


TUNING_PATH = Path(__file__).resolve().parent / "data" / "tuning.json"


def _load_tuning(h) -> None:
    """Register the autotuner's measured tiles (fourierd3_amd/data/tuning.json,
    written by ``python -m fourierd3_amd.autotune``) with the library."""
    import json
    if not TUNING_PATH.exists():
        return
    for e in json.loads(TUNING_PATH.read_text()).get("entries", []):
        ns = (C.c_int32 * 3)(*e["ns"])
        tiles = (C.c_int32 * 4)(*e["tiles"])
        h.fd3_tune_set(ns, int(e["n_species"]), int(e["n_rank"]), tiles)


def check(status: int) -> None:
    if status == FD3_OK:
        return
    msg = lib().fd3_last_error().decode(errors="replace")
    if status == FD3_ERR_ARG:
        raise ValueError(msg)
    if status == FD3_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if status == FD3_ERR_WORKSPACE:
        raise RuntimeError("workspace too small: " + msg)
    raise RuntimeError(msg)


def ptr(t) -> int | None:
    """Device (or host) address of a tensor, None for None."""
    return None if t is None else t.data_ptr()


def dtype_code(t) -> int:
    import torch
    if t.dtype == torch.float64:
        return FD3_DTYPE_F64
    if t.dtype == torch.float32:
        return FD3_DTYPE_F32
    raise ValueError(f"unsupported dtype {t.dtype}; expected float32 or float64")


def make_geom(cell_A, ns, order: int, interp: int) -> Geom:
    g = Geom()
    flat = [float(v) for v in list(cell_A.reshape(-1))]
    if len(flat) != 9:
        raise ValueError(f"cell of shape {tuple(cell_A.shape)} should be of shape (3, 3)")
    for i, v in enumerate(flat):
        g.cell[i] = v
    for i in range(3):
        g.ns[i] = int(ns[i])
    g.order = int(order)
    g.interp = int(interp)
    return g


def cell_array(cell_A) -> "C.Array":
    arr = (C.c_double * 9)()
    for i, v in enumerate([float(v) for v in list(cell_A.reshape(-1))]):
        arr[i] = v
    return arr


def profile_enable(on: bool = True) -> None:
    check(lib().fd3_profile_enable(1 if on else 0))


def profile_only(kernel: str | None) -> None:
    """Record only launches of `kernel` (None: every kernel)."""
    check(lib().fd3_profile_only(kernel.encode() if kernel else None))


def profile_report(max_entries: int = 64) -> dict:
    """{kernel name: (total ms, launches)} since profile_enable; synchronises."""
    import numpy as np
    names = C.create_string_buffer(64 * max_entries)
    ms = np.zeros(max_entries, np.float64)
    cnt = np.zeros(max_entries, np.int64)
    n = C.c_int()
    check(lib().fd3_profile_report(max_entries, names, ms.ctypes.data, cnt.ctypes.data, C.byref(n)))
    raw = names.raw
    out = {}
    for i in range(n.value):
        nm = raw[64 * i: 64 * i + 64].split(b"\0", 1)[0].decode()
        out[nm] = (float(ms[i]), int(cnt[i]))
    return out
