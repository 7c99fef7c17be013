"""Tile overrides of the mesh plan (fd3_tune_set / fd3_plan_tiles /
fd3_tune_clear) and the committed tuning table.  Host only: planning touches
no device."""
import json

import pytest

from fourierd3_amd import _lib as L
from fourierd3_amd import autotune as A

NS = [256, 256, 256]


@pytest.fixture
def clean_table():
    L.check(L.lib().fd3_tune_clear())
    yield
    L.check(L.lib().fd3_tune_clear())
    L._load_tuning(L.lib())


def test_default_tiles(clean_table):
    ty, tx, rt, kt = A.plan_tiles(NS, 4, 2)
    # z lines: M = 128 complex per line, 2048 values per block; y lines: 8192 per block
    assert (ty, tx) == (16, 32)
    # the split x k-space kernel (nx = ny = 256) takes five tiles per block
    assert rt == 2 and kt == 5
    # other meshes: 1, 2 or 4 ky tiles per block
    assert A.plan_tiles([128, 128, 128], 4, 2)[3] in (1, 2, 4)
    # four species: at most 8 spectra per thread -> rank tile 2
    assert A.plan_tiles(NS, 4, 6)[2] == 2
    assert A.plan_tiles(NS, 2, 6)[2] == 4


def test_override_taken_and_cleared(clean_table):
    base = A.plan_tiles(NS, 4, 2)
    A.set_tiles(NS, 4, 2, [8, 16, 1, 1])
    assert A.plan_tiles(NS, 4, 2) == [8, 16, 1, 1]
    # other keys are untouched
    assert A.plan_tiles(NS, 4, 3) == A.plan_tiles(NS, 4, 3)
    assert A.plan_tiles([128, 128, 128], 4, 2)[0] != 0
    L.check(L.lib().fd3_tune_clear())
    assert A.plan_tiles(NS, 4, 2) == base


@pytest.mark.parametrize("ns,tiles", [
    (NS, [3, 0, 0, 0]),                # not a power of two
    (NS, [64, 0, 0, 0]),               # 64 lines x 128 exceed the z-pass register budget
    (NS, [0, 64, 0, 0]),               # 64 lines x 256 exceed the y-pass budget
    (NS, [0, 0, 3, 0]),                # more ranks per tile than the model has
    (NS, [0, 0, 0, 65]),               # split x k-space kernel: at most 64 tiles per block
    ([128, 128, 128], [0, 0, 0, 32]),  # other meshes: at most 16 k tiles per block
    ([128, 128, 128], [0, 0, 0, 3]),   # ... a power of two
    ([32, 32, 32], [0, 0, 0, 8]),      # ... that divides the tile rows (32 / 8 ky rows per tile)
])
def test_invalid_override_ignored(clean_table, ns, tiles):
    base = A.plan_tiles(ns, 4, 2)
    A.set_tiles(ns, 4, 2, tiles)
    assert A.plan_tiles(ns, 4, 2) == base


@pytest.mark.parametrize("kt", [1, 2, 5, 8, 64])
def test_split_x_kspace_takes_any_tile_count(clean_table, kt):
    A.set_tiles(NS, 4, 2, [0, 0, 0, kt])
    assert A.plan_tiles(NS, 4, 2)[3] == kt


def test_zero_keeps_plan_choice(clean_table):
    base = A.plan_tiles(NS, 4, 2)
    A.set_tiles(NS, 4, 2, [0, 8, 0, 0])
    assert A.plan_tiles(NS, 4, 2) == [base[0], 8, base[2], base[3]]


def test_bad_arguments_rejected(clean_table):
    import ctypes as C
    ns = (C.c_int32 * 3)(*NS)
    t = (C.c_int32 * 4)()
    assert L.lib().fd3_plan_tiles(ns, 0, 2, t) != 0
    assert L.lib().fd3_plan_tiles((C.c_int32 * 3)(255, 256, 256), 4, 2, t) != 0
    assert L.lib().fd3_tune_set(ns, 4, 0, t) != 0


def test_committed_table_is_valid(clean_table):
    if not L.TUNING_PATH.exists():
        pytest.skip("no tuning table committed")
    table = json.loads(L.TUNING_PATH.read_text())
    for e in table["entries"]:
        assert len(e["ns"]) == 3 and len(e["tiles"]) == 4
        L.check(L.lib().fd3_tune_clear())
        A.set_tiles(e["ns"], e["n_species"], e["n_rank"], e["tiles"])
        # every committed tile is one the library takes
        assert A.plan_tiles(e["ns"], e["n_species"], e["n_rank"]) == e["tiles"]
