"""CPU tests of the batched form's planning (fd3_batch_create and the launch
merge): no device work is done when a plan is made, so these run without a
GPU."""
import numpy as np
import pytest

from fourierd3_amd import workloads as W
from fourierd3_amd.batch import BoxBatch


def test_batch_plan_merges_every_stage_across_boxes_and_species_sets():
    boxes = [W.strained(W.sio2_box(n, seed=i), 0.1, seed=i) for i, n in enumerate([60, 300, 900])]
    boxes += [W.strained(W.hcno_box(n, seed=10 + i), 0.1, seed=10 + i) for i, n in enumerate([80, 500, 1200])]
    bb = BoxBatch([b.numbers for b in boxes], [b.cell for b in boxes], device="cpu")
    info = bb.info()
    assert info["n_models"] == 2
    assert list(bb.model_of_box) == [0, 0, 0, 1, 1, 1]
    # one launch per stage and kernel variant, whatever the number of boxes
    per_box = info["box_launches"] / len(boxes)
    assert info["merged_launches"] < 2 * per_box
    # the packed workspace holds every box's own (single-box) workspace
    assert info["workspace_bytes"] > 0
    assert bb.offsets[-1] == sum(b.n_atoms for b in boxes)
    # compact species per box follow each model's ascending atomic numbers
    sp = bb.species.numpy()
    assert set(sp[: boxes[0].n_atoms]) == {0, 1} and set(sp[bb.offsets[3]: bb.offsets[4]]) == {0, 1, 2, 3}


def test_batch_plan_rejects_ewald_and_bad_inputs():
    boxes = [W.sio2_box(60, seed=1), W.sio2_box(90, seed=2)]
    with pytest.raises(NotImplementedError):
        BoxBatch([b.numbers for b in boxes], [b.cell for b in boxes], device="cpu", method="ewald")
    with pytest.raises(ValueError):
        BoxBatch([b.numbers for b in boxes], [boxes[0].cell], device="cpu")


def test_c4_plan_statistics():
    """The 256-box batching workload plans to a few dozen merged launches
    instead of ~6000 single-box launches."""
    boxes = W.batch_boxes(256, seed=0)
    bb = BoxBatch([b.numbers for b in boxes], [b.cell for b in boxes], device="cpu")
    info = bb.info()
    ns = {tuple(int(v) for v in bb.geoms[i].ns) for i in range(bb.n_boxes)}
    assert min(min(x) for x in ns) >= 16 and max(max(x) for x in ns) <= 128
    assert info["n_models"] == 2 and info["merged_launches"] * 50 < info["box_launches"]
