"""Write the golden fixtures of the parity tests: small seeded boxes with the
FP64 oracle's energy (eV), forces (eV/A) and 3x3 stress (eV/A^3).

The reference repository holds no golden vectors for this path, so these are
produced by the oracle (itself pinned by tests/test_oracle.py) and committed,
so that the GPU tests and later oracle changes are checked against fixed
numbers.  Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from fourierd3_amd import workloads as W  # noqa: E402
from oracle.fd3_oracle import OracleFourierD3  # noqa: E402

CASES = {
    # name: (box factory, oracle keyword arguments)
    "sio2_48_triclinic_spme": (lambda: W.strained(W.sio2_box(48, seed=101), 0.1, seed=101), {}),
    "hcno_64_triclinic_spme": (lambda: W.strained(W.hcno_box(64, seed=102), 0.1, seed=102), {}),
    "sio2_48_triclinic_pme4": (lambda: W.strained(W.sio2_box(48, seed=103), 0.1, seed=103),
                               {"method": "pme", "interpolation_nodes": 4}),
    "hcno_40_triclinic_ewald": (lambda: W.strained(W.hcno_box(40, seed=104), 0.1, seed=104),
                                {"method": "ewald", "k_cutoff": 6.0}),
}


def main():
    for name, (make, kw) in CASES.items():
        box = make()
        o = OracleFourierD3(box.numbers, box.cell, **kw)
        e, f, s = o.calculate(box.positions, box.cell)
        np.savez(os.path.join(HERE, name + ".npz"), numbers=box.numbers, positions=box.positions, cell=box.cell,
                 ns=np.asarray(o.ns), energy=np.float64(e), forces=f, stress=s)
        print(name, box.n_atoms, list(o.ns), e)


if __name__ == "__main__":
    main()
