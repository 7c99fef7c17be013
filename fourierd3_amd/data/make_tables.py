"""Convert the per-element D3 tables shipped with FourierD3 into ``tables.npz``.

The upstream tables (MIT licence, Copyright (c) 2026 Victoria Valeeva) live in
FourierD3's ``data/*.pt`` as torch zip archives indexed by atomic number:

* ``cnref.pt``  float32 (104, 7): reference coordination numbers, -1 = unused
  (loaded by FourierD3 ``src/fourierd3/utils.py:98-106``)
* ``rcov.pt``   float32 (119,): covalent radii in Bohr (``utils.py:87-95``)
* ``sqrtQz.pt`` float32 (119,): sqrt(Q_Z), the D3 r2r4 factors (``utils.py:76-84``)
* ``en.pt``     float64 (119,): Pauling electronegativities (``utils.py:109-117``)
* ``rcov_cn.pt`` float32 (119,): D4 covalent radii (``utils.py:120-127``)

The 21 MB ``reference-c6.pt`` (104, 104, 7, 7) is not part of the upstream
snapshot; ``fourierd3_amd.tables`` therefore builds a synthetic C6 table in the
same layout unless a real one is supplied.

Run once, from the repository root, with the upstream checkout at hand:
``python -m fourierd3_amd.data.make_tables /path/to/FourierD3/data``.
Only ``torch.load(weights_only=True)`` is used, which executes nothing from
the files.
"""
import sys
from pathlib import Path

import numpy as np


def main(src_dir: str) -> None:
    import torch

    src = Path(src_dir)
    out = {}
    for name in ("cnref", "rcov", "sqrtQz", "en", "rcov_cn"):
        t = torch.load(src / f"{name}.pt", weights_only=True)
        out[name] = t.numpy().copy()
    dst = Path(__file__).parent / "tables.npz"
    np.savez(dst, **out)
    print("wrote", dst, {k: (v.shape, v.dtype) for k, v in out.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/data")
