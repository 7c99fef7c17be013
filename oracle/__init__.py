"""FP64 CPU oracle for the FourierD3 D3 energy/force/stress path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import anything from here, and only as
the checker (or, for ``bench.py``, as the timed CPU baseline).  The product
package ``fourierd3_amd`` never imports it.

Modules
* ``fd3_oracle``   - restatement of the reference algorithm in float64 torch on
  the CPU (CN, C6 chain, SPME/PME/Ewald energy; forces and stress by autograd
  through a strain tensor, as the reference calculator does).
* ``neighbours``   - brute-force periodic neighbour list (matscipy ``ijS``
  semantics).
* ``direct_sum``   - the real-space lattice sum the k-space energy converges to
  (closed-form pin for the k-space formula).
* ``workloads``    - seeded synthetic boxes of the benchmark configurations.

Pinning.  The reference holds no golden vectors for this path, so
``tests/test_oracle.py`` pins each function against an independent route:
mpmath quadrature of the Green's function, closed-form B-splines and CN
values, brute-force neighbour enumeration, direct lattice sums, supercell
scaling, symmetry invariances and FP64 finite differences.

Parity unpinned (restated from the published algorithms; nothing in the
reference's files fixes them):
* torch-pme ``get_ns_mesh`` rounding, the even-p node offset and the Lagrange
  node set (torch-pme version unpinned in the reference's pyproject.toml:17);
* torch-pme's Ewald k-grid extent (``method='ewald'``);
* the real D3 C6 values: ``data/reference-c6.pt`` is not shipped with the
  reference, so the C6 table is the synthetic one of
  ``fourierd3_amd.tables.synthetic_c6ref`` unless ``FD3_C6REF`` names the file.
"""
