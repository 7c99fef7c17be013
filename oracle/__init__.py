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
"""
