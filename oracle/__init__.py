'''CPU oracle for the grid-density path.  TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the benchmark's cpu_baseline leg use this
package, always as the checker.  The product (orbkit_amd) never imports it.

  orbkit_oracle.c   FP64 C restatement of orbkit's AO kernel and MO loop
                    (orbkit/c_grid-based.c:9-79, orbkit/c_support.c:8-195,
                    orbkit/cy_core.pyx:49-101)
  core_oracle.py    numpy restatement of the orchestration
                    (orbkit/core.py:38-308, :314-605, :607-839)
  _ref/             orbkit's own c_lcreator compiled from its sources
                    (Makefile target `ref`), used to pin orbkit_oracle.c

Pinning status per function is listed in core_oracle.py's header.
'''
