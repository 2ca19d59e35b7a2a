"""CPU FP64 oracle for the heat-method geodesic + wiring-metric path.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or as the timed CPU baseline); the product package
``fastcorticalwiring_amd`` never imports it.

Modules and the reference code each restates:

* ``heat``      - fastcw_lean.py:375-534 (gradient basis, L = G^T A G, lumped M,
                  fused normalised divergence, SuperLU heat/Poisson solves,
                  shift-to-source) and distance_engines.py:17-21 (mean edge).
* ``geometry``  - lean_geometry.py:44-367 (face metric, bin width, buckets,
                  area/perimeter/radius, per_source_geometry, field_health);
                  the scalar kernels run from ``lean_geometry.c`` (gcc).
* ``clipping``  - the legacy polygon-clipping area/perimeter of
                  core_analysis.py:184-519, used as an independent brute-force
                  check of ``geometry``.
* ``analysis``  - core_analysis.py:651-665, 759-777, 859-978, 1336-1439,
                  1441-1632 (submesh, boundary, BFS order, per-source loop:
                  health, distance to boundary, MSD, geometry, samples, CSR).

Pinning (what checks each function independently of its own formula) is in
``tests/test_oracle.py``; every function there has at least one check by a
closed form, an invariant, a known answer from the reference's tests, or a
brute-force alternative algorithm.
"""
