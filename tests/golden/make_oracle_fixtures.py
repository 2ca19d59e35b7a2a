"""Write the FP64 oracle's outputs on the folded workloads as test fixtures.

For each of the folded surfaces (icosphere level 5, 6 and 7: 10,242 / 40,962 /
163,842 vertices, 10% masked cap; ``meshes.config_surface(1..3)``) this runs the
CPU FP64 oracle (``oracle.heat.HeatOracle`` + ``oracle.analysis.wiring_costs``,
the restatement of fastcw_lean.py:467-534 and core_analysis.py:1441-1632) on 256
seeded cortical sources and stores, per source:

* the distance field at 512 seeded probe vertices (float64);
* the health fractions and flag, distance to boundary, MSD unweighted / weighted
  (float32, as the reference's result arrays hold them, core_analysis.py:709-724);
* the radii, perimeters and supplementary (radius, area) samples in float64
  (lean_geometry.per_source_geometry, lean_geometry.py:302-340);
* the cleaned sample pairs with the boundary cap 0.1 (float32 CSR over the 256
  sources, core_analysis.py:1599-1632).

The submesh's SHA-256 is stored with them, so a test can tell that the surface it
rebuilt is the one the fixture was made from. Inputs are data generated by this
repository's own seeded surface code; nothing from the reference is imported.

    python tests/golden/make_oracle_fixtures.py [1 2 3]
"""

import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

N_SOURCES = 256
N_PROBES = 512
CAP = 0.1
Q = 3


def fixture_path(idx):
    return os.path.join(HERE, f"oracle_folded_c{idx}.npz")


def submesh_sha(V, F):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(V, dtype=np.float64).tobytes())
    h.update(np.ascontiguousarray(F, dtype=np.int32).tobytes())
    return h.hexdigest()


def sample_sources(n, seed=11):
    return np.sort(np.random.default_rng(seed).choice(n, N_SOURCES, replace=False)).astype(np.int32)


def sample_probes(n, seed=12):
    return np.sort(np.random.default_rng(seed).choice(n, N_PROBES, replace=False)).astype(np.int32)


def make(idx):
    from fastcorticalwiring_amd import meshes
    from oracle import analysis as oa
    from oracle import geometry as og
    from oracle import heat as oh

    t0 = time.perf_counter()
    s = meshes.config_surface(idx)
    mesh = oa.OracleMesh(s.vertices, s.faces, s.cortex_mask)
    H = oh.HeatOracle(mesh.V, mesh.F)
    src = sample_sources(mesh.n)
    probes = sample_probes(mesh.n)
    D = H.compute_distance_batch(src)
    rows = {int(v): i for i, v in enumerate(src)}
    scales = meshes.TWELVE_SCALES
    ref = oa.wiring_costs(mesh, lambda blk: D[[rows[int(b)] for b in blk]], scales, Q,
                          boundary_cap_fraction=CAP, sources_sub=src.tolist(), block=N_SOURCES)
    T = np.array(sorted(scales)) * float(np.sum(mesh.vertex_areas))
    fr = np.arange(1, Q + 1) / (Q + 1.0)
    geo = og.per_source_geometry(D, mesh.F, mesh.lean_area, mesh.g11, mesh.g12, mesh.g22, mesh.width, T, fr,
                                 1e-10, 60)
    orig = mesh.sub_to_orig[src]
    flag = ref["field_flag"][orig]
    for key in ("radii", "perim", "extra_r", "extra_a"):
        geo[key][flag > 0] = np.nan
    ip = ref["sample_indptr"]
    lens = ip[orig + 1] - ip[orig]
    indptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    sr = np.concatenate([ref["sample_radii"][ip[o]:ip[o + 1]] for o in orig]).astype(np.float32)
    sa = np.concatenate([ref["sample_areas"][ip[o]:ip[o + 1]] for o in orig]).astype(np.float32)
    out = dict(
        config=np.int32(idx), n=np.int32(mesh.n), m=np.int32(mesh.F.shape[0]),
        mesh_sha=np.frombuffer(bytes.fromhex(submesh_sha(mesh.V, mesh.F)), np.uint8),
        src=src, probes=probes, D_probe=np.ascontiguousarray(D[:, probes]),
        health_h=np.float64(mesh.health_h), width=np.float64(mesh.width), targets=T, fracs=fr,
        cap=np.float64(CAP), t=np.float64(H.t),
        flag=flag.astype(np.uint8), neg_frac=ref["neg_frac"][orig], viol_frac=ref["viol_frac"][orig],
        dist_to_boundary=ref["dist_to_boundary"][orig], msd_unweighted=ref["msd_unweighted"][orig],
        msd_weighted=ref["msd_weighted"][orig],
        radii=geo["radii"], perimeters=geo["perim"], extra_r=geo["extra_r"], extra_a=geo["extra_a"],
        sample_indptr=indptr, sample_r=sr, sample_a=sa,
    )
    np.savez_compressed(fixture_path(idx), **out)
    print(f"config {idx}: n={mesh.n} m={mesh.F.shape[0]} flagged={int(flag.sum())} "
          f"samples={indptr[-1]} in {time.perf_counter() - t0:.1f}s -> {fixture_path(idx)}", flush=True)


def load(idx):
    with np.load(fixture_path(idx)) as z:
        return {k: z[k] for k in z.files}


if __name__ == "__main__":
    for a in (sys.argv[1:] or ["1", "2", "3"]):
        make(int(a))
