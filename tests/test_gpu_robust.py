"""GPU parity of the robust operators (intrinsic Delaunay triangulation, potpourri3d's
use_robust=True, distance_engines.py:78-79, 119-125) against the robust oracle
(oracle/intrinsic.py) on a stretched, capped ellipsoid whose triangulation is far from Delaunay.

Tolerances as for the plain operators: 1e-9 relative on distances (two FP64 direct solves of the
same operators), 1e-7 on the metrics computed from those fields.
"""

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ell():
    from fastcorticalwiring_amd import meshes
    from fastcorticalwiring_amd.engine import HeatContext
    from oracle import analysis as oa
    from oracle import intrinsic as oi

    V, F = meshes.icosphere(4)
    V = V * np.array([20.0, 20.0, 80.0])
    mask = V[:, 2] < 0.8 * V[:, 2].max()  # a cap removed: the surface has a boundary
    mesh = oa.OracleMesh(V, F, mask)
    print("robust fixture:", mesh.n, "vertices", flush=True)
    ctx = HeatContext(mesh.V, mesh.F, block=128, robust=True)
    R = oi.RobustHeatOracle(mesh.V, mesh.F)
    yield mesh, ctx, R
    ctx.close()


def test_robust_context_reports_its_triangulation(ell):
    mesh, ctx, R = ell
    assert ctx.info["robust"] == 1
    assert ctx.info["intrinsic_flips"] > 100 and ctx.info["intrinsic_skipped"] == 0
    assert abs(ctx.t - R.t) <= 1e-12 * R.t  # t from the input surface's mean edge


def test_robust_distances_match_oracle(ell):
    mesh, ctx, R = ell
    src = np.arange(0, mesh.n, 7)[:200]
    D = ctx.distances(src)
    Dref = R.compute_distance_batch(src)
    e = np.max(np.abs(D - Dref)) / np.max(np.abs(Dref))
    print(f"robust distances: max abs / max {e:.2e}")
    assert e < 1e-9
    assert np.all(D[np.arange(src.size), src] == 0.0)


def test_robust_differs_from_plain_operators(ell):
    from fastcorticalwiring_amd.engine import HeatContext

    mesh, ctx, R = ell
    src = [3, 500, 1200]
    with HeatContext(mesh.V, mesh.F, block=64, streams=1) as plain:
        Dp = plain.distances(src)
    Dr = ctx.distances(src)
    assert np.max(np.abs(Dp - Dr)) > 1e-6 * np.max(Dr)


def test_robust_wiring_uses_the_surface_geometry(ell):
    """Metrics of the robust fields on the surface's own faces: the oracle loop fed with the
    robust oracle's fields."""
    from fastcorticalwiring_amd import meshes
    from oracle import analysis as oa

    mesh, ctx, R = ell
    scales = meshes.TWELVE_SCALES
    total = float(np.sum(mesh.vertex_areas))
    T = np.array(sorted(scales)) * total
    fr = np.arange(1, 4) / 4.0
    ctx.set_metrics(T, fr, mesh.boundary, mesh.vertex_areas, eps=1e-6, health_h=mesh.health_h,
                    bin_width=mesh.width, boundary_cap_fraction=None)
    src = np.sort(np.random.default_rng(5).choice(mesh.n, 64, replace=False))
    res = ctx.wiring(src)
    ref = oa.wiring_costs(mesh, R.compute_distance_batch, scales, sources_sub=src.tolist())
    orig = mesh.sub_to_orig[src]
    sc = res["scalars"]
    assert np.all(sc[:, 0] == ref["field_flag"][orig])
    ok = sc[:, 0] == 0
    assert ok.sum() > 32
    for i in range(len(scales)):
        e = np.max(rel_err(res["radii"][ok, i], ref["radius"][i][orig[ok]]))
        p = np.max(rel_err(res["perimeters"][ok, i], ref["perimeter"][i][orig[ok]]))
        assert e < 1e-7 and p < 1e-7, (i, e, p)
    e = np.max(rel_err(sc[ok, 5], ref["msd_weighted"][orig[ok]].astype(np.float64)))
    assert e < 1e-6  # the reference stores MSD as float32


@pytest.mark.parametrize("height,fins", [(5.0, 1), (0.05, 1), (3.0, 3)])
def test_robust_distances_on_non_manifold_meshes(height, fins):
    """Meshes with an edge of 3-5 faces (fins hinged on one edge of a capped ellipsoid): the robust
    operators are built on the tufted cover, and the device distances match the robust oracle's at
    1e-9; the potpourri-compatible engine (robust always) accepts the mesh."""
    import test_intrinsic as ti

    from fastcorticalwiring_amd import distance_engines as de
    from fastcorticalwiring_amd.engine import HeatContext
    from oracle import analysis as oa
    from oracle import intrinsic as oi

    V, F = ti.fin_mesh(level=4, height=height, fins=fins)
    mask = V[:, 2] < 0.8 * V[:, 2].max()
    mesh = oa.OracleMesh(V, F, mask)
    print(f"fin mesh: {mesh.n} vertices, height {height}, {fins} fin(s)", flush=True)
    R = oi.RobustHeatOracle(mesh.V, mesh.F)
    assert R.cover
    src = np.concatenate([np.arange(0, mesh.n, 29)[:60], [mesh.n - 1]])  # the last fin apex too
    with HeatContext(mesh.V, mesh.F, block=128, robust=True) as ctx:
        assert ctx.info["tufted_cover"] == 1 and ctx.info["nonmanifold_edges"] == 1
        assert ctx.info["op_faces"] == R.F.shape[0]
        D = ctx.distances(src)
    Dref = R.compute_distance_batch(src)
    e = np.max(np.abs(D - Dref)) / np.max(np.abs(Dref))
    print(f"  distances vs robust oracle: max abs / max {e:.2e}", flush=True)
    assert e < 1e-9
    eng = de.create_distance_engine("mi355x_potpourri", mesh.V, mesh.F, {"block": 64})
    d = eng.compute_distance(int(src[3]))
    Rp = oi.RobustHeatOracle(mesh.V, mesh.F, t=0.7 ** 2)
    ref = Rp.compute_distance_batch([int(src[3])])[0]
    assert np.max(np.abs(d - ref)) < 1e-9 * np.max(ref)
