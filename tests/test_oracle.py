"""Pins the CPU oracle independently of its own formulas (runs without a GPU).

Each oracle function is checked by something other than a retyped copy of it:
closed forms, invariants, the reference tests' known answers, or an independent
brute-force algorithm (polygon clipping, sparse-matrix operators, scipy solves).
"""

import math

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import spsolve

from conftest import rel_err


@pytest.fixture(scope="module")
def c1_mesh(surf_c1):
    from oracle import analysis as oa

    return oa.OracleMesh(surf_c1.vertices, surf_c1.faces, surf_c1.cortex_mask)


@pytest.fixture(scope="module")
def c2_mesh(surf_c2):
    from oracle import analysis as oa

    return oa.OracleMesh(surf_c2.vertices, surf_c2.faces, surf_c2.cortex_mask)


# --- operators (fastcw_lean.py:375-413) -------------------------------------------------------

def test_stiffness_equals_cotan_formula_on_one_triangle():
    from oracle import heat as oh

    rng = np.random.default_rng(0)
    V = rng.standard_normal((3, 3))
    F = np.array([[0, 1, 2]], np.int32)
    L, M, area, *_ = oh.consistent_operators(V, F)
    L = L.toarray()

    def angle(at, a, b):
        u, v = V[a] - V[at], V[b] - V[at]
        return math.acos(np.dot(u, v) / (np.linalg.norm(u) * np.linalg.norm(v)))

    for i, j, k in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
        assert L[i, j] == pytest.approx(-0.5 / math.tan(angle(k, i, j)), rel=1e-12)
    for i, j, k in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
        assert L[i, i] == pytest.approx(0.5 / math.tan(angle(j, i, k)) + 0.5 / math.tan(angle(k, i, j)), rel=1e-12)
    assert M.diagonal() == pytest.approx([area[0] / 3] * 3, rel=1e-15)


def test_operator_invariants(c2_mesh):
    from oracle import heat as oh

    L, M, area, *_ = oh.consistent_operators(c2_mesh.V, c2_mesh.F)
    assert abs(L - L.T).max() == 0.0
    scale = abs(L).max()
    assert np.max(np.abs(L @ np.ones(c2_mesh.n))) < 1e-12 * scale
    assert M.diagonal().sum() == pytest.approx(area.sum(), rel=1e-12)
    x = np.random.default_rng(1).standard_normal(c2_mesh.n)
    assert x @ (L @ x) > 0.0


def test_gradient_of_linear_field_is_exact_on_a_plane():
    from oracle import heat as oh

    g = np.linspace(0.0, 1.0, 7)
    X, Y = np.meshgrid(g, g)
    V = np.c_[X.ravel(), Y.ravel(), np.zeros(X.size)]
    rng = np.random.default_rng(2)
    V[:, :2] += rng.uniform(-0.03, 0.03, (V.shape[0], 2))
    F = []
    for r in range(6):
        for c in range(6):
            a = r * 7 + c
            F += [[a, a + 1, a + 8], [a, a + 8, a + 7]]
    F = np.array(F, np.int32)
    G, area = oh.gradient_operator(V, F)
    a = np.array([0.3, -1.7, 0.0])
    grad = (G @ (V @ a)).reshape(-1, 3)
    assert np.max(np.abs(grad - a)) < 1e-12


def test_fused_divergence_matches_sparse_form(c1_mesh):
    from oracle import heat as oh

    area, gb0, gb1, gb2 = oh.face_gradient_basis(c1_mesh.V, c1_mesh.F)
    u = np.random.default_rng(3).uniform(0.1, 1.0, (c1_mesh.n, 5))
    a = oh.fused_normalized_divergence(u, c1_mesh.F, area, gb0, gb1, gb2)
    b = oh.divergence_sparse(u, c1_mesh.V, c1_mesh.F)
    assert np.max(np.abs(a - b)) < 1e-12 * np.max(np.abs(b))
    # the hat-function gradients sum to zero on every face, so the divergence sums to zero
    assert np.max(np.abs(a.sum(axis=0))) < 1e-10 * np.max(np.abs(a))


def test_heat_oracle_solves_and_shift(c2_mesh):
    from oracle import heat as oh

    H = oh.HeatOracle(c2_mesh.V, c2_mesh.F)
    src = np.array([0, 11, 4000])
    u = H.heat_solve(src)
    A = (H.M + H.t * H.L).tocsr()
    rhs = np.zeros_like(u)
    rhs[src, np.arange(3)] = 1.0
    assert np.max(np.abs(A @ u - rhs)) < 1e-12
    assert np.all(u > 0)  # no negative heat on this near-Delaunay mesh
    D = H.compute_distance_batch(src)
    assert np.all(D[np.arange(3), src] == 0.0)
    assert np.min(D) > -c2_mesh.health_h
    # the Poisson solve of the same right-hand side by an independent direct solver
    div = oh.fused_normalized_divergence(u, H.F, H.area, H.gb0, H.gb1, H.gb2)
    P = (H.L + H.poisson_shift * H.M).tocsc()
    phi = np.column_stack([spsolve(P, div[:, j]) for j in range(3)])
    Dsp = (phi - phi[src, np.arange(3)][None, :]).T
    assert np.max(np.abs(D - Dsp)) < 1e-7 * np.max(D)


def test_heat_distance_converges_to_great_circle_on_spheres():
    from fastcorticalwiring_amd import meshes
    from oracle import heat as oh

    errs = []
    for level in (3, 4, 5):
        U, F = meshes.icosphere(level)
        V = 50.0 * U
        H = oh.HeatOracle(V, F)
        D = H.compute_distance_batch([0])[0]
        gc = 50.0 * np.arccos(np.clip(U @ U[0], -1, 1))
        far = gc > 20.0
        errs.append(np.max(np.abs(D[far] - gc[far]) / gc[far]))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 0.05


# --- lean geometry (lean_geometry.py:44-357) --------------------------------------------------

def test_face_cdf_closed_forms():
    from oracle import geometry as og

    assert og.face_cdf(0.0, 1.0, 2.0, 1.0, -1.0) == (0.0, 0.0)
    a, da = og.face_cdf(2.0, 0.0, 1.0, 1.0, 0.5)  # vertex order must not matter
    assert a == pytest.approx(0.125) and da == pytest.approx(0.5)
    a, da = og.face_cdf(0.0, 1.0, 2.0, 1.0, 1.5)
    assert a == pytest.approx(0.875) and da == pytest.approx(0.5)
    lo, _ = og.face_cdf(0.0, 1.0, 2.0, 1.0, 1.0 - 1e-12)
    hi, _ = og.face_cdf(0.0, 1.0, 2.0, 1.0, 1.0 + 1e-12)
    assert abs(lo - hi) < 1e-11
    assert og.face_cdf(0.0, 1.0, 2.0, 3.0, 5.0) == (3.0, 0.0)
    # the derivative integrates to the area: midpoint rule over a fine grid
    r = np.linspace(0, 2, 200001)
    dA = np.array([og.face_cdf(0.0, 0.7, 2.0, 1.0, x)[1] for x in r[::100]])
    assert np.trapezoid(dA, r[::100]) == pytest.approx(1.0, rel=1e-4)


def _euclid_fields(mesh, src):
    return np.linalg.norm(mesh.V[None, :, :] - mesh.V[np.asarray(src)][:, None, :], axis=2)


def test_area_perimeter_match_polygon_clipping(c2_mesh):
    from oracle import clipping as oc
    from oracle import geometry as og

    rng = np.random.default_rng(1)
    m = c2_mesh
    D = _euclid_fields(m, rng.choice(m.n, 3, replace=False))
    d_bad = D[0].copy()
    bad = rng.choice(m.n, m.n // 50, replace=False)
    d_bad[bad] += rng.uniform(0, 15, bad.size)
    for d, want_long in ((D[0], False), (D[1], False), (D[2], False), (d_bad, True)):
        radii = rng.uniform(0.5, np.nanmax(d) * 0.8, 12)
        A, dA, P, nl = og.area_perimeter(d, m.F, m.lean_area, m.g11, m.g12, m.g22, m.width, radii)
        if want_long:
            assert nl > 100
        for i, r in enumerate(radii):
            Ac, Pc = oc.clip_area_perimeter(m.V, m.F, d, r)
            assert abs(A[i] - Ac) / Ac < 1e-12
            assert abs(P[i] - Pc) / Pc < 1e-12
        # dA/dr against a central difference of the clipped area
        r = radii[0]
        h = 1e-5
        fd = (oc.clip_area_perimeter(m.V, m.F, d, r + h)[0] - oc.clip_area_perimeter(m.V, m.F, d, r - h)[0]) / (2 * h)
        assert dA[0] == pytest.approx(fd, rel=1e-5)


def test_per_source_geometry_against_clipping(c2_mesh):
    from fastcorticalwiring_amd import meshes
    from oracle import clipping as oc
    from oracle import geometry as og

    m = c2_mesh
    T = np.array(meshes.TWELVE_SCALES) * m.vertex_areas.sum()
    fr = np.array([0.25, 0.5, 0.75])
    D = _euclid_fields(m, [123, 4567])
    out = og.per_source_geometry(D, m.F, m.lean_area, m.g11, m.g12, m.g22, m.width, T, fr, 1e-12, 60)
    for row in range(2):
        for s in range(12):
            Ac, Pc = oc.clip_area_perimeter(m.V, m.F, D[row], out["radii"][row, s])
            assert abs(Ac - T[s]) / T[s] < 1e-10
            assert abs(out["perim"][row, s] - Pc) / Pc < 1e-10
        for e in range(33):
            Ac, _ = oc.clip_area_perimeter(m.V, m.F, D[row], out["extra_r"][row, e])
            assert abs(out["extra_a"][row, e] - Ac) / Ac < 1e-10
        r = out["radii"][row]
        assert np.all(np.diff(r) > 0)
        er = out["extra_r"][row].reshape(11, 3)
        assert np.all((er > r[:-1, None]) & (er < r[1:, None]))


def test_nan_and_infeasible_targets(c2_mesh):
    from oracle import geometry as og

    m = c2_mesh
    d = _euclid_fields(m, [7])[0]
    A, _, P, _ = og.area_perimeter(d, m.F, m.lean_area, m.g11, m.g12, m.g22, m.width, [np.nan, 0.0, -1.0])
    assert np.all(A == 0.0) and np.all(P == 0.0)
    total = m.lean_area.sum()
    r, _ = og.solve_radius(d, m.F, m.lean_area, m.width, [np.nan, 2 * total, 0.0, -5.0])
    assert np.isnan(r[0]) and np.isnan(r[1]) and r[2] == 0.0 and r[3] == 0.0


def test_field_health_screen(c2_mesh):
    from oracle import geometry as og

    m = c2_mesh
    chord = np.linalg.norm(m.V - m.V[5], axis=1)
    assert og.field_health(chord, m.V, 5, m.health_h)[2] is False
    neg = chord.copy()
    neg[: m.n // 10] = -2 * m.health_h
    n_frac, v_frac, bad = og.field_health(neg, m.V, 5, m.health_h)
    assert bad and n_frac == pytest.approx((m.n // 10) / m.n)
    short = 0.5 * chord
    assert og.field_health(short, m.V, 5, m.health_h)[2] is True
    nan = chord.copy()
    nan[3] = np.nan
    a, b, bad = og.field_health(nan, m.V, 5, m.health_h)
    assert bad and math.isnan(a) and math.isnan(b)


def _golden():
    import json
    import os

    with open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")) as f:
        return json.load(f)


def test_boundary_area_loss_fraction_known_answers():
    """Known answers of the reference's test_vertex_subsetting.py:52-58 (tests/golden fixture)."""
    from fastcorticalwiring_amd import mesh_prep
    from oracle import geometry as og

    g = _golden()["boundary_area_loss_fraction"]
    for fn in (og.boundary_area_loss_fraction, mesh_prep.boundary_area_loss_fraction):
        for c in g["cases"]:
            db = float(c["dist_to_boundary"])
            assert fn(c["radius"], db) == pytest.approx(c["expected"], abs=g["tolerance_abs"])
        assert math.isnan(fn(np.nan, 1.0)) and math.isnan(fn(0.0, 1.0))


def test_clean_vertex_samples():
    from oracle import analysis as oa

    pairs = [(3.0, 30.0), (1.0, 10.0), (2.0, 20.0), (2.0 + 1e-12, 21.0), (1.0, 11.0)]
    out = oa.clean_vertex_samples(pairs)
    assert out == [(1.0, 10.0), (2.0, 20.0), (3.0, 30.0)]
    assert oa.clean_vertex_samples([]) == []
    assert oa.clean_vertex_samples([(5.0, 1.0)]) == [(5.0, 1.0)]


def test_twelve_scales_match_reference_list():
    from fastcorticalwiring_amd import meshes

    ref = _golden()["twelve_scales"]["values"]  # run_pipeline.py:248-267, test_fastcw_lean.py:51-52
    assert list(meshes.TWELVE_SCALES) == ref
    logspaced = 0.002 * 25.0 ** (np.arange(12) / 11.0)
    assert np.max(np.abs(np.array(ref) - logspaced)) < 5e-9  # 8-decimal rounding of the log-spaced set


# --- loop semantics (core_analysis.py:1441-1632) ----------------------------------------------

def test_wiring_loop_layout_on_a_capped_mesh(c2_mesh):
    from fastcorticalwiring_amd import meshes
    from oracle import analysis as oa

    m = c2_mesh
    src = [0, 100, 2000]
    res = oa.wiring_costs(m, lambda blk: _euclid_fields(m, blk), meshes.TWELVE_SCALES, boundary_cap_fraction=0.2,
                          sources_sub=src)
    orig = m.sub_to_orig[src]
    assert np.isfinite(res["msd_weighted"][orig]).all()
    others = np.setdiff1d(np.arange(m.n_full), orig)
    assert np.isnan(res["msd_weighted"][others]).all() and np.isnan(res["radius"][:, others]).all()
    assert res["sample_indptr"][-1] == res["sample_radii"].size
    for oi in orig:
        r = res["sample_radii"][res["sample_indptr"][oi]:res["sample_indptr"][oi + 1]]
        assert np.all(np.diff(r) > 0)
    # Euclid fields on a mesh with a boundary: distance to boundary is the min over boundary vertices
    d = _euclid_fields(m, [100])[0]
    assert res["dist_to_boundary"][m.sub_to_orig[100]] == np.float32(d[m.boundary].min())


def test_bfs_order_small_known_answer():
    from fastcorticalwiring_amd import mesh_prep
    from oracle import analysis as oa

    F = np.array([[0, 1, 2], [1, 3, 2], [3, 4, 2], [5, 6, 7]], np.int32)
    expect = [0, 1, 2, 3, 4, 5, 6, 7]
    assert oa.bfs_order_all(8, F) == expect
    assert mesh_prep.bfs_order_all(8, F).tolist() == expect
    F2 = np.array([[4, 0, 3], [0, 1, 3], [1, 2, 3]], np.int32)
    assert oa.bfs_order_all(5, F2) == [0, 1, 3, 4, 2]


def test_normalize_scales_known_answer():
    """test_vertex_subsetting.py:60-62: scales are sorted ascending (tests/golden fixture)."""
    from fastcorticalwiring_amd import mesh_prep

    g = _golden()["normalize_scales"]
    assert tuple(mesh_prep.normalize_scales(g["input"])) == tuple(g["expected"])
