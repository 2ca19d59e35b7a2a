"""Robust operators: the intrinsic Delaunay triangulation (host library vs the oracle's
independent restatement) and the Gram form of its operators, pinned by invariants and known
answers (potpourri3d, the reference's robust engine, is not installed)."""

import numpy as np
import pytest

from fastcorticalwiring_amd import _lib, meshes
from oracle import heat as oh
from oracle import intrinsic as oi


def _ellipsoid(level=3, stretch=4.0):
    V, F = meshes.icosphere(level)
    return np.ascontiguousarray(V * np.array([20.0, 20.0, 20.0 * stretch])), np.ascontiguousarray(F, np.int32)


def _gram_from_vectors(V, F):
    area, g0, g1, g2 = oh.face_gradient_basis(V, F)
    g = (g0, g1, g2)
    G = np.stack([np.sum(g[a] * g[b], axis=1) for a in range(3) for b in range(a, 3)], axis=1)
    return area, G


def _L_from_gram(F, area, G, n):
    idx = {(0, 0): 0, (0, 1): 1, (0, 2): 2, (1, 1): 3, (1, 2): 4, (2, 2): 5}
    rows, cols, vals = [], [], []
    for a in range(3):
        for b in range(3):
            rows.append(F[:, a])
            cols.append(F[:, b])
            vals.append(area * G[:, idx[(min(a, b), max(a, b))]])
    import scipy.sparse as sp

    L = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    L.sum_duplicates()
    return L


def test_gram_from_lengths_equals_gram_from_gradients():
    """On a mesh that is already Delaunay nothing flips, and the Gram matrix from edge lengths
    equals the one from the 3D basis gradients (two independent formulas)."""
    V, F = meshes.icosphere(2)
    V = V * 30.0
    faces, lens, area, gram, st = _lib.host_intrinsic(V, F.astype(np.int32))
    assert st["flips"] == 0 and st["delta"] == 0.0
    assert np.array_equal(faces, F)
    a3, G3 = _gram_from_vectors(V, F)
    assert np.allclose(area, a3, rtol=1e-12)
    assert np.max(np.abs(gram - G3)) <= 1e-10 * np.max(np.abs(G3))


def test_library_and_oracle_reach_the_same_delaunay_triangulation():
    V, F = _ellipsoid()
    n = V.shape[0]
    faces, lens, area, gram, st = _lib.host_intrinsic(V, F)
    Fo, Lo, flips, delta = oi.intrinsic_delaunay(V, F)
    assert st["flips"] > 100 and flips > 100 and st["skipped"] == 0
    assert oi.is_delaunay(faces, lens) and oi.is_delaunay(Fo, Lo)
    assert not oi.is_delaunay(F, np.stack([np.linalg.norm(V[F[:, (c + 2) % 3]] - V[F[:, (c + 1) % 3]], axis=1)
                                           for c in range(3)], axis=1))
    # the Delaunay triangulation of a generic surface is unique: same edge set and operators
    e_lib = {tuple(sorted((int(f[a]), int(f[b])))) for f in faces for a, b in ((0, 1), (1, 2), (2, 0))}
    e_orc = {tuple(sorted((int(f[a]), int(f[b])))) for f in Fo for a, b in ((0, 1), (1, 2), (2, 0))}
    assert e_lib == e_orc
    L_lib = _L_from_gram(faces.astype(np.int64), area, gram, n)
    L_orc, M_orc, a_orc = oi.intrinsic_operators(Fo, Lo, n)
    assert abs(L_lib - L_orc).max() <= 1e-10 * abs(L_orc).max()
    # flips preserve the surface: total area unchanged, Laplacian rows sum to zero
    a3, _ = _gram_from_vectors(V, F)
    assert abs(area.sum() - a3.sum()) <= 1e-12 * a3.sum() and abs(a_orc.sum() - a3.sum()) <= 1e-12 * a3.sum()
    assert np.max(np.abs(L_orc @ np.ones(n))) <= 1e-9 * abs(L_orc).max()
    # Delaunay cotan Laplacian: no positive off-diagonal weights
    off = (L_orc - sp_diag(L_orc)).tocoo()
    assert off.data.max() <= 1e-12 * abs(L_orc).max()


def sp_diag(A):
    import scipy.sparse as sp

    return sp.diags(A.diagonal())


def test_planar_kite_flips_to_the_short_diagonal():
    """Two triangles over a long diagonal with obtuse opposite angles: the flip gives the other
    diagonal, whose planar length is known."""
    V = np.array([[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.3, 0.0], [0.0, -0.3, 0.0]])
    F = np.array([[0, 1, 2], [1, 0, 3]], np.int32)
    faces, lens, area, gram, st = _lib.host_intrinsic(V, F)
    assert st["flips"] == 1
    assert sorted(map(tuple, np.sort(faces, axis=1))) == [(0, 2, 3), (1, 2, 3)]
    # the new edge (2, 3) has length 0.6 and both triangles have area 0.3
    for f, l in zip(faces, lens):
        c = [i for i in range(3) if f[i] not in (2, 3)][0]  # corner opposite the new edge
        assert abs(l[c] - 0.6) < 1e-14
    assert np.allclose(area, 0.3, rtol=1e-13)
    Fo, Lo, flips, _ = oi.intrinsic_delaunay(V, F)
    assert flips == 1


def test_mollification_restores_the_triangle_inequality_margin():
    """A sliver (l_a + l_b - l_c far below eps): delta lifts every triangle to margin eps."""
    V = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [2.0, 1e-9, 0.0], [1.0, 1.0, 0.0]])
    F = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    faces, lens, area, gram, st = _lib.host_intrinsic(V, F, mollify=1e-3)
    Fo, Lo, flips, delta = oi.intrinsic_delaunay(V, F, mollify=1e-3)
    assert st["delta"] > 0 and abs(st["delta"] - delta) <= 1e-15
    eps = 1e-3 * oh.mean_edge_length(V, F)
    for l in lens:
        for c in range(3):
            assert l[(c + 1) % 3] + l[(c + 2) % 3] - l[c] >= eps * (1 - 1e-9)


def test_robust_oracle_distances_close_to_great_circles():
    """On a badly shaped sphere mesh the robust operators give distances at least as close to the
    great-circle distance as the plain cotan operators."""
    V, F = meshes.icosphere(3)
    rng = np.random.default_rng(1)
    V = V + 0.25 * rng.standard_normal(V.shape) * (0.5 * oh.mean_edge_length(V, F))
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    R = 50.0
    V *= R
    src = [0, 17, 300]
    exact = R * np.arccos(np.clip((V[src] / R) @ (V / R).T, -1, 1))
    d_plain = oh.HeatOracle(V, F).compute_distance_batch(src)
    rob = oi.RobustHeatOracle(V, F)
    d_rob = rob.compute_distance_batch(src)
    assert rob.flips > 0
    err_plain = np.mean(np.abs(d_plain - exact)) / R
    err_rob = np.mean(np.abs(d_rob - exact)) / R
    assert err_rob <= err_plain * 1.05 and err_rob < 0.05


# ---- tufted cover (non-manifold edges, inconsistent orientation) --------------------------------

def fin_mesh(level=3, height=5.0, fins=1, stretch=3.0):
    """A stretched icosphere (radius 30 mm) with `fins` extra triangles hinged on the edge of its
    face 0: that edge then has 2 + fins faces. A fin's apex sits `height` times the edge length
    from the edge's midpoint, along the face's normal rotated by a different angle per fin."""
    V, F = meshes.icosphere(level)
    V = V * 30.0 * np.array([1.0, 1.0, stretch])
    F = F.astype(np.int32)
    i, j, k = (int(x) for x in F[0])
    e = V[j] - V[i]
    L = np.linalg.norm(e)
    nrm = np.cross(e, V[k] - V[i])
    nrm /= np.linalg.norm(nrm)
    side = np.cross(nrm, e / L)
    extra_v, extra_f = [], []
    for q in range(fins):
        ang = 0.35 + 0.5 * q
        d = np.cos(ang) * nrm + np.sin(ang) * side
        extra_v.append(0.5 * (V[i] + V[j]) + height * L * d)
        extra_f.append([i, j, V.shape[0] + q])
    return (np.ascontiguousarray(np.vstack([V] + extra_v)), np.ascontiguousarray(np.vstack([F, extra_f]), np.int32))


def _surface_area(V, F):
    return 0.5 * np.linalg.norm(np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]]), axis=1).sum()


def _check_laplacian(L, n):
    """Symmetric, rows summing to zero, PSD, and no positive off-diagonal weight (every edge of
    the cover Delaunay)."""
    scale = abs(L).max()
    assert abs(L - L.T).max() <= 1e-14 * scale
    assert np.max(np.abs(L @ np.ones(n))) <= 1e-12 * scale
    assert np.linalg.eigvalsh(L.toarray()).min() >= -1e-12 * scale
    off = (L - sp_diag(L)).tocoo()
    assert off.data.max() <= 1e-12 * scale


def test_cover_of_a_manifold_mesh_is_the_plain_triangulation():
    """On an edge-manifold, consistently oriented mesh the tufted cover is two copies of the surface
    that flip alike, so merging the mirror copies gives back the plain intrinsic triangulation."""
    V, F = _ellipsoid()
    n = V.shape[0]
    Fo, Lo, flips, _ = oi.intrinsic_delaunay(V, F)
    Fc, Lc, Wc, flips_c, _, nm = oi.tufted_delaunay(V, F)
    assert nm == 0 and flips_c == 2 * flips and np.all(Wc == 1.0) and Fc.shape == Fo.shape
    L0, M0, _ = oi.intrinsic_operators(Fo, Lo, n)
    L1, M1, _ = oi.intrinsic_operators(Fc, Lc, n, Wc)
    assert abs(L1 - L0).max() <= 1e-12 * abs(L0).max()
    assert np.allclose(M1.diagonal(), M0.diagonal(), rtol=1e-14)
    faces, lens, w, area, gram, st = _lib.host_intrinsic_cover(V, F)
    assert not st["cover"] and np.all(w == 1.0)


def test_orientation_does_not_change_the_operators():
    """Reversing every other face makes neighbours disagree on orientation (no halfedge twins); the
    cover then gives the same Laplacian as the consistently oriented surface."""
    V, F = _ellipsoid()
    n = V.shape[0]
    F2 = F.copy()
    F2[::2] = F2[::2][:, [0, 2, 1]]
    with pytest.raises(_lib.FcwError, match="fcw_host_intrinsic_cover"):
        _lib.host_intrinsic(V, F2)
    f0, l0, a0, g0, _ = _lib.host_intrinsic(V, F)
    f2, l2, w2, a2, g2, st = _lib.host_intrinsic_cover(V, F2)
    assert st["cover"] and st["nonmanifold_edges"] == 0 and np.all(w2 == 1.0)
    L0 = _L_from_gram(f0.astype(np.int64), a0, g0, n)
    L2 = _L_from_gram(f2.astype(np.int64), a2, g2, n)
    assert abs(L2 - L0).max() <= 1e-12 * abs(L0).max()
    Fo, Lo, Wo, _, _, _ = oi.tufted_delaunay(V, F2)
    Lr, _, _ = oi.intrinsic_operators(Fo, Lo, n, Wo)
    assert abs(Lr - L0).max() <= 1e-12 * abs(L0).max()


@pytest.mark.parametrize("height,fins", [(5.0, 1), (0.05, 1), (0.08, 2), (3.0, 3)])
def test_fin_meshes_library_and_oracle_agree(height, fins):
    """Fins hinged on one edge (3 to 5 faces about it): library and oracle reach the same operators;
    the Laplacian is symmetric, PSD, annihilates constants, has no positive off-diagonal weight, and
    the weighted areas add up to the surface's. Low fins force flips across the non-manifold wedges,
    leaving cover faces of weight 1/2."""
    V, F = fin_mesh(height=height, fins=fins)
    n = V.shape[0]
    faces, lens, w, area, gram, st = _lib.host_intrinsic_cover(V, F)
    assert st["cover"] and st["nonmanifold_edges"] == 1
    Fo, Lo, Wo, flips, _, nm = oi.tufted_delaunay(V, F)
    assert nm == 1 and flips == st["flips"]
    assert sorted(np.unique(w, return_counts=True)[1].tolist()) == sorted(np.unique(Wo, return_counts=True)[1].tolist())
    if height < 1.0:
        assert np.any(w == 0.5)
    L_lib = _L_from_gram(faces.astype(np.int64), area, gram, n)
    L_orc, M_orc, a_orc = oi.intrinsic_operators(Fo, Lo, n, Wo)
    assert abs(L_lib - L_orc).max() <= 1e-10 * abs(L_orc).max()
    _check_laplacian(L_orc, n)
    _check_laplacian(L_lib, n)
    A = _surface_area(V, F)
    assert abs(area.sum() - A) <= 1e-12 * A and abs(M_orc.diagonal().sum() - A) <= 1e-12 * A
