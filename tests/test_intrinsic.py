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
