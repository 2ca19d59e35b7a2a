"""Mesh preparation on the device (fcw_mesh_prep) against the host restatement in mesh_prep.py
(itself pinned to the oracle in test_host.py): every array bit for bit — submesh, face areas and
longest edges, vertex areas, boundary, adjacency and the breadth-first order — on the test
surfaces, a surface of two components with a masked hole, degenerate faces and a non-zero start."""

import numpy as np
import pytest

from fastcorticalwiring_amd import mesh_prep as mp
from fastcorticalwiring_amd import meshes

pytestmark = pytest.mark.gpu


def _host(V, F, M, start=0):
    V, F, M = mp.validate_mesh_inputs(V, F, M)
    Vs, Fs, s2o, o2s = mp.build_cortex_submesh(V, F, M)
    fa, fl, va = mp.face_geometry(Vs, Fs)
    ptr, cols = mp.adjacency_csr(Vs.shape[0], Fs)
    return dict(vertices=Vs, faces=Fs, sub_to_orig=s2o, orig_to_sub=o2s, face_areas=fa, face_L=fl, vertex_areas=va,
                boundary=mp.boundary_indices(Fs), order=mp.bfs_order_all(Vs.shape[0], Fs, start=start),
                adj_ptr=ptr, adj=cols)


def _compare(V, F, M, start=0, label=""):
    from fastcorticalwiring_amd import _lib

    print(f"mesh prep: {label} n={len(V)} m={len(F)}", flush=True)
    h = _host(V, F, M, start)
    d = _lib.mesh_prep_device(V, F, M, start=start)
    for k, hv in h.items():
        dv = d[k]
        assert dv.shape == np.asarray(hv).shape, (label, k, dv.shape, np.asarray(hv).shape)
        assert np.array_equal(dv, hv), (label, k)
    return h


@pytest.mark.parametrize("cfg", [0, 1, 3])
def test_device_prep_equals_host_on_test_surfaces(cfg):
    s = meshes.config_surface(cfg)
    h = _compare(s.vertices, s.faces, s.cortex_mask, label=f"config {cfg}")
    assert h["order"].size == h["vertices"].shape[0]


def test_device_prep_components_holes_degenerate_faces_and_start():
    V1, F1 = meshes.icosphere(3)
    V2, F2 = meshes.icosphere(2)
    V = np.concatenate([V1 * 30.0, V2 * 10.0 + 100.0])
    F = np.concatenate([F1, F2 + len(V1)])
    M = np.ones(len(V), bool)
    M[np.flatnonzero(V1[:, 2] > 0.7)] = False  # a hole: the first sphere gets a boundary
    M[len(V1) + 5] = False
    # a degenerate face (two coincident vertices) that the submesh drops
    V = np.concatenate([V, V[:1], V[:1] + 1e-9])
    F = np.concatenate([F, [[0, len(V) - 2, len(V) - 1]]])
    M = np.concatenate([M, [True, True]])
    h = _compare(V, F, M, label="two components")
    assert h["boundary"].size > 0
    _compare(V, F, M, start=37, label="start 37")
    _compare(V, F, M, start=10 ** 9, label="start out of range")
