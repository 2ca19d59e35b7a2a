"""Host-side checks that need no GPU: the C-ABI library, the host analysis and
factorisation inside it, mesh preparation, workloads, adapter rules."""

import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import spsolve

from conftest import ROOT, rel_err


def _declared_functions():
    src = open(os.path.join(ROOT, "include", "fcw.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(fcw_\w+)\s*\(", src, flags=re.M)))


def test_library_loads_and_exports_every_declared_symbol():
    from fastcorticalwiring_amd import _lib

    L = _lib.lib()
    names = _declared_functions()
    assert len(names) >= 14
    for name in names:
        assert hasattr(L, name), name
    assert sorted(_lib.EXPORTED_SYMBOLS) == names


def test_build_info_self_report():
    from fastcorticalwiring_amd import _lib

    info = _lib.build_info()
    assert "gfx950" in info and "abi=1" in info


@pytest.mark.parametrize("leaf", [16, 64, 200])
def test_host_factorisation_matches_scipy(surf_c2, leaf):
    from fastcorticalwiring_amd import _lib
    from oracle import analysis as oa
    from oracle import heat as oh

    m = oa.OracleMesh(surf_c2.vertices, surf_c2.faces, surf_c2.cortex_mask)
    L, M, *_ = oh.consistent_operators(m.V, m.F)
    t = oh.mean_edge_length(m.V, m.F) ** 2
    rng = np.random.default_rng(0)
    B = rng.standard_normal((m.n, 3))
    X, st = _lib.host_factor_solve(m.V, m.F, t, 1.0, B, leaf_size=leaf)
    A = (M + t * L).tocsc()
    Xr = np.column_stack([spsolve(A, B[:, j]) for j in range(3)])
    assert np.max(np.abs(X - Xr)) < 1e-11 * np.max(np.abs(Xr))
    assert st["nnodes"] > 1 and st["factor_nnz"] > m.n
    # Poisson system with a zero-mean right-hand side (the divergence always sums to zero)
    B0 = B - B.mean(axis=0)
    X, _ = _lib.host_factor_solve(m.V, m.F, 1.0, 1e-10, B0, leaf_size=leaf)
    P = (L + 1e-10 * M).tocsc()
    Xr = np.column_stack([spsolve(P, B0[:, j]) for j in range(3)])
    d = (X - X[0]) - (Xr - Xr[0])  # compare up to the constant (near-null) mode
    assert np.max(np.abs(d)) < 1e-8 * np.max(np.abs(Xr - Xr[0]))


def test_host_factorisation_rejects_zero_area_faces():
    from fastcorticalwiring_amd import _lib

    V = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [0, 1, 0]], float)
    F = np.array([[0, 1, 2], [0, 1, 3]], np.int32)
    with pytest.raises(_lib.FcwError, match="Zero-area"):
        _lib.host_factor_solve(V, F, 1.0, 1.0, np.ones((4, 1)))
    with pytest.raises(_lib.FcwError, match="out of range"):
        _lib.host_factor_solve(V, np.array([[0, 1, 7]], np.int32), 1.0, 1.0, np.ones((4, 1)))


def test_host_factorisation_on_disconnected_mesh():
    """Two separate spheres: the dissection must split components with an empty separator."""
    from fastcorticalwiring_amd import _lib, meshes
    from oracle import heat as oh

    U, F = meshes.icosphere(2)
    V = np.vstack([U * 10, U * 10 + 50])
    F2 = np.vstack([F, F + U.shape[0]]).astype(np.int32)
    L, M, *_ = oh.consistent_operators(V, F2)
    B = np.random.default_rng(1).standard_normal((V.shape[0], 2))
    X, st = _lib.host_factor_solve(V, F2, 0.3, 1.0, B, leaf_size=8)
    Xr = np.column_stack([spsolve((M + 0.3 * L).tocsc(), B[:, j]) for j in range(2)])
    assert np.max(np.abs(X - Xr)) < 1e-11 * np.max(np.abs(Xr))


def test_mesh_prep_matches_oracle(surf_c2):
    from fastcorticalwiring_amd import mesh_prep as mp
    from oracle import analysis as oa

    V, F, M = mp.validate_mesh_inputs(surf_c2.vertices, surf_c2.faces, surf_c2.cortex_mask)
    Vs, Fs, s2o, o2s = mp.build_cortex_submesh(V, F, M)
    Vo, Fo, s2oo, o2so = oa.build_cortex_submesh(V, F, M)
    assert np.array_equal(Vs, Vo) and np.array_equal(Fs, Fo) and np.array_equal(s2o, s2oo)
    assert np.array_equal(mp.boundary_indices(Fs), oa.boundary_indices(Fo))
    assert mp.bfs_order_all(Vs.shape[0], Fs).tolist() == oa.bfs_order_all(Vo.shape[0], Fo)
    a1, l1, w1 = mp.face_geometry(Vs, Fs)
    a2, l2, w2 = oa.face_geometry(Vo, Fo)
    assert np.array_equal(a1, a2) and np.array_equal(w1, w2)


def test_prep_helpers_from_adjacency_and_lengths(surf_c2):
    """The analysis takes the mean edge and the bin width from the adjacency and the longest edges
    that the device preparation returns: the same values as from the faces."""
    from fastcorticalwiring_amd import mesh_prep as mp

    V, F, M = mp.validate_mesh_inputs(surf_c2.vertices, surf_c2.faces, surf_c2.cortex_mask)
    Vs, Fs, _, _ = mp.build_cortex_submesh(V, F, M)
    ptr, cols = mp.adjacency_csr(Vs.shape[0], Fs)
    _, L, _ = mp.face_geometry(Vs, Fs)
    assert mp.mean_edge_from_adjacency(Vs, ptr, cols) == mp.mean_edge_length(Vs, Fs)
    assert mp.bin_width_from_lengths(L) == mp.default_bin_width(Vs, Fs)


def test_mesh_input_validation_errors():
    from fastcorticalwiring_amd import mesh_prep as mp

    V = np.zeros((4, 3))
    with pytest.raises(ValueError):
        mp.validate_mesh_inputs(V, np.array([[0, 1, 5]]), np.ones(4, bool))
    with pytest.raises(ValueError):
        mp.validate_mesh_inputs(V, np.array([[0, 1, 2]]), np.zeros(4, bool))
    with pytest.raises(ValueError):
        mp.validate_mesh_inputs(V, np.array([[0, 1, -1]]), np.ones(4, bool))
    with pytest.raises(ValueError):
        mp.validate_mesh_inputs(V, np.array([[0, 1, 2]]), np.ones(3, bool))


@pytest.mark.parametrize("idx,nv,nf", [(0, 2562, 5120), (1, 10242, 20480), (2, 40962, 81920)])
def test_workloads_are_seeded_and_sized(idx, nv, nf):
    from fastcorticalwiring_amd import meshes

    a = meshes.config_surface(idx)
    b = meshes.config_surface(idx)
    assert a.vertices.shape == (nv, 3) and a.faces.shape == (nf, 3)
    assert np.array_equal(a.vertices, b.vertices) and np.array_equal(a.cortex_mask, b.cortex_mask)
    if idx == 0:
        assert a.cortex_mask.all()
        assert np.allclose(np.linalg.norm(a.vertices, axis=1), 50.0)
    else:
        drop = 1.0 - a.cortex_mask.mean()
        assert 0.08 < drop < 0.12
        r = np.linalg.norm(a.vertices, axis=1)
        assert r.min() >= 70 * 0.85 - 1e-9 and r.max() <= 70 * 1.15 + 1e-9


def test_serving_stream_is_seeded():
    from fastcorticalwiring_amd import meshes

    s1 = meshes.serving_stream()
    assert s1 == meshes.serving_stream()
    assert len(s1) == 40 and all(r["level"] in (5, 6, 7) for r in s1)
    assert all(0.05 <= r["cap_frac"] <= 0.20 for r in s1)
    gaps = np.diff([r["arrival_s"] for r in s1])
    assert np.all(gaps >= 0) and 2.0 < gaps.mean() < 9.0


def test_engine_kwargs_rules_raise_before_any_gpu_work(surf_c1):
    from fastcorticalwiring_amd import distance_engines as de

    V, F = surf_c1.vertices, surf_c1.faces
    with pytest.raises(ValueError, match="Unsupported"):
        de.create_distance_engine("mi355x_heat", V, F, {"allow_eigen_fallback": True})
    # the potpourri contract (distance_engines.py:78-89): robust only, diffusion_length_mm >= 0.5
    with pytest.raises(ValueError, match="use_robust=False"):
        de.create_distance_engine("mi355x_potpourri", V, F, {"use_robust": False})
    with pytest.raises(ValueError, match="diffusion_length_mm instead"):
        de.create_distance_engine("mi355x_potpourri", V, F, {"t_coef": 1.0})
    with pytest.raises(ValueError, match="at least 0.5"):
        de.create_distance_engine("mi355x_potpourri", V, F, {"diffusion_length_mm": 0.4})
    with pytest.raises(ValueError, match="either"):
        de.create_distance_engine("mi355x_heat", V, F, {"t_coef": 1.0, "diffusion_length_mm": 0.7})
    with pytest.raises(ValueError):
        de.create_distance_engine("mi355x_heat", V, F, {"t_coef": -1.0})
    with pytest.raises(ValueError):
        de.create_distance_engine("mi355x_heat", V, F, {"diffusion_length_mm": float("nan")})
    with pytest.raises(ValueError, match="precision"):
        de.create_distance_engine("mi355x_heat", V, F, {"precision": "fp16"})
    with pytest.raises(ValueError, match="Unknown engine_type"):
        de.create_distance_engine("potpourri_gpu", V, F)


def test_base_engine_has_no_batch_interface():
    """The reference asserts the ABC exposes no batching (test_batch_heat_engine.py:12-16)."""
    from fastcorticalwiring_amd import distance_engines as de

    assert not hasattr(de.BaseDistanceEngine, "compute_distance_batch")
    assert not hasattr(de.BaseDistanceEngine, "supports_batching")
    assert issubclass(de.MI355XHeatDistanceEngine, de.BaseDistanceEngine)


def test_register_into_foreign_module():
    import types

    from fastcorticalwiring_amd import distance_engines as de

    mod = types.ModuleType("foreign_engines")

    class Base:
        def __init__(self, vertices, faces, **kw):
            self.vertices = np.ascontiguousarray(vertices, dtype=np.float64)
            self.faces = np.ascontiguousarray(faces, dtype=np.int32)
            self.engine_kwargs = dict(kw)

    mod.BaseDistanceEngine = Base
    mod.ENGINE_REGISTRY = {}
    cls = de.register_into(mod)
    assert mod.ENGINE_REGISTRY["mi355x_heat"] is cls and issubclass(cls, Base)
    assert issubclass(mod.ENGINE_REGISTRY["mi355x_potpourri"], Base)


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_host_factorisation_independent_of_thread_count():
    """The tree is factored level by level, wide levels one front per thread: the solution must be
    bit-identical for any OpenMP thread count (each run in its own process)."""
    script = (
        "import hashlib, numpy as np\n"
        "from fastcorticalwiring_amd import _lib, meshes\n"
        "V, F = meshes.icosphere(4)\n"
        "B = np.random.default_rng(2).standard_normal((V.shape[0], 2))\n"
        "X, st = _lib.host_factor_solve(V * 20.0, F.astype(np.int32), 0.5, 1.0, B, leaf_size=16)\n"
        "print(int(st['nnodes']), hashlib.sha256(X.tobytes()).hexdigest())\n")
    outs = []
    for threads in ("1", "3", "8"):
        env = dict(os.environ, OMP_NUM_THREADS=threads, PYTHONPATH=ROOT)
        r = subprocess.run([sys.executable, "-c", script], env=env, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        outs.append(r.stdout.strip())
    assert int(outs[0].split()[0]) > 64  # levels wide enough to spread over the threads
    assert outs[0] == outs[1] == outs[2]


def test_library_load_sets_passive_openmp_waits_unless_set():
    """libgomp reads OMP_WAIT_POLICY when the library loads it: the loader defaults it to PASSIVE
    (spinning waiters slow the factorisation ten-fold when other work shares the cores) and keeps
    a caller's value."""
    script = ("import os\nfrom fastcorticalwiring_amd import _lib\n_lib.lib()\n"
              "print(os.environ.get('OMP_WAIT_POLICY'))\n")
    env = {k: v for k, v in os.environ.items() if k != "OMP_WAIT_POLICY"}
    env["PYTHONPATH"] = ROOT
    r = subprocess.run([sys.executable, "-c", script], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "PASSIVE"
    env["OMP_WAIT_POLICY"] = "ACTIVE"
    r = subprocess.run([sys.executable, "-c", script], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "ACTIVE"


def test_host_factorisation_with_threaded_dissection():
    """Meshes of 20,000 vertices and more build the top two dissection levels on threads: the
    factorisation must still solve the system exactly."""
    from fastcorticalwiring_amd import _lib, meshes
    from oracle import heat as oh

    V, F = meshes.icosphere(6)
    V = V * 50.0
    F = F.astype(np.int32)
    assert V.shape[0] >= 20000
    L, M, *_ = oh.consistent_operators(V, F)
    B = np.random.default_rng(3).standard_normal((V.shape[0], 1))
    X, st = _lib.host_factor_solve(V, F, 0.7, 1.0, B)
    Xr = spsolve((M + 0.7 * L).tocsc(), B[:, 0])
    assert np.max(np.abs(X[:, 0] - Xr)) < 1e-11 * np.max(np.abs(Xr))
    assert st["nnodes"] > 500


def test_threaded_dissection_builds_the_same_tree():
    """The threaded nested dissection (halves of the top dissections built on their own threads)
    gives the same ordering, supernodes and panels as the sequential recursion, on a 163,842-vertex
    icosphere where the default rule threads two levels."""
    from fastcorticalwiring_amd import _lib, meshes

    V, F = meshes.icosphere(7)
    V = V * 70.0
    F = F.astype(np.int32)
    seq = _lib.host_analyse(V, F, par_depth=0)
    for depth in (2, 3, -1):
        par = _lib.host_analyse(V, F, par_depth=depth)
        assert np.array_equal(seq["perm"], par["perm"]), depth
        for key in ("node_start", "node_s", "node_b"):
            assert np.array_equal(seq[key], par[key]), (depth, key)
        assert (seq["nnodes"], seq["nlevels"], seq["factor_nnz"]) == (par["nnodes"], par["nlevels"],
                                                                      par["factor_nnz"])
        assert seq["panel_hash"] == par["panel_hash"], depth
    assert sorted(seq["perm"].tolist()) == list(range(V.shape[0]))
