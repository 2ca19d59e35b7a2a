"""A vertex of degree 80 on the path: a wheel surface (one hub, then rings of 80 vertices around
it). The hub's row meets 81 distinct vertices, more than a divergence row group may stage (64), so
it forms a group of its own whose values are read from global memory: in the fused divergence and
in the refined forms' grouped residual. Every other row of the wheel borders a normal group. Checked
against the oracle: the divergence alone, and distances from the hub, from the ring around it and
from the rim in FP64 and FP32+IR (every field crosses the hub's row)."""

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

K = 80  # hub degree
RINGS = 12


def _wheel():
    """Hub at the origin, ring 1 at radius 1, each further ring one angular step out and rotated by
    half a step (near-equilateral faces away from the hub), gently curved in z; faces oriented
    counterclockwise seen from +z."""
    step = 2.0 * np.pi / K
    V = [[0.0, 0.0, 0.0]]
    for k in range(RINGS):
        r = 1.0 + k * step
        ang = (np.arange(K) + 0.5 * (k % 2)) * step
        V += np.stack([r * np.cos(ang), r * np.sin(ang), 0.05 * (k / RINGS) * np.sin(3.0 * ang)], axis=1).tolist()
    F = [[0, 1 + i, 1 + (i + 1) % K] for i in range(K)]
    for k in range(RINGS - 1):
        a, b = 1 + k * K, 1 + (k + 1) * K
        for i in range(K):
            j = (i + 1) % K
            if k % 2 == 0:  # ring k+1 sits half a step ahead
                F += [[a + i, b + i, a + j], [a + j, b + i, b + j]]
            else:  # ring k+1 sits half a step behind
                F += [[a + i, b + j, a + j], [a + i, b + i, b + j]]
    return np.asarray(V, dtype=np.float64), np.asarray(F, dtype=np.int64)


@pytest.fixture(scope="module")
def wheel():
    from fastcorticalwiring_amd import mesh_prep as mp
    from oracle import analysis as oa
    from oracle import heat as oh

    V, F = _wheel()
    V, F, M = mp.validate_mesh_inputs(V, F, np.ones(V.shape[0], dtype=bool))
    mesh = oa.OracleMesh(V, F, M)
    # the hub keeps its degree in the operator mesh: one row beyond the row-group limits
    assert np.sum(np.any(mesh.F == 0, axis=1)) == K
    return mesh, oh.HeatOracle(mesh.V, mesh.F)


def test_hub_row_divergence_matches_oracle(wheel):
    from fastcorticalwiring_amd.engine import HeatContext
    from oracle import heat as oh

    mesh, _ = wheel
    rng = np.random.default_rng(11)
    X = mesh.V - mesh.V.mean(axis=0)
    cols = []
    for j in range(16):
        c = X[rng.integers(mesh.n)]
        d = np.linalg.norm(X - c, axis=1)
        cols.append(np.exp(-d / (0.3 + rng.random())) if j % 2 == 0 else rng.standard_normal(mesh.n))
    U = np.ascontiguousarray(np.stack(cols, axis=1))
    with HeatContext(mesh.V, mesh.F, block=64) as ctx:
        got = ctx.divergence(U)
    area, gb0, gb1, gb2 = oh.face_gradient_basis(mesh.V, mesh.F)
    ref = oh.fused_normalized_divergence(U, mesh.F, area, gb0, gb1, gb2)
    err = np.max(np.abs(got - ref), axis=0) / np.max(np.abs(ref), axis=0)
    print(f"wheel divergence: max rel {err.max():.2e}, hub row {np.max(np.abs(got[0] - ref[0])):.2e}", flush=True)
    assert err.max() < 1e-12
    assert np.max(np.abs(got[0] - ref[0])) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32_ir", 1e-5)])
def test_hub_distances_match_oracle(wheel, precision, tol):
    from fastcorticalwiring_amd.engine import HeatContext

    mesh, H = wheel
    rim = mesh.n - K
    src = np.array([0, 1, 2, K // 2, 1 + K, 3 * K, rim, rim + K // 3, mesh.n - 1])
    Dref = H.compute_distance_batch(src)
    with HeatContext(mesh.V, mesh.F, block=64, precision=precision) as ctx:
        D = ctx.distances(src)
    assert np.all(D[np.arange(src.size), src] == 0.0)
    far = Dref > 5 * mesh.health_h
    err = np.max(np.abs(D - Dref)) / np.max(Dref)
    rel = np.max(rel_err(D[far], Dref[far]))
    print(f"wheel {precision}: max abs/max {err:.2e} rel(d>5h) {rel:.2e}", flush=True)
    assert err < tol and rel < tol
