"""The fused gradient -> normalise -> divergence alone (fcw_divergence_block) against the oracle's
restatement of fused_normalized_divergence (fastcw_lean.py:416-442), on fields that take every path
of the kernel: smooth fields, faces whose gradient vanishes exactly (no term), and row groups whose
values span 130 decades (|grad u|^2 of the values scaled to the group's largest falls below 2^-767,
where the estimate of 1/|grad u| is taken on rescaled values). Gradients stay far above 1e-150 so
that the oracle's plain squared norm is exact too."""

import numpy as np
import pytest

from fastcorticalwiring_amd import mesh_prep as mp
from fastcorticalwiring_amd.engine import HeatContext
from oracle import heat as oh

pytestmark = pytest.mark.gpu


def _fields(V, k, seed):
    rng = np.random.default_rng(seed)
    n = V.shape[0]
    X = V - V.mean(axis=0)
    cols = []
    for j in range(k):
        kind = j % 4
        c = X[rng.integers(n)]
        d = np.linalg.norm(X - c, axis=1)
        if kind == 0:  # smooth, decaying over the whole surface
            u = np.exp(-d / (5.0 + 10.0 * rng.random()))
        elif kind == 1:  # clamped: flat patches where every face has grad u = 0 exactly
            u = np.clip(np.cos(d / 7.0), -0.3, 0.4)
        elif kind == 2:  # huge vertices beside values of 1e-30: 130 decades inside their row groups
            u = 1e-30 * (1.0 + X @ rng.standard_normal(3))
            u[rng.integers(n, size=3)] = 1e100
        else:  # rough
            u = rng.standard_normal(n)
        cols.append(u)
    return np.ascontiguousarray(np.stack(cols, axis=1))


@pytest.mark.parametrize("workload", [0, 1])
def test_divergence_matches_oracle_on_every_path(workload, surf_c1, surf_c2):
    surf = surf_c1 if workload == 0 else surf_c2
    V, F, M = mp.validate_mesh_inputs(surf.vertices, surf.faces, surf.cortex_mask)
    Vs, Fs, _, _ = mp.build_cortex_submesh(V, F, M)
    ctx = HeatContext(Vs, Fs, block=64)
    try:
        U = _fields(Vs, 64, seed=7 + workload)
        got = ctx.divergence(U)
    finally:
        ctx.close()
    area, gb0, gb1, gb2 = oh.face_gradient_basis(Vs, Fs)
    ref = oh.fused_normalized_divergence(U, Fs, area, gb0, gb1, gb2)
    scale = np.max(np.abs(ref), axis=0)
    err = np.max(np.abs(got - ref), axis=0) / scale
    assert np.all(np.isfinite(got))
    assert err.max() < 1e-12, err.max()
    # flat patches: some faces contribute nothing, and rows inside a patch are exactly zero
    flat = U[:, 1]
    inside = np.all(np.isin(Fs, np.where(flat == flat.max())[0]), axis=1)
    assert inside.any()
    rows = np.setdiff1d(np.unique(Fs[inside]), np.unique(Fs[~inside]))
    if rows.size:
        assert np.all(got[rows, 1] == 0.0)
    # the divergence of any face field sums to zero over the vertices (sum of the basis gradients)
    tot = np.abs(got.sum(axis=0)) / np.abs(got).sum(axis=0)
    assert tot.max() < 1e-12, tot.max()


def test_divergence_argument_errors(surf_c1):
    V, F, M = mp.validate_mesh_inputs(surf_c1.vertices, surf_c1.faces, surf_c1.cortex_mask)
    Vs, Fs, _, _ = mp.build_cortex_submesh(V, F, M)
    ctx = HeatContext(Vs, Fs, block=64)
    try:
        with pytest.raises(ValueError):
            ctx.divergence(np.zeros((Vs.shape[0], 65)))
        with pytest.raises(ValueError):
            ctx.divergence(np.zeros((Vs.shape[0] + 1, 4)))
        # a constant field has no gradient anywhere: no terms at all
        assert np.all(ctx.divergence(np.full((Vs.shape[0], 3), 2.5)) == 0.0)
    finally:
        ctx.close()
