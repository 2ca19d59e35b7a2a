"""GPU parity: libfcw.so (HIP, gfx950) against the CPU FP64 oracle on the same inputs.

Tolerances: the FP64 device path is compared to the SuperLU FP64 oracle at 1e-9
relative on distances (both are direct FP64 solves; only summation order and
the factorisation differ) and at 1e-5 relative on every metric (the target
stated for the hot path); the metric kernels on identical input fields are
compared at 1e-9.
"""

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TWELVE = None


def _scales():
    from fastcorticalwiring_amd import meshes

    return meshes.TWELVE_SCALES


@pytest.fixture(scope="module")
def c1():
    from fastcorticalwiring_amd.engine import HeatContext
    from fastcorticalwiring_amd import meshes
    from oracle import analysis as oa
    from oracle import heat as oh

    s = meshes.config_surface(0)
    mesh = oa.OracleMesh(s.vertices, s.faces, s.cortex_mask)
    ctx = HeatContext(mesh.V, mesh.F, block=256)
    H = oh.HeatOracle(mesh.V, mesh.F)
    yield s, mesh, ctx, H
    ctx.close()


@pytest.fixture(scope="module")
def c2():
    from fastcorticalwiring_amd.engine import HeatContext
    from fastcorticalwiring_amd import meshes
    from oracle import analysis as oa
    from oracle import heat as oh

    s = meshes.config_surface(1)
    mesh = oa.OracleMesh(s.vertices, s.faces, s.cortex_mask)
    ctx = HeatContext(mesh.V, mesh.F, block=128)
    H = oh.HeatOracle(mesh.V, mesh.F)
    yield s, mesh, ctx, H
    ctx.close()


def _set_metrics(ctx, mesh, scales, q=3, cap=None):
    total = float(np.sum(mesh.vertex_areas))
    T = np.array(sorted(scales)) * total
    fr = np.arange(1, q + 1) / (q + 1.0)
    ctx.set_metrics(T, fr, mesh.boundary, mesh.vertex_areas, eps=1e-6, health_h=mesh.health_h,
                    bin_width=mesh.width, boundary_cap_fraction=cap)
    return T, fr


def test_c1_all_sources_distances(c1):
    s, mesh, ctx, H = c1
    print("c1: all", mesh.n, "sources", flush=True)
    src = np.arange(mesh.n)
    D = ctx.distances(src)
    Dref = H.compute_distance_batch(src)
    assert D.shape == Dref.shape
    assert np.all(D[np.arange(mesh.n), src] == 0.0)
    err = np.max(np.abs(D - Dref)) / np.max(Dref)
    far = Dref > 5 * mesh.health_h
    rel = np.max(rel_err(D[far], Dref[far]))
    print(f"c1 max abs/max {err:.2e} rel(d>5h) {rel:.2e}")
    assert err < 1e-9 and rel < 1e-9


def test_c2_distances_ragged_blocks(c2):
    s, mesh, ctx, H = c2
    rng = np.random.default_rng(7)
    for k in (1, 63, 128, 200):
        src = rng.choice(mesh.n, k, replace=False)
        print("c2: k =", k, flush=True)
        D = ctx.distances(src)
        Dref = H.compute_distance_batch(src)
        far = Dref > 5 * mesh.health_h
        assert np.max(np.abs(D - Dref)) / np.max(Dref) < 1e-9
        assert np.max(rel_err(D[far], Dref[far])) < 1e-9


def test_block_composition_does_not_change_results(c2):
    """Distances are bit-identical whatever else shares the RHS block (no atomics in the solve)."""
    s, mesh, ctx, H = c2
    a = ctx.distance_block([5, 17, 300])
    b = ctx.distance_block([300] + list(range(1000, 1100)) + [17, 5])
    assert np.array_equal(a[0], b[-1]) and np.array_equal(a[1], b[-2]) and np.array_equal(a[2], b[0])


def test_streams_and_sparse_rhs_give_identical_distances(c1):
    from fastcorticalwiring_amd.engine import HeatContext

    s, mesh, ctx, H = c1
    src = np.arange(0, mesh.n, 5)[:300]
    ref = ctx.distances(src)
    for streams, sparse in ((1, False), (1, True), (3, True)):
        with HeatContext(mesh.V, mesh.F, block=128, streams=streams, sparse_rhs=sparse) as c:
            D = c.distances(src)
        assert np.max(np.abs(D - ref)) <= 1e-12 * np.max(ref), (streams, sparse)


def test_split_depth_sweeps_match_unsplit(c2):
    """The narrow levels' depth split over several workgroups (partial products added in a fixed
    order by a second launch) changes only rounding: distances agree with the unsplit sweeps to
    1e-13 relative, are reproducible run to run and do not depend on the block's composition, and
    the wiring metrics agree to 1e-9."""
    from fastcorticalwiring_amd.engine import HeatContext

    s, mesh, ctx, H = c2
    src = np.arange(0, mesh.n, 11)[:256]
    with HeatContext(mesh.V, mesh.F, block=128, streams=1, split_depth=-1) as off, \
            HeatContext(mesh.V, mesh.F, block=128, streams=1, split_depth=1) as on:
        assert off.info["split_levels"] == 0 and on.info["split_levels"] > 0
        print("split launches per block:", on.info["split_levels"], flush=True)
        D0 = off.distances(src)
        D1 = on.distances(src)
        assert np.max(np.abs(D1 - D0)) <= 1e-13 * np.max(D0)
        assert np.array_equal(on.distances(src), D1)
        # the same sources in another block composition
        D2 = on.distances(np.concatenate([src[64:128], src[:64]]))
        assert np.array_equal(np.concatenate([D2[64:], D2[:64]]), D1[:128])
        for c in (off, on):
            _set_metrics(c, mesh, _scales())
        w0, w1 = off.wiring(src), on.wiring(src)
        ok = w0["scalars"][:, 0] == 0
        assert np.array_equal(w0["scalars"][:, 0], w1["scalars"][:, 0]) and ok.sum() > 100
        assert np.max(rel_err(w1["radii"][ok], w0["radii"][ok])) < 1e-9
        assert np.max(rel_err(w1["perimeters"][ok], w0["perimeters"][ok])) < 1e-9


def test_fp32_refined_poisson_meets_the_target(c2):
    """FCW_FP32_IR: Poisson system in FP32 sweeps refined against the FP64 residual; distances
    within 1e-5 relative of the FP64 oracle (d > 5h, as compare_fastcw_lean.py:235-239), and
    closer with more refinement steps."""
    from fastcorticalwiring_amd.engine import HeatContext

    s, mesh, ctx, H = c2
    src = np.arange(3, mesh.n, 41)[:128]
    Dref = H.compute_distance_batch(src)
    far = Dref > 5 * mesh.health_h
    errs = []
    for steps in (0, 1, 2):
        with HeatContext(mesh.V, mesh.F, block=128, precision="fp32_ir", ir_steps=steps) as c:
            D = c.distances(src)
        assert np.all(D[np.arange(src.size), src] == 0.0)
        errs.append(np.max(rel_err(D[far], Dref[far])))
        print("ir_steps", steps, f"rel(d>5h) {errs[-1]:.2e}", flush=True)
    assert errs[1] < 1e-5 and errs[2] < 1e-5
    assert errs[2] <= errs[0]


def test_geometry_fields_euclid_and_non_lipschitz(c2):
    from oracle import geometry as og

    s, mesh, ctx, H = c2
    T, fr = _set_metrics(ctx, mesh, _scales())
    rng = np.random.default_rng(1)
    src = rng.choice(mesh.n, 16, replace=False)
    D = np.linalg.norm(mesh.V[None, :, :] - mesh.V[src][:, None, :], axis=2)
    bad = rng.choice(mesh.n, mesh.n // 50, replace=False)
    D[-1, bad] += rng.uniform(0, 15, bad.size)  # forces many long faces on the last field
    out = ctx.geometry_fields(D, src)
    ref = og.per_source_geometry(D, mesh.F, mesh.lean_area, mesh.g11, mesh.g12, mesh.g22, mesh.width, T, fr)
    for key, rk in (("radii", "radii"), ("perimeters", "perim"), ("extra_r", "extra_r"), ("extra_a", "extra_a")):
        a, b = out[key], ref[rk]
        assert np.array_equal(np.isnan(a), np.isnan(b)), key
        fin = np.isfinite(b)
        e = np.max(rel_err(a[fin], b[fin])) if fin.any() else 0.0
        print(key, f"{e:.2e}")
        assert e < 1e-9, key
    assert np.all(out["scalars"][:, 8] == 0)  # no band overflow


@pytest.mark.parametrize("nscales", [1, 5, 8, 9, 13, 16, 17, 24])
def test_geometry_fields_any_scale_count(c2, nscales):
    """The metric kernel's per-scale Newton runs one wave per scale, assigned by the number of
    scales (at most one per wave, the largest alone and the rest paired, or round robin): every
    assignment against the oracle's per_source_geometry."""
    from fastcorticalwiring_amd.engine import HeatContext
    from oracle import geometry as og

    s, mesh, _, H = c2
    total = float(np.sum(mesh.vertex_areas))
    T = np.geomspace(0.001, 0.05, nscales) * total
    fr = np.array([0.25, 0.5, 0.75])
    rng = np.random.default_rng(5)
    src = rng.choice(mesh.n, 8, replace=False)
    D = np.linalg.norm(mesh.V[None, :, :] - mesh.V[src][:, None, :], axis=2)
    with HeatContext(mesh.V, mesh.F, block=64) as ctx:
        ctx.set_metrics(T, fr, mesh.boundary, mesh.vertex_areas, eps=1e-6, health_h=mesh.health_h,
                        bin_width=mesh.width)
        out = ctx.geometry_fields(D, src)
    ref = og.per_source_geometry(D, mesh.F, mesh.lean_area, mesh.g11, mesh.g12, mesh.g22, mesh.width, T, fr)
    for key, rk in (("radii", "radii"), ("perimeters", "perim"), ("extra_r", "extra_r"), ("extra_a", "extra_a")):
        a, b = out[key], ref[rk]
        assert a.shape == b.shape, key
        assert np.array_equal(np.isnan(a), np.isnan(b)), key
        fin = np.isfinite(b)
        e = np.max(rel_err(a[fin], b[fin])) if fin.any() else 0.0
        assert e < 1e-9, (key, e)


def test_wiring_metrics_match_oracle_loop(c2):
    from oracle import analysis as oa

    s, mesh, ctx, H = c2
    scales = _scales()
    T, fr = _set_metrics(ctx, mesh, scales, cap=0.1)
    rng = np.random.default_rng(3)
    src = np.sort(rng.choice(mesh.n, 150, replace=False))
    res = ctx.wiring(src)
    ref = oa.wiring_costs(mesh, H.compute_distance_batch, scales, boundary_cap_fraction=0.1,
                          sources_sub=src.tolist())
    orig = mesh.sub_to_orig[src]
    sc = res["scalars"]
    assert np.all(sc[:, 0] == ref["field_flag"][orig])
    ok = sc[:, 0] == 0
    for i in range(len(scales)):
        e = np.max(rel_err(res["radii"][ok, i], ref["radius"][i][orig[ok]]))
        p = np.max(rel_err(res["perimeters"][ok, i], ref["perimeter"][i][orig[ok]]))
        assert e < 1e-5 and p < 1e-5, (i, e, p)
    assert np.max(rel_err(sc[ok, 4], ref["msd_unweighted"][orig[ok]])) < 1e-5
    assert np.max(rel_err(sc[ok, 5], ref["msd_weighted"][orig[ok]])) < 1e-5
    assert np.max(rel_err(sc[ok, 3], ref["dist_to_boundary"][orig[ok]])) < 1e-5
    # cleaned samples: same count per vertex, values within float32 rounding + tolerance
    for row, oi in enumerate(orig):
        lo, hi = ref["sample_indptr"][oi], ref["sample_indptr"][oi + 1]
        cnt = int(sc[row, 7])
        assert cnt == hi - lo, (row, cnt, hi - lo)
        assert np.max(rel_err(res["sample_r"][row, :cnt], ref["sample_radii"][lo:hi]), initial=0) < 1e-5
        assert np.max(rel_err(res["sample_a"][row, :cnt], ref["sample_areas"][lo:hi]), initial=0) < 1e-5


def test_health_screen_flags_corrupted_field(c2):
    s, mesh, ctx, H = c2
    _set_metrics(ctx, mesh, _scales())
    src = np.array([10, 20, 30])
    D = H.compute_distance_batch(src)
    D[1, : mesh.n // 2] = -5.0 * mesh.health_h  # a grossly negative field
    out = ctx.geometry_fields(D, src, health_screen=True)
    sc = out["scalars"]
    assert sc[1, 0] == 1 and np.isnan(out["radii"][1]).all() and np.isnan(sc[1, 4])
    assert sc[0, 0] == 0 and sc[2, 0] == 0 and np.isfinite(out["radii"][[0, 2]]).all()
    assert sc[1, 1] > 1e-3


def test_closed_mesh_distance_to_boundary_is_inf(c1):
    s, mesh, ctx, H = c1
    _set_metrics(ctx, mesh, _scales())
    res = ctx.wiring(np.array([0, 1, 2]))
    assert np.all(np.isinf(res["scalars"][:, 3]))


def test_errors_are_raised_not_faulted(c1):
    from fastcorticalwiring_amd._lib import FcwError

    s, mesh, ctx, H = c1
    with pytest.raises(FcwError, match="out of range"):
        ctx.distance_block([mesh.n])
    with pytest.raises(FcwError, match="block"):
        ctx.distance_block(np.arange(ctx.block + 1) % mesh.n)


def test_adapter_contract_and_prefetch(c2):
    from fastcorticalwiring_amd import distance_engines as de
    from fastcorticalwiring_amd import mesh_prep

    s, mesh, ctx, H = c2
    eng = de.create_distance_engine("mi355x_heat", mesh.V, mesh.F, {"block": 64})
    calls = []
    orig_block = eng.context.distance_block

    def counting(block):
        calls.append(len(block))
        return orig_block(block)

    eng.context.distance_block = counting
    order = mesh_prep.bfs_order_all(mesh.n, mesh.F)[:130]
    for sidx in order:
        d = eng.compute_distance(int(sidx))
        assert d.shape == (mesh.n,) and d.dtype == np.float64 and d.flags.c_contiguous
        assert d[int(sidx)] == 0.0 and np.all(np.isfinite(d))
    assert calls == [64, 64, 64]
    ref = H.compute_distance(int(order[100]))
    assert np.max(np.abs(eng.compute_distance(int(order[100])) - ref)) < 1e-9 * np.max(ref)
    with pytest.raises(ValueError):
        eng.compute_distance(mesh.n)


def test_analysis_end_to_end_c1(c1):
    from fastcorticalwiring_amd.analysis import WiringAnalysis
    from oracle import analysis as oa

    s, mesh, ctx, H = c1
    an = WiringAnalysis(s.vertices, s.faces, s.cortex_mask)
    scales = _scales()
    an.compute_all_wiring_costs(scale=scales)
    ref = oa.wiring_costs(mesh, H.compute_distance_batch, scales)
    for i, sc in enumerate(sorted(scales)):
        a = an.radius_function[float(sc)]
        b = ref["radius"][i]
        assert np.max(rel_err(a, b)) < 1e-5
    assert np.max(rel_err(an.msd_weighted, ref["msd_weighted"])) < 1e-5
    assert np.array_equal(an.sample_indptr, ref["sample_indptr"])
    assert np.max(rel_err(an.sample_radii_flat, ref["sample_radii"])) < 1e-5
    assert an.n_flagged_fields == 0


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
def test_graph_replays_match_eager_launches(c2, precision):
    """Full blocks replay the captured heat backward sweep and divergence + Poisson solve: every
    replay, over several calls with eager work (partial blocks, another context's launches) and
    synchronisations between them, gives the eager launches' distances bit for bit and their
    metrics to 1e-12 (the metric kernel's LDS atomics may add in another order)."""
    from fastcorticalwiring_amd.engine import HeatContext

    s, mesh, ctx, H = c2
    src = np.arange(1, mesh.n, 13)[:600]  # two full 256-source blocks and a partial one
    with HeatContext(mesh.V, mesh.F, block=256, streams=2, precision=precision, graphs=False) as eager, \
            HeatContext(mesh.V, mesh.F, block=256, streams=2, precision=precision, graphs=True) as graph:
        D0 = eager.distances(src)
        for c in (eager, graph):
            _set_metrics(c, mesh, _scales())
        w0 = eager.wiring(src)
        for rep in range(3):
            D1 = graph.distances(src)
            assert np.array_equal(D1, D0), rep
            w1 = graph.wiring(src)
            assert np.array_equal(w1["scalars"][:, [0, 7, 8]], w0["scalars"][:, [0, 7, 8]]), rep
            for key in ("radii", "perimeters", "extra_a"):
                fin = np.isfinite(w0[key])
                assert np.array_equal(fin, np.isfinite(w1[key]))
                assert np.max(rel_err(w1[key][fin], w0[key][fin])) < 1e-12, (rep, key)
            ctx.distances(src[:37])  # other eager work on the device between the replays
    print(f"{precision}: 3 rounds of graph replays equal the eager launches", flush=True)


def test_bf16x3_refined_poisson_converges_to_the_oracle(c2):
    """FCW_BF16X3_IR: the Poisson sweeps as three split-BF16 matrix-core products (hi*hi + hi*lo +
    lo*hi, FP32 accumulation), refined against the FP64 residual. Each product keeps about 16
    significant bits, so the unrefined solve is far coarser than FP32's; the refinement steps
    contract the error until the distances meet 1e-5 relative of the FP64 oracle (d > 5h), and
    with enough steps the metrics do too."""
    from fastcorticalwiring_amd.engine import HeatContext

    s, mesh, ctx, H = c2
    src = np.arange(3, mesh.n, 41)[:128]
    Dref = H.compute_distance_batch(src)
    far = Dref > 5 * mesh.health_h
    errs = {}
    for steps in (0, 2, 4, 6):
        with HeatContext(mesh.V, mesh.F, block=128, precision="bf16x3_ir", ir_steps=steps) as c:
            D = c.distances(src)
        assert np.all(D[np.arange(src.size), src] == 0.0)
        errs[steps] = float(np.max(rel_err(D[far], Dref[far])))
        print(f"bf16x3 ir_steps {steps}: rel(d>5h) {errs[steps]:.2e}", flush=True)
    assert errs[6] < 1e-5 and errs[6] <= errs[2] <= errs[0]
