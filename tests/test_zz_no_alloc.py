"""The allocation-free wiring entry point (fcw_wiring_sources_noalloc) against fcw_wiring_sources.

Both run the same kernels on the same context; only where the call's source positions, sweep
plans and device outputs live differs (a static device array instead of device allocations). Two
calls in a row, the second with fewer sources, check that nothing the first call left in the array
leaks into the second. Tolerances: flags and sample counts equal, every value within 1e-12
relative (the metric kernel's LDS atomics may add in another order from run to run).
"""

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _same(a, b, tag):
    assert np.array_equal(a["scalars"][:, [0, 7, 8]], b["scalars"][:, [0, 7, 8]]), tag
    for key in ("radii", "perimeters", "extra_r", "extra_a", "sample_r", "sample_a"):
        x, y = a[key], b[key]
        if key.startswith("sample"):
            cnt = a["scalars"][:, 7].astype(int)
            for row in range(x.shape[0]):
                assert np.max(rel_err(x[row, :cnt[row]], y[row, :cnt[row]]), initial=0.0) < 1e-6, (tag, key, row)
            continue
        assert np.array_equal(np.isnan(x), np.isnan(y)), (tag, key)
        fin = np.isfinite(y)
        assert np.max(rel_err(x[fin], y[fin]), initial=0.0) < 1e-12, (tag, key)
    x, y = a["scalars"][:, 1:6], b["scalars"][:, 1:6]
    fin = np.isfinite(y)
    assert np.array_equal(fin, np.isfinite(x)), tag
    assert np.max(rel_err(x[fin], y[fin]), initial=0.0) < 1e-12, tag


def test_noalloc_entry_point_matches_the_allocating_one():
    from fastcorticalwiring_amd import meshes
    from fastcorticalwiring_amd.engine import HeatContext
    from oracle import analysis as oa

    s = meshes.config_surface(1)
    mesh = oa.OracleMesh(s.vertices, s.faces, s.cortex_mask)
    T = np.array(meshes.TWELVE_SCALES) * float(mesh.vertex_areas.sum())
    rng = np.random.default_rng(5)
    with HeatContext(mesh.V, mesh.F, block=256, streams=4) as ctx:
        ctx.set_metrics(T, np.array([0.25, 0.5, 0.75]), mesh.boundary, mesh.vertex_areas, health_h=mesh.health_h,
                        bin_width=mesh.width, boundary_cap_fraction=0.1)
        for n_src in (1500, 200):  # six blocks over four streams, then one partial block
            src = rng.choice(mesh.n, n_src, replace=False)
            print("noalloc call with", n_src, "sources", flush=True)
            a = ctx.wiring_noalloc(src)
            b = ctx.wiring(src)
            _same(a, b, n_src)
        # the same sources again, other order: rows follow the sources
        c = ctx.wiring_noalloc(src[::-1])
        _same({k: v[::-1] for k, v in c.items()}, b, "reversed")


def test_noalloc_entry_point_rejects_what_does_not_fit():
    from fastcorticalwiring_amd import meshes
    from fastcorticalwiring_amd._lib import FcwError
    from fastcorticalwiring_amd.engine import HeatContext
    from oracle import analysis as oa

    s = meshes.config_surface(0)
    mesh = oa.OracleMesh(s.vertices, s.faces, s.cortex_mask)
    T = np.array(meshes.TWELVE_SCALES) * float(mesh.vertex_areas.sum())
    with HeatContext(mesh.V, mesh.F, block=256, streams=2) as ctx:
        with pytest.raises(FcwError, match="set_metrics"):
            ctx._metrics = {"S": 12, "q": 3}
            ctx.wiring_noalloc(np.arange(4))
        ctx.set_metrics(T, np.array([0.25, 0.5, 0.75]), mesh.boundary, mesh.vertex_areas, health_h=mesh.health_h,
                        bin_width=mesh.width)
        with pytest.raises(FcwError, match="at most 4096 sources"):
            ctx.wiring_noalloc(np.arange(4097) % mesh.n)
        ctx.set_profiling(True)
        with pytest.raises(FcwError, match="profiling"):
            ctx.wiring_noalloc(np.arange(4))
        ctx.set_profiling(False)
        # 32 scales with 8 samples between them: about 7 KB of outputs per source, so 4096 sources
        # exceed the 16 MiB array and the call says so
        T32 = np.geomspace(0.001, 0.05, 32) * float(mesh.vertex_areas.sum())
        ctx.set_metrics(T32, np.arange(1, 9) / 9.0, mesh.boundary, mesh.vertex_areas, health_h=mesh.health_h,
                        bin_width=mesh.width)
        with pytest.raises(FcwError, match="bytes of device scratch"):
            ctx.wiring_noalloc(np.arange(4096) % mesh.n)
        out = ctx.wiring_noalloc(np.arange(64))
        assert np.all(out["scalars"][:, 0] == 0) and np.all(np.isfinite(out["radii"]))
