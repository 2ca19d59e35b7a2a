"""Python handle on one libfcw context (one cortical submesh resident on one GPU).

``HeatContext`` owns the C-ABI context: construction builds the operators and
both factorisations and uploads them; the methods run the device pipeline.
There is no CPU path: every compute call goes through libfcw.so and raises if
the library or the GPU is unavailable.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


class HeatContext:
    """Heat-method geodesics + wiring metrics on one mesh (fastcw_lean.py:467-534 counterpart).

    Parameters mirror ``ConsistentBatchHeatEngine(vertices, faces, t_coef, poisson_shift)``
    (fastcw_lean.py:481); ``t`` overrides ``t_coef * mean_edge**2``. ``robust=True`` builds the
    operators on the intrinsic Delaunay triangulation of the mollified surface, as potpourri3d's
    ``use_robust=True`` (distance_engines.py:78-79, 119-125).
    """

    def __init__(self, vertices, faces, t_coef=1.0, t=None, poisson_shift=1e-10, block=256, leaf_size=64,
                 precision="fp64", ir_steps=None, device=0, host_threads=0, streams=8, sparse_rhs=True, robust=False,
                 split_depth=0, band_cap=0, graphs=True, tile_rows=0, index_staging=True):
        V = np.ascontiguousarray(vertices, dtype=np.float64)
        F = np.ascontiguousarray(faces, dtype=np.int32)
        if V.ndim != 2 or V.shape[1] != 3:
            raise ValueError(f"Invalid vertex array shape {V.shape}; expected (N, 3).")
        if F.ndim != 2 or F.shape[1] != 3:
            raise ValueError(f"Invalid face array shape {F.shape}; expected (M, 3).")
        L = _lib.lib()
        opt = _lib.Options()
        L.fcw_default_options(ctypes.byref(opt))
        opt.t = float(t) if t is not None else 0.0
        opt.t_coef = float(t_coef)
        opt.poisson_shift = float(poisson_shift)
        opt.precision = {"fp64": _lib.FCW_FP64, "fp32_ir": _lib.FCW_FP32_IR, "bf16x3_ir": _lib.FCW_BF16X3_IR}[precision]
        opt.block = int(block)
        opt.leaf_size = int(leaf_size)
        opt.host_threads = int(host_threads)
        opt.device = int(device)
        opt.ir_steps = -1 if ir_steps is None else int(ir_steps)
        opt.streams = int(streams)
        opt.sparse_rhs = 1 if sparse_rhs else 0
        opt.robust = 1 if robust else 0
        opt.split_depth = int(split_depth)
        if band_cap:
            opt.band_cap = int(band_cap)
        opt.graphs = 1 if graphs else 0
        opt.tile_rows = int(tile_rows)
        opt.index_staging = 1 if index_staging else 0
        handle = ctypes.c_void_p()
        _lib.check(L.fcw_create(_lib.dptr(V), V.shape[0], _lib.iptr(F), F.shape[0], ctypes.byref(opt),
                                ctypes.byref(handle)))
        self._h = handle
        self.n = V.shape[0]
        self.m = F.shape[0]
        info = _lib.Info()
        _lib.check(L.fcw_get_info(self._h, ctypes.byref(info)))
        self.info = info.as_dict()
        self.block = int(self.info["block"])
        self.t = float(self.info["t"])
        self.mean_edge = float(self.info["mean_edge"])
        self._metrics = None

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().fcw_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    # -- distances -------------------------------------------------------------
    def distance_block(self, sources):
        """(k, n) float64 distances, row per source, shifted to 0 at the source (k <= block)."""
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.int64).reshape(-1), dtype=np.int32)
        if src.size == 0:
            return np.empty((0, self.n))
        out = np.empty((src.size, self.n), dtype=np.float64)
        _lib.check(_lib.lib().fcw_distance_block(self._h, _lib.iptr(src), src.size, _lib.dptr(out)))
        return out

    def divergence(self, u):
        """Fused gradient -> normalise -> divergence of the (n, k) fields u (k <= block), as
        fused_normalized_divergence (fastcw_lean.py:416-442): (n, k) float64."""
        U = np.ascontiguousarray(u, dtype=np.float64)
        if U.ndim == 1:
            U = U[:, None]
        if U.shape[0] != self.n or not 1 <= U.shape[1] <= self.block:
            raise ValueError(f"fields must be (n, k) with n = {self.n} and 1 <= k <= {self.block}")
        out = np.empty_like(U)
        _lib.check(_lib.lib().fcw_divergence_block(self._h, _lib.dptr(U), U.shape[1], _lib.dptr(out)))
        return out

    def distances(self, sources):
        src = np.asarray(sources, dtype=np.int64).reshape(-1)
        parts = [self.distance_block(src[i:i + self.block]) for i in range(0, src.size, self.block)]
        return np.concatenate(parts, axis=0) if parts else np.empty((0, self.n))

    # -- metrics ---------------------------------------------------------------
    def set_metrics(self, targets, extra_fracs, boundary, vertex_areas, eps=1e-6, health_h=None, bin_width=None,
                    boundary_cap_fraction=None, rtol=1e-10, max_iter=60):
        T = np.ascontiguousarray(targets, dtype=np.float64)
        fr = np.ascontiguousarray(extra_fracs, dtype=np.float64)
        bnd = np.ascontiguousarray(boundary, dtype=np.int32)
        w = np.ascontiguousarray(vertex_areas, dtype=np.float64)
        if w.shape[0] != self.n:
            raise ValueError("vertex_areas must have one entry per vertex")
        if health_h is None or bin_width is None:
            raise ValueError("health_h and bin_width are required")
        cap = -1.0 if boundary_cap_fraction is None else float(boundary_cap_fraction)
        _lib.check(_lib.lib().fcw_set_metrics(
            self._h, _lib.dptr(T), T.size, _lib.dptr(fr) if fr.size else None, fr.size,
            _lib.iptr(bnd) if bnd.size else None, bnd.size, _lib.dptr(w), float(eps), float(health_h),
            float(bin_width), cap, float(rtol), int(max_iter)))
        self._metrics = dict(S=T.size, q=fr.size)

    def _alloc_out(self, k):
        S, q = self._metrics["S"], self._metrics["q"]
        ne = max(0, S - 1) * q
        arrs = dict(
            scalars=np.empty((k, _lib.NSCALARS)), radii=np.empty((k, S)), perimeters=np.empty((k, S)),
            extra_r=np.empty((k, max(ne, 0))), extra_a=np.empty((k, max(ne, 0))),
            sample_r=np.empty((k, S + ne), np.float32), sample_a=np.empty((k, S + ne), np.float32))
        mo = _lib.MetricsOut()
        for key in ("scalars", "radii", "perimeters", "extra_r", "extra_a"):
            setattr(mo, key, _lib.dptr(arrs[key]) if arrs[key].size else None)
        mo.sample_r = _lib.fptr(arrs["sample_r"])
        mo.sample_a = _lib.fptr(arrs["sample_a"])
        return arrs, mo

    def wiring(self, sources):
        """Metrics for every source (all blocks on the device, one copy back)."""
        if self._metrics is None:
            raise RuntimeError("set_metrics() must be called first")
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.int64).reshape(-1), dtype=np.int32)
        arrs, mo = self._alloc_out(src.size)
        if src.size:
            _lib.check(_lib.lib().fcw_wiring_sources(self._h, _lib.iptr(src), src.size, ctypes.byref(mo)))
        return arrs

    def wiring_noalloc(self, sources):
        """As wiring(), through the entry point that allocates no device memory while it runs
        (fcw_wiring_sources_noalloc: at most 4096 sources per call)."""
        if self._metrics is None:
            raise RuntimeError("set_metrics() must be called first")
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.int64).reshape(-1), dtype=np.int32)
        arrs, mo = self._alloc_out(src.size)
        if src.size:
            _lib.check(_lib.lib().fcw_wiring_sources_noalloc(self._h, _lib.iptr(src), src.size, ctypes.byref(mo)))
        return arrs

    def geometry_fields(self, D, sources=None, health_screen=False):
        """Metrics of caller-supplied fields (k, n); the health screen only when asked."""
        if self._metrics is None:
            raise RuntimeError("set_metrics() must be called first")
        D = np.ascontiguousarray(np.atleast_2d(D), dtype=np.float64)
        k = D.shape[0]
        src = None if sources is None else np.ascontiguousarray(np.asarray(sources).reshape(-1), dtype=np.int32)
        arrs, mo = self._alloc_out(k)
        _lib.check(_lib.lib().fcw_geometry_fields(self._h, _lib.dptr(D), _lib.iptr(src), k, 1 if health_screen else 0,
                                                  ctypes.byref(mo)))
        return arrs

    # -- timing ----------------------------------------------------------------
    def set_profiling(self, on=True):
        _lib.check(_lib.lib().fcw_set_profiling(self._h, 1 if on else 0))

    def stage_times(self):
        out = np.zeros(6)
        _lib.lib().fcw_stage_times(self._h, _lib.dptr(out), 6)
        return dict(heat_solve=out[0], fused_div=out[1], poisson_solve=out[2], metrics=out[3], wall=out[4],
                    issue=out[5])

    KERNEL_KINDS = ("fwd_gemm", "bwd_gemm", "metrics", "fwd_gemm32", "bwd_gemm32", "divergence", "chunk_stats",
                    "residual")

    def kernel_times(self):
        """{kind: (total ms, launches)} since profiling was switched on."""
        k = len(self.KERNEL_KINDS)
        out = np.zeros(2 * k)
        _lib.lib().fcw_kernel_times(self._h, _lib.dptr(out), 2 * k)
        return {name: (out[2 * i], int(out[2 * i + 1])) for i, name in enumerate(self.KERNEL_KINDS)}

    def kernel_log(self):
        """Structured array of every timed launch: kind, system, level, wgs, entries, ms."""
        L = _lib.lib()
        n = L.fcw_kernel_log(self._h, None, 0)
        buf = np.zeros((max(n, 1), 6))
        n = L.fcw_kernel_log(self._h, _lib.dptr(buf), n)
        rec = np.zeros(n, dtype=[("kind", "U12"), ("system", "i4"), ("level", "i4"), ("wgs", "i8"),
                                 ("entries", "f8"), ("ms", "f8")])
        for i in range(n):
            rec[i] = (self.KERNEL_KINDS[int(buf[i, 0])], int(buf[i, 1]), int(buf[i, 2]), int(buf[i, 3]), buf[i, 4],
                      buf[i, 5])
        return rec
