"""Whole-hemisphere wiring metrics on the GPU: the counterpart of
``FastCorticalWiringAnalysis`` (core_analysis.py:526-724) and its
``compute_all_wiring_costs`` (core_analysis.py:1441-1672).

The per-source loop body (distance -> health screen -> distance to boundary
-> MSD -> radii / perimeters / samples) runs on the device for RHS blocks of
sources (``HeatContext.wiring``); this class does the once-per-mesh host
preparation and scatters the per-source outputs into the reference's result
layout: float32 arrays over the full vertex space with NaN for non-cortical,
unprocessed and flagged vertices (core_analysis.py:709-724, 1593-1595),
``field_flag``, and the CSR (radius, area) samples (core_analysis.py:1621-1632).
"""

from __future__ import annotations

import time
import warnings

import numpy as np

from . import _lib
from . import mesh_prep as mp
from .engine import HeatContext

GEOMETRY_METHOD = "lean_closed_form_gpu"


class WiringAnalysis:
    DEFAULT_SCALES = (0.0006, 0.001, 0.002, 0.004, 0.008, 0.016)

    def __init__(self, vertices, faces, cortex_mask, engine_kwargs=None, eps=1e-6, metadata=None):
        V, F, M = mp.validate_mesh_inputs(vertices, faces, cortex_mask)
        self.metadata = dict(metadata or {})
        self.eps = float(eps)
        self.engine_kwargs = dict(engine_kwargs or {})
        self.vertices_full = V
        self.faces_full = F
        self.cortex_mask_full = M
        self.n_vertices_full = V.shape[0]
        t0 = time.perf_counter()
        # submesh, geometry, boundary, adjacency and source order on the device (fcw_mesh_prep),
        # bit-identical to mesh_prep's host functions
        d = _lib.mesh_prep_device(V, F, M, start=0, device=int(self.engine_kwargs.get("device", 0)))
        self.vertices, self.faces, self.sub_to_orig, self.orig_to_sub = (d["vertices"], d["faces"], d["sub_to_orig"],
                                                                          d["orig_to_sub"])
        self.n_vertices = self.vertices.shape[0]
        self.n_faces = self.faces.shape[0]
        if self.n_vertices == 0:
            raise ValueError("Cortex mask produced an empty cortical submesh.")
        if self.n_faces == 0:
            raise ValueError("No valid cortical faces remain after applying cortex mask.")
        self.boundary_indices = d["boundary"]
        self.face_areas, self.face_L, self.vertex_areas_sub = d["face_areas"], d["face_L"], d["vertex_areas"]
        self.vertex_areas = np.zeros(self.n_vertices_full)
        self.vertex_areas[self.sub_to_orig] = self.vertex_areas_sub
        self._health_h = float(np.median(self.face_L))
        self._lean_width = mp.bin_width_from_lengths(self.face_L)
        self._bfs_order = d["order"]
        self.mean_edge_length = mp.mean_edge_from_adjacency(self.vertices, d["adj_ptr"], d["adj"])
        kw = dict(self.engine_kwargs)
        t = None
        self.diffusion_length_mm = None
        if "diffusion_length_mm" in kw:
            L = float(kw.pop("diffusion_length_mm"))
            if not np.isfinite(L) or L <= 0.0:
                raise ValueError("diffusion_length_mm must be finite and positive.")
            self.diffusion_length_mm = L
            t = L * L
            self.t_coef = (L / self.mean_edge_length) ** 2
        else:
            self.t_coef = float(kw.pop("t_coef", 1.0))
        self.use_robust = bool(kw.pop("use_robust", False))
        self.prep_seconds = time.perf_counter() - t0
        self.ctx = HeatContext(self.vertices, self.faces, t_coef=self.t_coef, t=t,
                               poisson_shift=float(kw.pop("poisson_shift", 1e-10)), block=int(kw.pop("block", 256)),
                               leaf_size=int(kw.pop("leaf_size", 64)), precision=str(kw.pop("precision", "fp64")),
                               ir_steps=kw.pop("ir_steps", None), device=int(kw.pop("device", 0)),
                               host_threads=int(kw.pop("host_threads", 0)), streams=int(kw.pop("streams", 8)),
                               robust=self.use_robust, split_depth=int(kw.pop("split_depth", 0)),
                               tile_rows=int(kw.pop("tile_rows", 0)))
        if kw:
            raise ValueError(f"Unsupported engine kwargs: {', '.join(sorted(kw))}")
        self.field_flag = np.zeros(self.n_vertices_full, dtype=np.uint8)
        self.n_flagged_fields = 0
        self.active_scales = tuple(self.DEFAULT_SCALES)
        self.msd_unweighted = np.full(self.n_vertices_full, np.nan, np.float32)
        self.msd_weighted = np.full(self.n_vertices_full, np.nan, np.float32)
        self.dist_to_boundary = np.full(self.n_vertices_full, np.nan, np.float32)
        self.radius_function = {}
        self.perimeter_function = {}
        self.sample_radii_flat = None
        self.sample_areas_flat = None
        self.sample_indptr = None
        self.last_run = {}

    normalize_scales = staticmethod(mp.normalize_scales)
    scale_token = staticmethod(mp.scale_token)
    boundary_area_loss_fraction = staticmethod(mp.boundary_area_loss_fraction)

    def close(self):
        """Free the device context (the result arrays stay); further computation raises."""
        self.ctx.close()

    def compute_all_wiring_costs(self, scale=None, vertex_subset=None, n_samples_between_scales=3,
                                 boundary_cap_fraction=None, verbose=False):
        scales = mp.normalize_scales(scale, self.DEFAULT_SCALES)
        m = int(n_samples_between_scales)
        if m < 0:
            raise ValueError("n_samples_between_scales must be >= 0.")
        if boundary_cap_fraction is not None:
            cap = float(boundary_cap_fraction)
            if not np.isfinite(cap) or cap < 0.0 or cap >= 0.5:
                raise ValueError("boundary_cap_fraction must be in [0, 0.5), or None.")
        else:
            cap = None
        nF = self.n_vertices_full
        self.active_scales = scales
        total_area = float(np.sum(self.vertex_areas_sub))
        targets = np.array([total_area * float(s) for s in scales], dtype=np.float64)
        fracs = np.arange(1, m + 1, dtype=np.float64) / (m + 1.0)
        order = self._bfs_order
        if vertex_subset is not None:
            keep = set()
            for raw in vertex_subset:
                try:
                    oi = int(raw)
                except Exception:
                    continue
                if oi < 0 or oi >= nF or not self.cortex_mask_full[oi]:
                    continue
                si = int(self.orig_to_sub[oi])
                if si >= 0:
                    keep.add(si)
            order = np.asarray([v for v in order.tolist() if v in keep], dtype=np.int64)
        self.ctx.set_metrics(targets, fracs, self.boundary_indices, self.vertex_areas_sub, eps=self.eps,
                             health_h=self._health_h, bin_width=self._lean_width, boundary_cap_fraction=cap)
        t0 = time.perf_counter()
        res = self.ctx.wiring(order)
        wall = time.perf_counter() - t0
        sc = res["scalars"]
        orig = self.sub_to_orig[order]
        flag = sc[:, _lib.S_FLAG] > 0.5
        ok = ~flag
        self.msd_unweighted = np.full(nF, np.nan, np.float32)
        self.msd_weighted = np.full(nF, np.nan, np.float32)
        self.dist_to_boundary = np.full(nF, np.nan, np.float32)
        self.field_flag = np.zeros(nF, dtype=np.uint8)
        self.field_flag[orig[flag]] = 1
        self.n_flagged_fields = int(flag.sum())
        self.msd_unweighted[orig[ok]] = sc[ok, _lib.S_MSD_U].astype(np.float32)
        self.msd_weighted[orig[ok]] = sc[ok, _lib.S_MSD_W].astype(np.float32)
        self.dist_to_boundary[orig[ok]] = sc[ok, _lib.S_DIST_BOUNDARY].astype(np.float32)
        self.radius_function = {}
        self.perimeter_function = {}
        for i, s in enumerate(scales):
            r = np.full(nF, np.nan, np.float32)
            p = np.full(nF, np.nan, np.float32)
            r[orig[ok]] = res["radii"][ok, i].astype(np.float32)
            p[orig[ok]] = res["perimeters"][ok, i].astype(np.float32)
            self.radius_function[float(s)] = r
            self.perimeter_function[float(s)] = p
        counts = np.zeros(nF, dtype=np.int64)
        counts[orig[ok]] = sc[ok, _lib.S_N_SAMPLES].astype(np.int64)
        indptr = np.zeros(nF + 1, dtype=np.int64)
        np.cumsum(counts, out=indptr[1:])
        tot = int(indptr[-1])
        sr = np.empty(tot, np.float32)
        sa = np.empty(tot, np.float32)
        if tot:
            rows = np.flatnonzero(ok)
            cnt = sc[rows, _lib.S_N_SAMPLES].astype(np.int64)
            col = np.arange(res["sample_r"].shape[1])[None, :]
            sel = col < cnt[:, None]
            # destination offsets follow the full-vertex CSR order
            dst_rows = orig[rows]
            dst_start = indptr[dst_rows]
            dst = (dst_start[:, None] + col)[sel]
            sr[dst] = res["sample_r"][rows][sel]
            sa[dst] = res["sample_a"][rows][sel]
        self.sample_indptr = indptr
        self.sample_radii_flat = sr
        self.sample_areas_flat = sa
        self.n_samples_between_scales = m
        self.boundary_cap_fraction = cap
        self.last_run = dict(n_sources=int(order.size), wall_seconds=wall, n_flagged=self.n_flagged_fields,
                             n_overflow=int(np.sum(sc[:, _lib.S_OVERFLOW] > 0)),
                             band_evals=float(np.sum(sc[ok, _lib.S_N_EVAL])))
        self.last_scalars = sc
        self.last_order = order
        if self.n_flagged_fields:
            msg = (f"{self.n_flagged_fields} of {order.size} source vertices had corrupted distance fields; "
                   "all metrics for them were set to NaN (see field_flag).")
            warnings.warn(msg, RuntimeWarning)
        if verbose:
            print(f"{order.size} sources in {wall:.3f}s ({order.size / max(wall, 1e-12):.1f} sources/s)")
        return (self.msd_unweighted, self.msd_weighted), self.radius_function, self.perimeter_function

    def provenance(self):
        return {
            "geometry_method": GEOMETRY_METHOD,
            "engine": "mi355x_heat",
            "use_robust": self.use_robust,
            "diffusion_length_mm": self.diffusion_length_mm,
            "t_coef": self.t_coef,
            "mean_edge_length_mm": self.mean_edge_length,
            "n_flagged_fields": int(self.n_flagged_fields),
            "health_neg_frac_tol": 1e-3,
            "health_chord_frac_tol": 1e-3,
            "backend": _lib.build_info(),
        }

    def get_metric_arrays(self):
        out = {"msd_unweighted": self.msd_unweighted, "msd_weighted": self.msd_weighted,
               "dist_to_boundary": self.dist_to_boundary}
        for s in self.active_scales:
            tok = mp.scale_token(s)
            if float(s) in self.radius_function:
                out[f"radius_{tok}"] = self.radius_function[float(s)]
                out[f"perimeter_{tok}"] = self.perimeter_function[float(s)]
        return out

    def get_vertex_samples(self, orig_idx):
        if self.sample_indptr is None:
            return np.empty(0, np.float32), np.empty(0, np.float32)
        lo, hi = int(self.sample_indptr[int(orig_idx)]), int(self.sample_indptr[int(orig_idx) + 1])
        return self.sample_radii_flat[lo:hi], self.sample_areas_flat[lo:hi]
