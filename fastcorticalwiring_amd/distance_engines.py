"""Geodesic distance-engine adapters, with the MI355X heat-method backend.

Same interface as the reference module (distance_engines.py:54-69, 257-273):
``BaseDistanceEngine(vertices, faces, **engine_kwargs)`` with ``name`` and
``compute_distance(source_idx) -> (n,) float64``, a name -> class
``ENGINE_REGISTRY`` and ``create_distance_engine``. The base class has no batch
methods (the reference's test_batch_heat_engine.py:12-16 asserts that); the
GPU backend batches *behind* the single-source call.

``register_into(module)`` adds the backend to another ``distance_engines``
module's registry (the reference's own, when used as a drop-in there).
"""

from __future__ import annotations

import warnings
from abc import ABC, abstractmethod

import numpy as np

from . import mesh_prep

ENGINE_NAME = "mi355x_heat"
# Same backend behind the reference's production-engine contract (PotpourriDistanceEngine,
# distance_engines.py:72-125): robust operators always, t = diffusion_length_mm^2.
POTPOURRI_NAME = "mi355x_potpourri"
DEFAULT_DIFFUSION_LENGTH_MM = 0.7         # distance_engines.py:12-14
MIN_DIFFUSION_LENGTH_MM = 0.5
RECOMMENDED_MIN_DIFFUSION_LENGTH_MM = 0.7


class BaseDistanceEngine(ABC):
    """Minimal interface expected by the wiring pipeline (distance_engines.py:54-69)."""

    def __init__(self, vertices, faces, **engine_kwargs):
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float64)
        self.faces = np.ascontiguousarray(faces, dtype=np.int32)
        self.engine_kwargs = dict(engine_kwargs or {})

    @property
    @abstractmethod
    def name(self) -> str:
        raise NotImplementedError

    @abstractmethod
    def compute_distance(self, source_idx: int) -> np.ndarray:
        raise NotImplementedError


def _backend_mixin(base, potpourri=False):
    """The heat backend body, parameterised on the base class it must derive from. potpourri:
    the kwargs contract of the reference's PotpourriDistanceEngine (distance_engines.py:75-89)."""

    class _Hip(base):
        """MI355X heat-method distances (fastcw_lean.py:467-534 operators) behind compute_distance.

        engine_kwargs:
          t_coef (float, default 1.0): t = t_coef * mean_edge**2 (fastcw_lean.py:485-489), or
          diffusion_length_mm (float): t = L**2 (the potpourri convention, distance_engines.py:119-124);
          poisson_shift (1e-10), block (256 sources per solve), precision ('fp64' | 'fp32_ir' | 'bf16x3_ir'),
          ir_steps (refinement steps of the refined forms: 2 for 'fp32_ir', 6 for 'bf16x3_ir'), leaf_size (64), device (0), host_threads (0),
          streams (1: this adapter solves one block per miss, so more workspaces would sit idle),
          use_robust (False: operators on the surface's own triangulation, as ConsistentBatchHeatEngine;
          True: on its intrinsic Delaunay triangulation, as potpourri3d's robust mode),
          split_depth (0: narrow tree levels split their depth when streams <= 2, 1 always, -1 never),
          tile_rows (0: sweep row tile heights chosen per tree level; 16, 32 or 64 on every level).
        Sources are solved in RHS blocks: a miss solves the next ``block`` sources of the planned
        order (default: the reference's BFS order, core_analysis.py:1384-1417), so the caller's
        one-source-at-a-time loop reads from the cache.
        """

        _ALLOWED = {"t_coef", "diffusion_length_mm", "poisson_shift", "block", "precision", "leaf_size", "device",
                    "ir_steps", "host_threads", "streams", "use_robust", "split_depth", "tile_rows"} | (
                        {"allow_eigen_fallback"} if potpourri else set())
        _NAME = POTPOURRI_NAME if potpourri else ENGINE_NAME

        def __init__(self, vertices, faces, **engine_kwargs):
            super().__init__(vertices, faces, **engine_kwargs)
            kw = dict(self.engine_kwargs)
            unknown = sorted(set(kw) - self._ALLOWED)
            if unknown:
                raise ValueError(f"Unsupported engine kwargs for {self._NAME} engine: {', '.join(unknown)}")
            if potpourri:
                kw.pop("allow_eigen_fallback", None)  # no Eigen/SuiteSparse backends to choose between
                if "use_robust" in kw and not bool(kw.pop("use_robust")):
                    raise ValueError("use_robust=False is not supported; robust mode is always used.")
                if "t_coef" in kw:
                    raise ValueError("Set diffusion_length_mm instead of t_coef.")
                L = float(kw.pop("diffusion_length_mm", DEFAULT_DIFFUSION_LENGTH_MM))
                if not np.isfinite(L) or L < MIN_DIFFUSION_LENGTH_MM:
                    raise ValueError(f"diffusion_length_mm must be finite and at least {MIN_DIFFUSION_LENGTH_MM} mm.")
                if L < RECOMMENDED_MIN_DIFFUSION_LENGTH_MM:
                    warnings.warn(f"diffusion_length_mm={L} is below the validated {RECOMMENDED_MIN_DIFFUSION_LENGTH_MM} mm.",
                                  RuntimeWarning)
                kw["diffusion_length_mm"] = L
                robust = True
            else:
                robust = bool(kw.pop("use_robust", False))
            if "t_coef" in kw and "diffusion_length_mm" in kw:
                raise ValueError("Set either t_coef or diffusion_length_mm, not both.")
            self.mean_edge_length = mesh_prep.mean_edge_length(self.vertices, self.faces)
            t = None
            if "diffusion_length_mm" in kw:
                L = float(kw.pop("diffusion_length_mm"))
                if not np.isfinite(L) or L <= 0.0:
                    raise ValueError("diffusion_length_mm must be finite and positive.")
                self.diffusion_length_mm = L
                self.t_coef = (L / self.mean_edge_length) ** 2
                t = L * L
            else:
                tc = float(kw.pop("t_coef", 1.0))
                if not np.isfinite(tc) or tc <= 0.0:
                    raise ValueError("t_coef must be finite and positive.")
                self.t_coef = tc
                self.diffusion_length_mm = None
            shift = float(kw.pop("poisson_shift", 1e-10))
            if not np.isfinite(shift) or shift <= 0.0:
                raise ValueError("poisson_shift must be finite and positive.")
            precision = str(kw.pop("precision", "fp64")).lower()
            if precision not in ("fp64", "fp32_ir", "bf16x3_ir"):
                raise ValueError("precision must be 'fp64', 'fp32_ir' or 'bf16x3_ir'.")
            self.use_robust = robust
            try:
                from . import _lib
                from .engine import HeatContext

                _lib.lib()
            except ImportError as exc:
                raise ImportError(f"libfcw.so is required for the '{self._NAME}' engine.") from exc
            info = _lib.build_info()
            if "gfx950" not in info:
                raise RuntimeError(f"{self._NAME} backend self-check failed: {info!r}")
            self.backend_info = info
            self._ctx = HeatContext(self.vertices, self.faces, t_coef=self.t_coef, t=t, poisson_shift=shift,
                                    block=int(kw.pop("block", 256)), leaf_size=int(kw.pop("leaf_size", 64)),
                                    precision=precision, ir_steps=kw.pop("ir_steps", None),
                                    device=int(kw.pop("device", 0)), host_threads=int(kw.pop("host_threads", 0)),
                                    streams=int(kw.pop("streams", 1)), robust=robust,
                                    split_depth=int(kw.pop("split_depth", 0)), tile_rows=int(kw.pop("tile_rows", 0)))
            self.t = self._ctx.t
            self._plan = None
            self._plan_pos = None
            self._cache = {}

        @property
        def name(self) -> str:
            return self._NAME

        @property
        def context(self):
            return self._ctx

        def plan_sources(self, order):
            """Declare the order in which sources will be requested (enables exact prefetch)."""
            order = np.asarray(order, dtype=np.int64).reshape(-1)
            self._plan = order
            self._plan_pos = {int(v): i for i, v in enumerate(order.tolist())}
            self._cache = {}

        def _ensure_plan(self):
            if self._plan is None:
                self.plan_sources(mesh_prep.bfs_order_all(self.vertices.shape[0], self.faces))

        def compute_distance(self, source_idx: int) -> np.ndarray:
            src = int(source_idx)
            n = self.vertices.shape[0]
            if src < 0 or src >= n:
                raise ValueError(f"source_idx out of range: {src} for mesh with {n} vertices")
            hit = self._cache.get(src)
            if hit is None:
                self._ensure_plan()
                i = self._plan_pos.get(src)
                if i is None:
                    block = [src]
                else:
                    block = self._plan[i:i + self._ctx.block].tolist()
                D = self._ctx.distance_block(block)
                self._cache = {int(s): D[j] for j, s in enumerate(block)}
                hit = self._cache[src]
            return np.ascontiguousarray(hit, dtype=np.float64)

    return _Hip


MI355XHeatDistanceEngine = _backend_mixin(BaseDistanceEngine)
MI355XHeatDistanceEngine.__name__ = "MI355XHeatDistanceEngine"
MI355XHeatDistanceEngine.__qualname__ = "MI355XHeatDistanceEngine"
MI355XPotpourriDistanceEngine = _backend_mixin(BaseDistanceEngine, potpourri=True)
MI355XPotpourriDistanceEngine.__name__ = "MI355XPotpourriDistanceEngine"
MI355XPotpourriDistanceEngine.__qualname__ = "MI355XPotpourriDistanceEngine"

ENGINE_REGISTRY = {
    ENGINE_NAME: MI355XHeatDistanceEngine,
    POTPOURRI_NAME: MI355XPotpourriDistanceEngine,
}


def create_distance_engine(engine_type, vertices, faces, engine_kwargs=None):
    """Construct a distance engine by name (distance_engines.py:265-273)."""
    key = str(engine_type).strip().lower()
    if key not in ENGINE_REGISTRY:
        valid = ", ".join(sorted(ENGINE_REGISTRY.keys()))
        raise ValueError(f"Unknown engine_type '{engine_type}'. Valid options: {valid}")
    cls = ENGINE_REGISTRY[key]
    return cls(vertices, faces, **dict(engine_kwargs or {}))


def register_into(module, name=ENGINE_NAME, potpourri_name=POTPOURRI_NAME):
    """Register the backend in another distance_engines module, subclassing *its*
    BaseDistanceEngine: ``name`` with this module's kwargs, ``potpourri_name`` (unless None) with
    the reference's potpourri-engine contract. Returns the ``name`` class."""
    base = getattr(module, "BaseDistanceEngine")
    cls = _backend_mixin(base)
    cls.__name__ = cls.__qualname__ = "MI355XHeatDistanceEngine"
    module.ENGINE_REGISTRY[name] = cls
    if potpourri_name:
        pcls = _backend_mixin(base, potpourri=True)
        pcls.__name__ = pcls.__qualname__ = "MI355XPotpourriDistanceEngine"
        module.ENGINE_REGISTRY[potpourri_name] = pcls
    if not issubclass(cls, base):  # pragma: no cover - defensive
        warnings.warn("registered backend does not derive from the module's BaseDistanceEngine")
    return cls
