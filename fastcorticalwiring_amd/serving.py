"""Continuous-batching hemisphere service on one GPU.

Requests (one hemisphere mesh each) arrive over time. A request is admitted as
soon as an active slot and enough of the HBM budget are free; its host work
(submesh, BFS order, analysis + FP64 factorisation inside ``fcw_create``)
runs on a preparation thread while other hemispheres are on the GPU. Every
admitted hemisphere gets its own libfcw context (own streams, factors resident
in HBM) and a worker thread that feeds its sources to the device in chunks of
RHS blocks; with several hemispheres active their kernels run concurrently on
the device. Latency = completion - arrival.

This replaces the reference's subject-level process pool (run_pipeline.py:176-342,
one process per pinned core, one subject per process) for the GPU.
"""

from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import mesh_prep as mp
from . import meshes

__all__ = ["Admission", "HemisphereRequest", "RequestResult", "ServingEngine", "estimate_device_bytes",
           "stream_requests"]


@dataclass
class HemisphereRequest:
    rid: int
    arrival_s: float
    vertices: np.ndarray
    faces: np.ndarray
    cortex_mask: np.ndarray
    scales: tuple = meshes.TWELVE_SCALES
    n_samples_between_scales: int = 3
    boundary_cap_fraction: float | None = None
    meta: dict = field(default_factory=dict)


@dataclass
class RequestResult:
    rid: int
    arrival_s: float
    admitted_s: float = float("nan")
    ready_s: float = float("nan")
    done_s: float = float("nan")
    n_sources: int = 0
    sub_to_orig: np.ndarray | None = None
    outputs: dict | None = None
    error: str | None = None
    device_bytes: int = 0

    @property
    def latency_s(self):
        return self.done_s - self.arrival_s


def estimate_device_bytes(n, m, block=256, streams=2, scales=12, band_cap=6144, factor_nnz_per_vertex=80.0):
    """Upper estimate of one context's device memory, for admission before it is created."""
    ld = ((block + 63) // 64) * 64
    per_ws = 5.0 * n * ld * 8 + 1.0 * n * ld * 8 + ld * scales * 5 * band_cap * 8
    panels = 2.0 * 1.3 * factor_nnz_per_vertex * n * 8
    mesh = m * 8 * 20 + n * 8 * 12
    outputs = n * (12 * 2 + 33 * 2 + 9 + 45) * 8
    return int(streams * per_ws + panels + mesh + outputs)


class Admission:
    """At most max_active hemispheres on the device at once, and a new one only while the sum of the
    active ones' estimated device memory plus its own fits the budget (one is always admitted when
    none is active, so an oversized request still runs, alone)."""

    def __init__(self, max_active, hbm_budget_bytes):
        self.max_active = int(max_active)
        self.hbm_budget = int(hbm_budget_bytes)
        self._lock = threading.Condition()
        self.in_use = 0
        self.active = 0

    def admit(self, need):
        with self._lock:
            while self.active >= self.max_active or (self.active > 0 and self.in_use + need > self.hbm_budget):
                self._lock.wait()
            self.active += 1
            self.in_use += need

    def release(self, need):
        with self._lock:
            self.active -= 1
            self.in_use -= need
            self._lock.notify_all()


class _Prepared:
    def __init__(self, req: HemisphereRequest, block, streams, host_threads, precision, leaf_size=64, tuner=None,
                 tile_rows=0):
        from .engine import HeatContext

        from . import _lib

        V, F, M = mp.validate_mesh_inputs(req.vertices, req.faces, req.cortex_mask)
        d = _lib.mesh_prep_device(V, F, M)  # submesh, geometry, boundary and order on the device
        self.V, self.F, self.s2o, self.o2s = d["vertices"], d["faces"], d["sub_to_orig"], d["orig_to_sub"]
        fl, va = d["face_L"], d["vertex_areas"]
        self.order = d["order"]
        if tuner is not None:  # the tuned block width, leaf size (and tile height) of this mesh's size class
            cfg = tuner.choose(self.V, self.F, self.order)
            block, leaf_size, tile_rows = cfg["block"], cfg["leaf_size"], cfg.get("tile_rows", tile_rows)
        self.ctx = HeatContext(self.V, self.F, block=block, precision=precision, host_threads=host_threads,
                               streams=streams, leaf_size=leaf_size, tile_rows=tile_rows)
        self.ctx_streams = streams
        targets = np.array(sorted(req.scales)) * float(va.sum())
        q = int(req.n_samples_between_scales)
        fracs = np.arange(1, q + 1) / (q + 1.0)
        self.ctx.set_metrics(targets, fracs, d["boundary"], va, health_h=float(np.median(fl)),
                             bin_width=mp.bin_width_from_lengths(fl),
                             boundary_cap_fraction=req.boundary_cap_fraction)


class ServingEngine:
    """Admit, prepare and run hemisphere requests with overlapping device work.

    hbm_budget_bytes : admission budget for the sum of active contexts' estimated device memory
    max_active       : concurrent hemispheres on the device
    chunk_blocks     : RHS blocks a worker submits per call (smaller = finer interleaving)
    streams          : RHS blocks in flight per hemisphere context (the active contexts share the
                       device's hardware queues)
    leaf_size        : nested-dissection leaf size of every context
    tile_rows        : sweep row tile height of every context (0: chosen per tree level)
    tuner            : an ``autotune.Tuner``: block width and leaf size per mesh size class instead
    """

    def __init__(self, hbm_budget_bytes=96 << 30, max_active=3, block=256, chunk_blocks=8, prep_threads=2,
                 host_threads=8, precision="fp64", keep_outputs=True, prepare=None, streams=4, on_done=None,
                 leaf_size=64, tuner=None, tile_rows=0):
        self.hbm_budget = int(hbm_budget_bytes)
        self.max_active = int(max_active)
        self.block = int(block)
        self.chunk_blocks = int(chunk_blocks)
        self.prep_threads = int(prep_threads)
        self.host_threads = int(host_threads)
        self.precision = precision
        self.keep_outputs = keep_outputs
        self.streams = int(streams)
        self.leaf_size = int(leaf_size)
        self.tile_rows = int(tile_rows)
        self.tuner = tuner
        self._prepare = prepare or (lambda req: _Prepared(req, self.block, self.streams, self.host_threads,
                                                           self.precision, self.leaf_size, self.tuner,
                                                           self.tile_rows))
        self._admission = Admission(self.max_active, self.hbm_budget)
        self.trace = []
        self.on_done = on_done  # called with each RequestResult as it completes

    def _admit(self, need):
        self._admission.admit(need)

    def _release(self, need):
        self._admission.release(need)

    def _serve_one(self, req: HemisphereRequest, res: RequestResult, t0: float, prep_sem: threading.Semaphore):
        n_est = int(np.count_nonzero(req.cortex_mask))
        need = estimate_device_bytes(n_est, 2 * n_est, self.block, self.streams)
        res.device_bytes = need
        self._admit(need)
        res.admitted_s = time.perf_counter() - t0
        prep = None
        try:
            with prep_sem:
                prep = self._prepare(req)
            res.ready_s = time.perf_counter() - t0
            order = prep.order
            res.n_sources = int(order.size)
            res.sub_to_orig = prep.s2o
            step = self.chunk_blocks * prep.ctx.block
            parts = []
            for i in range(0, order.size, step):
                parts.append(prep.ctx.wiring(order[i:i + step]))
                self.trace.append((time.perf_counter() - t0, req.rid, i))
            if self.keep_outputs:
                res.outputs = {key: np.concatenate([p[key] for p in parts], axis=0) for key in parts[0]}
                res.outputs["order"] = order
        except Exception as exc:  # the failure is reported per request; the service keeps going
            res.error = f"{type(exc).__name__}: {exc}"
        finally:
            res.done_s = time.perf_counter() - t0
            if prep is not None:
                prep.ctx.close()
            self._release(need)
            if self.on_done is not None:
                self.on_done(res)

    def run(self, requests, time_scale=1.0):
        """Replay requests at their arrival times (scaled); returns results in request order."""
        reqs = sorted(requests, key=lambda r: r.arrival_s)
        results = {r.rid: RequestResult(r.rid, r.arrival_s * time_scale) for r in reqs}
        prep_sem = threading.Semaphore(self.prep_threads)
        threads = []
        t0 = time.perf_counter()
        for r in reqs:
            delay = r.arrival_s * time_scale - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            th = threading.Thread(target=self._serve_one, args=(r, results[r.rid], t0, prep_sem), daemon=True)
            th.start()
            threads.append(th)
        for th in threads:
            th.join()
        self.wall_s = time.perf_counter() - t0
        return [results[r.rid] for r in requests]

    @staticmethod
    def summary(results):
        lat = np.array([r.latency_s for r in results if r.error is None])
        done = [r.done_s for r in results if r.error is None]
        first = min(r.arrival_s for r in results)
        span = max(done) - first if done else float("nan")
        return {
            "requests": len(results),
            "failed": sum(r.error is not None for r in results),
            "p50_s": float(np.percentile(lat, 50)) if lat.size else None,
            "p95_s": float(np.percentile(lat, 95)) if lat.size else None,
            "p99_s": float(np.percentile(lat, 99)) if lat.size else None,
            "mean_s": float(lat.mean()) if lat.size else None,
            "hemispheres_per_hour": 3600.0 * len(done) / span if done else None,
            "sources": int(sum(r.n_sources for r in results)),
            "span_s": span,
        }


def stream_requests(n_requests=40, seed=2024, mean_gap_s=5.0):
    """The serving workload: folded icospheres of level 5/6/7 (30/40/30%), caps 5-20%, own seeds."""
    out = []
    for i, spec in enumerate(meshes.serving_stream(n_requests, mean_gap_s, seed)):
        lmax = 32 if spec["level"] >= 7 else 16
        s = meshes.make_surface(spec["level"], 70.0, 0.15, lmax, 0.3, spec["cap_frac"], spec["seed"])
        out.append(HemisphereRequest(i, spec["arrival_s"], s.vertices, s.faces, s.cortex_mask,
                                     meta=dict(level=spec["level"], cap=spec["cap_frac"])))
    return out
