"""Autotuner for the solve configuration of one mesh.

Times the same sample of sources (a contiguous run of the BFS order, identical
for every candidate) through the whole per-source pipeline for each
(RHS block width, nested-dissection leaf size) pair and returns the fastest;
then, for that pair, each sweep row tile height given (``tile_rows``: 0 lets
the library choose per tree level, 16/32/64 forces one height on every level).
The leaf size sets the supernode shapes (leaf panels (s+b) x s with s <= leaf,
and how much update-row traffic the levels above the leaves carry); the block
width sets the panel products' N dimension and the number of sources per
launch. With several streams the sample is long enough to keep every stream
busy (16 blocks of the widest candidate at least), so the rate measured is the
throughput rate, not one block's latency.

``Tuner`` keeps the choice per mesh size (vertex count rounded to a power of
two), which is how the serving loop uses it: the first hemisphere of a size
class pays for the trials, later ones reuse the result.
"""

from __future__ import annotations

import itertools
import math
import threading
import time

import numpy as np

from . import mesh_prep as mp
from . import meshes

DEFAULT_BLOCKS = (128, 256, 512)
DEFAULT_LEAF_SIZES = (32, 64, 128)


def _default_context(V, F, block, leaf_size, precision, streams, tile_rows=0):
    from .engine import HeatContext

    return HeatContext(V, F, block=block, leaf_size=leaf_size, precision=precision, streams=streams,
                       tile_rows=tile_rows)


def autotune(vertices, faces, blocks=DEFAULT_BLOCKS, leaf_sizes=DEFAULT_LEAF_SIZES, n_sources=None, scales=None,
             precision="fp64", streams=8, repeats=2, log=None, make_context=None, order=None, tile_rows=(0,)):
    """Time every (block, leaf size) candidate with the library's own tile heights, then the best
    pair with each other height in ``tile_rows``; returns (best, trials). Each trial records the
    sources/s over the sample, the context's setup seconds and its factor size.

    make_context(V, F, block, leaf_size, precision, streams[, tile_rows=...]) builds the context (the
    device one by default; tile_rows is passed only when it is not 0); the timing only needs its
    ``block``, ``info``, ``set_metrics``, ``wiring`` and ``close``.
    """
    V = np.ascontiguousarray(vertices, dtype=np.float64)
    F = np.ascontiguousarray(faces, dtype=np.int32)
    make_context = make_context or _default_context
    fa, fl, va = mp.face_geometry(V, F)
    if order is None:
        order = mp.bfs_order_all(V.shape[0], F)
    if n_sources is None:
        n_sources = 16 * max(blocks) * max(1, streams) // 8
    src = np.resize(np.asarray(order), max(int(n_sources), 2 * max(blocks)))
    T = np.array(sorted(scales or meshes.TWELVE_SCALES)) * float(va.sum())
    fr = np.array([0.25, 0.5, 0.75])
    bnd = mp.boundary_indices(F)
    hh, w = float(np.median(fl)), mp.default_bin_width(V, F)
    trials = []

    def trial(blk, leaf, tr):
        t0 = time.perf_counter()
        ctx = (make_context(V, F, blk, leaf, precision, streams, tile_rows=tr) if tr
               else make_context(V, F, blk, leaf, precision, streams))
        setup = time.perf_counter() - t0
        try:
            ctx.set_metrics(T, fr, bnd, va, health_h=hh, bin_width=w)
            ctx.wiring(src[: 2 * ctx.block])  # warm-up
            best = float("inf")
            for _ in range(repeats):
                t0 = time.perf_counter()
                ctx.wiring(src)
                best = min(best, time.perf_counter() - t0)
            rate = src.size / best
            trials.append(dict(block=blk, leaf_size=leaf, tile_rows=int(tr), sources_per_s=rate, setup_s=setup,
                               nnodes=int(ctx.info["nnodes"]), factor_nnz=int(ctx.info["factor_nnz"])))
        finally:
            ctx.close()
        if log:
            log(f"autotune: block {blk:4d} leaf {leaf:4d} tile rows {tr or 'chosen':>6}: {rate:10.1f} sources/s "
                f"(setup {setup:.2f}s)")

    for leaf, blk in itertools.product(leaf_sizes, blocks):
        trial(blk, leaf, 0)
    first = max(trials, key=lambda t: t["sources_per_s"])
    for tr in tile_rows:
        if tr:
            trial(first["block"], first["leaf_size"], int(tr))
    best = max(trials, key=lambda t: t["sources_per_s"])
    return best, trials


def size_class(n):
    """Meshes whose vertex counts round to the same power of two share a tuned configuration."""
    return int(round(math.log2(max(int(n), 1))))


class Tuner:
    """Tuned (block, leaf size, tile rows) per mesh size class, computed on first use and cached."""

    def __init__(self, precision="fp64", streams=8, blocks=DEFAULT_BLOCKS, leaf_sizes=DEFAULT_LEAF_SIZES, log=None,
                 make_context=None, n_sources=None, tile_rows=(0,)):
        self.precision = precision
        self.streams = streams
        self.blocks = tuple(blocks)
        self.leaf_sizes = tuple(leaf_sizes)
        self.log = log
        self.make_context = make_context
        self.n_sources = n_sources
        self.tile_rows = tuple(tile_rows)
        self._lock = threading.Lock()
        self.choices = {}
        self.trials = {}

    def choose(self, vertices, faces, order=None):
        key = size_class(np.asarray(vertices).shape[0])
        with self._lock:  # one hemisphere of a class tunes, the others wait for its result
            if key not in self.choices:
                best, trials = autotune(vertices, faces, self.blocks, self.leaf_sizes, self.n_sources,
                                        precision=self.precision, streams=self.streams, log=self.log,
                                        make_context=self.make_context, order=order, tile_rows=self.tile_rows)
                self.choices[key] = dict(block=best["block"], leaf_size=best["leaf_size"])
                if best["tile_rows"]:
                    self.choices[key]["tile_rows"] = best["tile_rows"]
                self.trials[key] = trials
            return dict(self.choices[key])
