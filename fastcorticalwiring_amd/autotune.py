"""Autotuner for the solve configuration of one mesh.

Times a fixed sample of sources (the first ``n_sources`` of the BFS order,
identical for every candidate) through the full per-source pipeline for each
(RHS block width, nested-dissection leaf size) pair and returns the fastest.
The leaf size sets the supernode shapes (leaf panel (s+b) x s with s <= leaf);
the block width sets the panel-product N dimension and the number of sources
per launch.
"""

from __future__ import annotations

import itertools
import time

import numpy as np

from . import mesh_prep as mp
from . import meshes
from .engine import HeatContext


def autotune(vertices, faces, blocks=(128, 256, 512), leaf_sizes=(32, 64, 128), n_sources=2048, scales=None,
             precision="fp64", repeats=2, log=None):
    V = np.ascontiguousarray(vertices, dtype=np.float64)
    F = np.ascontiguousarray(faces, dtype=np.int32)
    fa, fl, va = mp.face_geometry(V, F)
    order = mp.bfs_order_all(V.shape[0], F)
    src = np.resize(order, max(n_sources, max(blocks)))
    T = np.array(sorted(scales or meshes.TWELVE_SCALES)) * float(va.sum())
    fr = np.array([0.25, 0.5, 0.75])
    bnd = mp.boundary_indices(F)
    hh, w = float(np.median(fl)), mp.default_bin_width(V, F)
    trials = []
    for leaf, blk in itertools.product(leaf_sizes, blocks):
        t0 = time.perf_counter()
        ctx = HeatContext(V, F, block=blk, leaf_size=leaf, precision=precision)
        setup = time.perf_counter() - t0
        ctx.set_metrics(T, fr, bnd, va, health_h=hh, bin_width=w)
        ctx.wiring(src[:blk])  # warm-up
        best = float("inf")
        for _ in range(repeats):
            t0 = time.perf_counter()
            ctx.wiring(src)
            best = min(best, time.perf_counter() - t0)
        rate = src.size / best
        trials.append(dict(block=blk, leaf_size=leaf, sources_per_s=rate, setup_s=setup,
                           nnodes=ctx.info["nnodes"], factor_nnz=ctx.info["factor_nnz"]))
        if log:
            log(f"autotune: block {blk:4d} leaf {leaf:4d}: {rate:10.1f} sources/s (setup {setup:.2f}s)")
        ctx.close()
    best = max(trials, key=lambda t: t["sources_per_s"])
    return best, trials
