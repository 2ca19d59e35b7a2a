"""The autotuner's selection logic on the CPU, with a stand-in context whose speed depends on
(block, leaf size) in a known way; the device trials run in bench.py."""

import time

import numpy as np


class _StubCtx:
    made = []

    def __init__(self, V, F, block, leaf_size, precision, streams, tile_rows=0):
        self.block = block
        self.leaf = leaf_size
        self.tile = tile_rows
        self.info = {"nnodes": 10, "factor_nnz": 100}
        self.closed = False
        _StubCtx.made.append(self)

    def set_metrics(self, *a, **k):
        pass

    def wiring(self, src):
        # fastest at block 256, leaf 64: cost per source grows away from it
        cost = 1.0 + abs(np.log2(self.block / 256)) + 0.5 * abs(np.log2(self.leaf / 64))
        cost -= 0.4 if self.tile == 32 else 0.0  # forcing 32-row tiles is faster for this stand-in
        time.sleep(2e-6 * cost * len(src))
        return {}

    def close(self):
        self.closed = True


def test_autotune_picks_the_fastest_candidate_and_closes_every_context():
    from fastcorticalwiring_amd import autotune, meshes

    V, F = meshes.icosphere(2)
    _StubCtx.made = []
    best, trials = autotune.autotune(V * 30, F, blocks=(128, 256, 512), leaf_sizes=(32, 64, 128), n_sources=2048,
                                     make_context=_StubCtx, repeats=1)
    assert len(trials) == 9 and len(_StubCtx.made) == 9 and all(c.closed for c in _StubCtx.made)
    assert (best["block"], best["leaf_size"]) == (256, 64)
    assert best["sources_per_s"] == max(t["sources_per_s"] for t in trials)


def test_tuner_caches_per_size_class():
    from fastcorticalwiring_amd import autotune, meshes

    _StubCtx.made = []
    tuner = autotune.Tuner(blocks=(128, 256), leaf_sizes=(64,), make_context=_StubCtx, n_sources=512)
    V, F = meshes.icosphere(2)  # 162 vertices
    a = tuner.choose(V * 30, F)
    n1 = len(_StubCtx.made)
    b = tuner.choose(V * 31, F)  # same size class: no new trials
    assert a == b == {"block": 256, "leaf_size": 64} and len(_StubCtx.made) == n1 == 2
    V3, F3 = meshes.icosphere(3)  # 642 vertices: another class
    tuner.choose(V3 * 30, F3)
    assert len(_StubCtx.made) == 4 and len(tuner.choices) == 2
    assert autotune.size_class(162) != autotune.size_class(642)


def test_autotune_tries_tile_heights_on_the_best_pair():
    from fastcorticalwiring_amd import autotune, meshes

    V, F = meshes.icosphere(2)
    _StubCtx.made = []
    best, trials = autotune.autotune(V * 30, F, blocks=(128, 256), leaf_sizes=(64,), n_sources=1024,
                                     make_context=_StubCtx, repeats=1, tile_rows=(0, 64, 32))
    # two (block, leaf) trials with the chosen heights, then the two forced heights on the best pair
    assert [(t["block"], t["tile_rows"]) for t in trials] == [(128, 0), (256, 0), (256, 64), (256, 32)]
    assert (best["block"], best["tile_rows"]) == (256, 32)
    assert all(c.closed for c in _StubCtx.made)
