"""Seeded synthetic cortical-like surfaces for parity tests, benchmarks and the serving stream.

The reference ships no meshes; its tests build an icosphere and fold it radially
(lean_testkit.py:66-99). The workloads here are folded icospheres of levels 4-7:

* ``icosphere(level)`` - unit icosphere by edge-midpoint subdivision of the
  icosahedron (10*4^level + 2 vertices, 20*4^level faces).
* ``fold_sh`` - radial displacement R*(1 + amp*f), f a seeded sum of real
  spherical harmonics of degree 1..lmax with coefficients N(0,1)/l, scaled to
  max|f| = 1 so the surface stays star-shaped (no self-intersection).
* ``jitter_tangent`` - each vertex moved in its tangent plane by a seeded
  uniform offset of length <= frac*h (h = median unit-sphere edge), then
  re-projected onto the sphere.
* ``cap_mask`` - drops a seeded geodesic cap holding ``frac`` of the surface
  area (vertex-area weighted), which gives the cortical submesh a boundary.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

__all__ = [
    "icosphere",
    "real_sh_field",
    "fold_sh",
    "jitter_tangent",
    "cap_mask",
    "SyntheticSurface",
    "make_surface",
    "config_surface",
    "serving_stream",
    "TWELVE_SCALES",
]

# The 12 log-spaced area fractions 0.002 .. 0.05 (0.002 * 25**(i/11) rounded to 8 decimals),
# as the reference lists them: run_pipeline.py:248-267, verify_lean_integration.py:139-140,
# test_fastcw_lean.py:51-52.
TWELVE_SCALES = (0.002, 0.00267988, 0.00359088, 0.00481157, 0.00644721, 0.00863888,
                 0.01157558, 0.01551059, 0.02078326, 0.02784833, 0.03731509, 0.05)


def icosphere(level: int):
    """Unit icosphere, float64 vertices and int32 faces (outward orientation)."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    V = np.array(
        [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
         (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)],
        dtype=np.float64,
    )
    V /= np.linalg.norm(V, axis=1)[:, None]
    F = np.array(
        [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)],
        dtype=np.int64,
    )
    for _ in range(int(level)):
        nF = F.shape[0]
        edges = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]], axis=0)
        edges.sort(axis=1)
        uniq, inv = np.unique(edges, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = V[uniq[:, 0]] + V[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1)[:, None]
        base = V.shape[0]
        V = np.concatenate([V, mid], axis=0)
        m01 = inv[:nF] + base
        m12 = inv[nF:2 * nF] + base
        m20 = inv[2 * nF:] + base
        F = np.concatenate(
            [
                np.stack([F[:, 0], m01, m20], axis=1),
                np.stack([F[:, 1], m12, m01], axis=1),
                np.stack([F[:, 2], m20, m12], axis=1),
                np.stack([m01, m12, m20], axis=1),
            ],
            axis=0,
        )
    return V, F.astype(np.int32)


def _legendre_normalized(lmax: int, x: np.ndarray):
    """Orthonormal associated Legendre values P[l][m](x) for 0 <= m <= l <= lmax.

    Standard three-term recursion for the 4*pi-orthonormal functions (without
    the Condon-Shortley phase); stable for the degrees used here (<= 64).
    """
    s = np.sqrt(np.maximum(0.0, 1.0 - x * x))
    P = {}
    pmm = np.full_like(x, 1.0 / math.sqrt(4.0 * math.pi))
    for m in range(lmax + 1):
        if m > 0:
            pmm = pmm * s * math.sqrt((2.0 * m + 1.0) / (2.0 * m))
        P[(m, m)] = pmm
        if m + 1 <= lmax:
            P[(m + 1, m)] = math.sqrt(2.0 * m + 3.0) * x * pmm
        for l in range(m + 2, lmax + 1):
            a = math.sqrt((4.0 * l * l - 1.0) / (l * l - m * m))
            b = math.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1.0) ** 2 - 1.0))
            P[(l, m)] = a * (x * P[(l - 1, m)] - b * P[(l - 2, m)])
    return P


def real_sh_field(U: np.ndarray, lmax: int, seed: int):
    """Seeded sum over 1 <= l <= lmax of real spherical harmonics, coefficients N(0,1)/l."""
    U = np.asarray(U, dtype=np.float64)
    x = np.clip(U[:, 2], -1.0, 1.0)
    phi = np.arctan2(U[:, 1], U[:, 0])
    rng = np.random.default_rng(seed)
    P = _legendre_normalized(lmax, x)
    f = np.zeros(U.shape[0])
    for l in range(1, lmax + 1):
        c = rng.standard_normal(2 * l + 1) / l
        f += c[l] * P[(l, 0)]
        for m in range(1, l + 1):
            f += c[l + m] * math.sqrt(2.0) * P[(l, m)] * np.cos(m * phi)
            f += c[l - m] * math.sqrt(2.0) * P[(l, m)] * np.sin(m * phi)
    return f


def fold_sh(U: np.ndarray, R: float, amp: float, lmax: int, seed: int):
    """Radially folded sphere: R*(1 + amp*f/max|f|) along each unit direction."""
    f = real_sh_field(U, lmax, seed)
    f = f / max(float(np.max(np.abs(f))), 1e-300)
    return U * (R * (1.0 + amp * f))[:, None]


def _median_edge(V, F):
    e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]], axis=0)
    return float(np.median(np.linalg.norm(V[e[:, 1]] - V[e[:, 0]], axis=1)))


def jitter_tangent(U: np.ndarray, F: np.ndarray, frac: float, seed: int):
    """Move unit-sphere vertices tangentially by <= frac*h (uniform in a disc), then re-project."""
    U = np.asarray(U, dtype=np.float64)
    h = _median_edge(U, F)
    rng = np.random.default_rng(seed)
    # Tangent basis per vertex.
    ref = np.where(np.abs(U[:, 2:3]) < 0.9, np.array([[0.0, 0.0, 1.0]]), np.array([[1.0, 0.0, 0.0]]))
    t1 = np.cross(U, ref)
    t1 /= np.linalg.norm(t1, axis=1)[:, None]
    t2 = np.cross(U, t1)
    rad = frac * h * np.sqrt(rng.uniform(0.0, 1.0, U.shape[0]))
    ang = rng.uniform(0.0, 2.0 * math.pi, U.shape[0])
    P = U + (rad * np.cos(ang))[:, None] * t1 + (rad * np.sin(ang))[:, None] * t2
    return P / np.linalg.norm(P, axis=1)[:, None]


def _vertex_areas(V, F):
    p0, p1, p2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    a = 0.5 * np.linalg.norm(np.cross(p1 - p0, p2 - p0), axis=1)
    va = np.zeros(V.shape[0])
    for i in range(3):
        np.add.at(va, F[:, i], a / 3.0)
    return va


def cap_mask(U: np.ndarray, V: np.ndarray, F: np.ndarray, frac: float, seed: int):
    """Cortex mask (True = kept) that drops a seeded cap holding ``frac`` of the area.

    The cap is {angle(u, c) < theta} on the unit directions U with a seeded
    centre c; theta is chosen so the removed vertex area is ``frac`` of the
    total vertex area of the folded surface V.
    """
    if frac <= 0.0:
        return np.ones(U.shape[0], dtype=bool)
    rng = np.random.default_rng(seed)
    c = rng.standard_normal(3)
    c /= np.linalg.norm(c)
    ang = np.arccos(np.clip(U @ c, -1.0, 1.0))
    va = _vertex_areas(V, F)
    order = np.argsort(ang, kind="stable")
    cum = np.cumsum(va[order])
    n_drop = int(np.searchsorted(cum, frac * cum[-1]))
    mask = np.ones(U.shape[0], dtype=bool)
    mask[order[:n_drop]] = False
    return mask


@dataclass
class SyntheticSurface:
    vertices: np.ndarray  # (N, 3) float64, mm
    faces: np.ndarray  # (M, 3) int32
    cortex_mask: np.ndarray  # (N,) bool
    description: str


def make_surface(level: int, radius: float, fold_amp: float = 0.0, lmax: int = 0,
                 jitter: float = 0.0, cap_frac: float = 0.0, seed: int = 0) -> SyntheticSurface:
    """Build one synthetic surface. Seeds: fold = seed, jitter = seed+1, cap = seed+2."""
    U, F = icosphere(level)
    if jitter > 0.0:
        U = jitter_tangent(U, F, jitter, seed + 1)
    if fold_amp > 0.0 and lmax > 0:
        V = fold_sh(U, radius, fold_amp, lmax, seed)
    else:
        V = U * float(radius)
    mask = cap_mask(U, V, F, cap_frac, seed + 2)
    desc = (f"icosphere L{level} R={radius}mm fold={fold_amp}xSH(l<={lmax}) jitter={jitter}h "
            f"cap={cap_frac} seed={seed}")
    return SyntheticSurface(np.ascontiguousarray(V), np.ascontiguousarray(F), mask, desc)


def config_surface(index: int, seed: int = 12345) -> SyntheticSurface:
    """The four single-mesh workloads.

    0: L4 sphere R=50, closed.   1: L5 folded (l<=16), 10% cap.
    2: L6 folded (l<=16), 10% cap.   3: L7 folded (l<=32), 0.3h jitter, 10% cap.
    """
    if index == 0:
        return make_surface(4, 50.0, seed=seed)
    if index == 1:
        return make_surface(5, 70.0, 0.15, 16, 0.0, 0.10, seed)
    if index == 2:
        return make_surface(6, 70.0, 0.15, 16, 0.0, 0.10, seed)
    if index == 3:
        return make_surface(7, 70.0, 0.15, 32, 0.3, 0.10, seed)
    raise ValueError(f"unknown workload index {index}")


def serving_stream(n_requests: int = 40, mean_gap_s: float = 5.0, seed: int = 2024,
                   levels=(5, 6, 7), mix=(0.3, 0.4, 0.3), cap_range=(0.05, 0.20)):
    """Seeded hemisphere requests: (arrival_s, level, cap_frac, mesh_seed) with Poisson arrivals."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(mean_gap_s, n_requests)
    arrivals = np.cumsum(gaps) - gaps[0]
    lv = rng.choice(np.asarray(levels), size=n_requests, p=np.asarray(mix))
    caps = rng.uniform(cap_range[0], cap_range[1], n_requests)
    seeds = rng.integers(0, 2 ** 31 - 1, n_requests)
    return [
        dict(arrival_s=float(a), level=int(l), cap_frac=float(c), seed=int(s))
        for a, l, c, s in zip(arrivals, lv, caps, seeds)
    ]
