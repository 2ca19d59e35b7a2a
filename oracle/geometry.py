"""Oracle: per-source disc geometry (test infrastructure only).

Restates lean_geometry.py:44-367. The per-face scalar kernels
(buckets, area_at, perimeter_at, solve_radius, per_source_geometry) run from
``lean_geometry.c``; the once-per-mesh helpers are numpy here.
"""

import math

import numpy as np

from . import _clib


def precompute_face_metric(V, F):
    """(area, g11, g12, g22): face area and inverse-Gram coefficients (lean_geometry.py:44-63)."""
    V = np.asarray(V, dtype=np.float64)
    F = np.asarray(F, dtype=np.int64)
    e1 = V[F[:, 1]] - V[F[:, 0]]
    e2 = V[F[:, 2]] - V[F[:, 0]]
    a11 = np.einsum("ij,ij->i", e1, e1)
    a22 = np.einsum("ij,ij->i", e2, e2)
    a12 = np.einsum("ij,ij->i", e1, e2)
    det = a11 * a22 - a12 * a12
    area = 0.5 * np.sqrt(np.maximum(det, 0.0))
    inv = np.zeros_like(det)
    good = det > 0.0
    inv[good] = 1.0 / det[good]
    return (np.ascontiguousarray(area), np.ascontiguousarray(a22 * inv),
            np.ascontiguousarray(-a12 * inv), np.ascontiguousarray(a11 * inv))


def default_bin_width(V, F, quantile=0.999, pad=1.05):
    """pad x quantile of per-face max edge length (lean_geometry.py:99-114)."""
    V = np.asarray(V, dtype=np.float64)
    F = np.asarray(F, dtype=np.int64)
    L = np.maximum.reduce([
        np.linalg.norm(V[F[:, 1]] - V[F[:, 0]], axis=1),
        np.linalg.norm(V[F[:, 2]] - V[F[:, 1]], axis=1),
        np.linalg.norm(V[F[:, 0]] - V[F[:, 2]], axis=1),
    ])
    return float(np.quantile(L, quantile) * pad)


def face_cdf(d0, d1, d2, A, r):
    """Scalar (area of {d <= r} in a linear face, dA/dr) (lean_geometry.py:66-88)."""
    a, b, c = sorted((d0, d1, d2))
    if r >= c:
        return A, 0.0
    if r <= a:
        return 0.0, 0.0
    if r < b:
        den = (b - a) * (c - a)
        x = r - a
        return A * x * x / den, 2.0 * A * x / den
    den = (c - a) * (c - b)
    y = c - r
    return A * (1.0 - y * y / den), 2.0 * A * y / den


def area_perimeter(d, F, area, g11, g12, g22, width, radii):
    """(A(r), dA/dr(r), P(r), n_long) for each radius, via the bucketed kernels."""
    d = np.ascontiguousarray(d, dtype=np.float64)
    F32 = np.ascontiguousarray(F, dtype=np.int32)
    radii = np.ascontiguousarray(np.atleast_1d(radii), dtype=np.float64)
    A = np.empty_like(radii)
    dA = np.empty_like(radii)
    P = np.empty_like(radii)
    nl = np.zeros(1, dtype=np.int64)
    rc = _clib.lib().oc_area_perimeter(
        _clib.dptr(d), _clib.iptr(F32), F32.shape[0], _clib.dptr(np.ascontiguousarray(area)),
        _clib.dptr(np.ascontiguousarray(g11)), _clib.dptr(np.ascontiguousarray(g12)),
        _clib.dptr(np.ascontiguousarray(g22)), float(width), _clib.dptr(radii), radii.size,
        _clib.dptr(A), _clib.dptr(dA), _clib.dptr(P), _clib.lptr(nl))
    if rc:
        raise MemoryError("oracle buckets allocation failed")
    return A, dA, P, int(nl[0])


def solve_radius(d, F, area, width, targets, rtol=1e-10, max_iter=60):
    """(r, n_eval) per target (lean_geometry.py:247-299)."""
    d = np.ascontiguousarray(d, dtype=np.float64)
    F32 = np.ascontiguousarray(F, dtype=np.int32)
    targets = np.ascontiguousarray(np.atleast_1d(targets), dtype=np.float64)
    r = np.empty_like(targets)
    ne = np.zeros(targets.size, dtype=np.int64)
    rc = _clib.lib().oc_solve_radius(
        _clib.dptr(d), _clib.iptr(F32), F32.shape[0], _clib.dptr(np.ascontiguousarray(area)), float(width),
        _clib.dptr(targets), targets.size, float(rtol), int(max_iter), _clib.dptr(r), _clib.lptr(ne))
    if rc:
        raise MemoryError("oracle buckets allocation failed")
    return r, ne


def per_source_geometry(D, F, area, g11, g12, g22, width, sorted_targets, extra_fracs, rtol=1e-10,
                        max_iter=60):
    """Batched lean_geometry.per_source_geometry (lean_geometry.py:302-340) over rows of D (ns, n).

    Returns dict of radii (ns,S), perim (ns,S), extra_r/extra_a (ns,(S-1)q), n_eval (ns,).
    """
    D = np.ascontiguousarray(np.atleast_2d(D), dtype=np.float64)
    ns, n = D.shape
    F32 = np.ascontiguousarray(F, dtype=np.int32)
    T = np.ascontiguousarray(sorted_targets, dtype=np.float64)
    fr = np.ascontiguousarray(extra_fracs, dtype=np.float64)
    S, q = T.size, fr.size
    ne = max(0, S - 1) * q
    out = dict(radii=np.empty((ns, S)), perim=np.empty((ns, S)), extra_r=np.empty((ns, ne)),
               extra_a=np.empty((ns, ne)), n_eval=np.zeros(ns, dtype=np.int64))
    rc = _clib.lib().oc_per_source_geometry(
        _clib.dptr(D), ns, n, _clib.iptr(F32), F32.shape[0], _clib.dptr(np.ascontiguousarray(area)),
        _clib.dptr(np.ascontiguousarray(g11)), _clib.dptr(np.ascontiguousarray(g12)),
        _clib.dptr(np.ascontiguousarray(g22)), float(width), _clib.dptr(T), S, _clib.dptr(fr), q,
        float(rtol), int(max_iter), _clib.dptr(out["radii"]), _clib.dptr(out["perim"]),
        _clib.dptr(out["extra_r"]), _clib.dptr(out["extra_a"]), _clib.lptr(out["n_eval"]))
    if rc < 0:
        raise MemoryError("oracle per_source_geometry allocation failed")
    return out


def field_health(d, V, src, h, neg_frac_tol=1e-3, chord_frac_tol=1e-3):
    """(neg_frac, chord_violation_frac, is_bad) (lean_geometry.py:343-357)."""
    d = np.asarray(d, dtype=np.float64)
    if not np.isfinite(d).all():
        return float("nan"), float("nan"), True
    chord = np.linalg.norm(V - V[int(src)], axis=1)
    neg = float(np.mean(d < -h))
    viol = float(np.mean(d < 0.9 * chord - h))
    return neg, viol, bool(neg > neg_frac_tol or viol > chord_frac_tol)


def boundary_area_loss_fraction(radius, boundary_distance):
    """Circular-segment area fraction lost to a straight boundary (core_analysis.py:759-777)."""
    r = float(radius)
    d = float(boundary_distance)
    if not np.isfinite(r) or r <= 0.0:
        return np.nan
    if not np.isfinite(d):
        return 0.0
    if d >= r:
        return 0.0
    x = min(1.0, max(0.0, d / r))
    return float((math.acos(x) - x * math.sqrt(max(0.0, 1.0 - x * x))) / math.pi)
