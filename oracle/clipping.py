"""Oracle: brute-force polygon-clipping disc area and isoline length (test infrastructure only).

An independent algorithm for the quantities lean_geometry computes in closed
form: each face is clipped against the half-space {d_lin <= r} of the linearly
interpolated field and the clipped polygon's area / cut-segment length are
measured with 3-D cross products, over every face (no buckets, no bands).
This is the legacy method of core_analysis.py:184-519 without its eps-snapping
heuristics (which only act when |d - r| <= eps); the reference's own test
compares the two methods at 1e-12 (test_fastcw_lean.py:28-42).
"""

import numpy as np


def _tri_area(a, b, c):
    return 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=-1)


def _cut(pa, pb, sa, sb):
    """Point where the linear field crosses zero on edge a->b (sa, sb of opposite signs)."""
    t = sa / (sa - sb)
    return pa + (pb - pa) * t[:, None]


def clip_area_perimeter(V, F, d, r):
    """(area of {d <= r}, length of {d = r}) over all faces, for scalar r; NaN faces skipped."""
    V = np.asarray(V, dtype=np.float64)
    F = np.asarray(F, dtype=np.int64)
    d = np.asarray(d, dtype=np.float64)
    P = V[F]  # (m, 3, 3)
    D = d[F]  # (m, 3)
    ok = np.all(np.isfinite(D), axis=1)
    P, D = P[ok], D[ok]
    s = D - float(r)
    inside = s <= 0.0
    nin = inside.sum(axis=1)
    full = _tri_area(P[:, 0], P[:, 1], P[:, 2])
    area = np.where(nin == 3, full, 0.0)
    perim = np.zeros(P.shape[0])
    for want_in, nsel in ((True, 1), (False, 2)):
        sel = np.flatnonzero(nin == nsel)
        if sel.size == 0:
            continue
        flag = inside[sel] if want_in else ~inside[sel]
        a_idx = np.argmax(flag, axis=1)  # the lone vertex (inside for nin==1, outside for nin==2)
        b_idx = (a_idx + 1) % 3
        c_idx = (a_idx + 2) % 3
        rows = np.arange(sel.size)
        Ps, Ss = P[sel], s[sel]
        pa, pb, pc = Ps[rows, a_idx], Ps[rows, b_idx], Ps[rows, c_idx]
        sa, sb, sc = Ss[rows, a_idx], Ss[rows, b_idx], Ss[rows, c_idx]
        q1 = _cut(pa, pb, sa, sb)
        q2 = _cut(pa, pc, sa, sc)
        corner = _tri_area(pa, q1, q2)
        if want_in:
            area[sel] = corner
        else:
            area[sel] = full[sel] - corner
        perim[sel] = np.linalg.norm(q1 - q2, axis=1)
    return float(area.sum()), float(perim.sum())
