// Intrinsic Delaunay triangulation of a triangle mesh, for the robust Laplacian.
//
// The production heat method of the reference calls potpourri3d's
// MeshHeatMethodDistanceSolver(V, F, t_coef, use_robust=True) (distance_engines.py:78-79,
// 119-125). Its robust mode builds the operators on an intrinsic Delaunay triangulation of
// the input surface after intrinsic mollification (Sharp & Crane, "A Laplacian for Nonmanifold
// Triangle Meshes", SGP 2020, section 4.5; Fisher, Springborn, Bobenko, Schroder, "An algorithm
// for the construction of intrinsic Delaunay triangulations", SIGGRAPH 2006 course notes):
//
//  1. mollify: eps = factor * (mean edge length); delta = max over triangles and corners of
//     eps - (l_a + l_b - l_c), clamped at 0, is added to every edge length, so every triangle
//     satisfies the triangle inequality with margin eps;
//  2. flip to Delaunay: while some interior edge has opposite angles summing to more than pi
//     (cot alpha + cot beta < 0), replace it by the other diagonal of its two triangles, with the
//     new length measured in the triangles' planar layout. The vertex set is unchanged and the
//     surface metric is exactly preserved; only the connectivity and edge lengths change.
//
// Degenerate flips (both triangles of the edge having the same opposite vertex) are skipped and
// counted; the cotan operators of the result need only the faces and their edge lengths.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "fcw_internal.h"

namespace fcw {

namespace {

inline double tri_area(double a, double b, double c) {
    // Heron's formula in the form stable for needle triangles (a >= b >= c)
    if (a < b) std::swap(a, b);
    if (a < c) std::swap(a, c);
    if (b < c) std::swap(b, c);
    const double q = (a + (b + c)) * (c - (a - b)) * (c + (a - b)) * (a + (b - c));
    return 0.25 * std::sqrt(std::max(q, 0.0));
}

// cot of the angle opposite side c in a triangle with sides a, b, c
inline double cot_opposite(double a, double b, double c) {
    const double A = tri_area(a, b, c);
    return (a * a + b * b - c * c) / (4.0 * A);
}

}  // namespace

int intrinsic_delaunay(int n, const double* V, int m, const int* F, double mollify, Intrinsic& out, std::string& err) {
    out = Intrinsic();
    out.faces.assign(F, F + 3 * (size_t)m);
    std::vector<int>& T = out.faces;
    const size_t nh = 3 * (size_t)m;
    // halfedge h = 3f + c runs from T[h] to T[3f + (c + 1) % 3]
    auto next = [](size_t h) { return h - h % 3 + (h % 3 + 1) % 3; };
    auto prev = [](size_t h) { return h - h % 3 + (h % 3 + 2) % 3; };
    std::vector<int64_t> twin(nh, -1);
    std::vector<double> len(nh);
    {
        std::unordered_map<uint64_t, int64_t> dir;
        dir.reserve(nh * 2);
        auto key = [](int a, int b) { return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b; };
        for (size_t h = 0; h < nh; ++h) {
            const int a = T[h], b = T[next(h)];
            if (a == b) {
                err = "face with a repeated vertex";
                return 1;
            }
            if (!dir.emplace(key(a, b), (int64_t)h).second) {
                err = "non-manifold or inconsistently oriented edge";
                return 1;
            }
        }
        for (size_t h = 0; h < nh; ++h) {
            auto it = dir.find(key(T[next(h)], T[h]));
            if (it != dir.end()) twin[h] = it->second;
        }
    }
    double sum = 0.0;
    int64_t ne = 0;
    for (size_t h = 0; h < nh; ++h) {
        const double* a = V + 3 * (size_t)T[h];
        const double* b = V + 3 * (size_t)T[next(h)];
        len[h] = std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
        if (twin[h] < 0 || (int64_t)h < twin[h]) {
            sum += len[h];
            ++ne;
        }
    }
    (void)n;
    // 1. mollification
    const double eps = mollify * (ne ? sum / ne : 0.0);
    double delta = 0.0;
    for (int f = 0; f < m; ++f) {
        const double l0 = len[3 * (size_t)f], l1 = len[3 * (size_t)f + 1], l2 = len[3 * (size_t)f + 2];
        delta = std::max({delta, eps - (l0 + l1 - l2), eps - (l1 + l2 - l0), eps - (l2 + l0 - l1)});
    }
    if (delta > 0.0)
        for (double& l : len) l += delta;
    out.delta = delta;

    // 2. flips
    auto non_delaunay = [&](size_t h) {
        const int64_t t = twin[h];
        if (t < 0) return false;
        const double ck = cot_opposite(len[next(h)], len[prev(h)], len[h]);
        const double cl = cot_opposite(len[next((size_t)t)], len[prev((size_t)t)], len[(size_t)t]);
        return ck + cl < -1e-12;
    };
    const int max_sweeps = 1000;
    bool any = true;
    for (int sweep = 0; any && sweep < max_sweeps; ++sweep) {
        any = false;
        for (size_t h = 0; h < nh; ++h) {
            const int64_t t = twin[h];
            if (t < (int64_t)h || !non_delaunay(h)) continue;
            // f1 = (i, j, k) with h: i->j, h1: j->k, h2: k->i; f2 = (j, i, l) with t: j->i, t1: i->l, t2: l->j
            const size_t h1 = next(h), h2 = prev(h), t1 = next((size_t)t), t2 = prev((size_t)t);
            const int i = T[h], j = T[h1], k = T[h2], l = T[t2];
            if (k == l) {
                ++out.skipped;
                continue;
            }
            // planar layout: i at the origin, j on the x axis, k above, l below
            const double c = len[h];
            const double bk = len[h2], ak = len[h1], bl = len[t1], al = len[t2];
            const double xk = (c * c + bk * bk - ak * ak) / (2.0 * c), yk = std::sqrt(std::max(bk * bk - xk * xk, 0.0));
            const double xl = (c * c + bl * bl - al * al) / (2.0 * c), yl = -std::sqrt(std::max(bl * bl - xl * xl, 0.0));
            const double lkl = std::hypot(xk - xl, yk - yl);
            const size_t f1 = h / 3, f2 = (size_t)t / 3;
            // outer halfedges: (k->i, i->l) go to f1, (l->j, j->k) go to f2
            const int64_t tw_h2 = twin[h2], tw_t1 = twin[t1], tw_t2 = twin[t2], tw_h1 = twin[h1];
            const double l_h2 = len[h2], l_t1 = len[t1], l_t2 = len[t2], l_h1 = len[h1];
            // new f1 = (k, i, l): slots k->i, i->l, l->k ; new f2 = (l, j, k): slots l->j, j->k, k->l
            const size_t a0 = 3 * f1, a1 = a0 + 1, a2 = a0 + 2, b0 = 3 * f2, b1 = b0 + 1, b2 = b0 + 2;
            T[a0] = k, T[a1] = i, T[a2] = l;
            T[b0] = l, T[b1] = j, T[b2] = k;
            len[a0] = l_h2, len[a1] = l_t1, len[a2] = lkl;
            len[b0] = l_t2, len[b1] = l_h1, len[b2] = lkl;
            twin[a0] = tw_h2, twin[a1] = tw_t1, twin[a2] = (int64_t)b2;
            twin[b0] = tw_t2, twin[b1] = tw_h1, twin[b2] = (int64_t)a2;
            if (tw_h2 >= 0) twin[(size_t)tw_h2] = (int64_t)a0;
            if (tw_t1 >= 0) twin[(size_t)tw_t1] = (int64_t)a1;
            if (tw_t2 >= 0) twin[(size_t)tw_t2] = (int64_t)b0;
            if (tw_h1 >= 0) twin[(size_t)tw_h1] = (int64_t)b1;
            ++out.flips;
            any = true;
        }
    }
    if (any) {
        err = "intrinsic Delaunay flips did not converge";
        return 1;
    }
    // length of the edge opposite corner c = halfedge (c + 1) % 3 of the face
    out.len.resize(nh);
    for (int f = 0; f < m; ++f)
        for (int c = 0; c < 3; ++c) out.len[3 * (size_t)f + c] = len[3 * (size_t)f + (c + 1) % 3];
    return 0;
}

void intrinsic_operators(int m, const double* len, double* area, double* gram) {
    for (int f = 0; f < m; ++f) {
        const double l0 = len[3 * (size_t)f], l1 = len[3 * (size_t)f + 1], l2 = len[3 * (size_t)f + 2];
        const double A = tri_area(l0, l1, l2);
        area[f] = A;
        // grad B_c . grad B_c = l_c^2 / (4 A^2); grad B_a . grad B_b = (l_k^2 - l_a^2 - l_b^2) / (8 A^2)
        const double s2 = 1.0 / (4.0 * A * A), s8 = 0.5 * s2;
        const double q0 = l0 * l0, q1 = l1 * l1, q2 = l2 * l2;
        double* G = gram + 6 * (size_t)f;  // G00 G01 G02 G11 G12 G22
        G[0] = q0 * s2;
        G[1] = (q2 - q0 - q1) * s8;
        G[2] = (q1 - q0 - q2) * s8;
        G[3] = q1 * s2;
        G[4] = (q0 - q1 - q2) * s8;
        G[5] = q2 * s2;
    }
}

}  // namespace fcw
