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
//
// Non-manifold edges (three or more faces) and inconsistently oriented neighbours have no
// halfedge twin, so such a mesh is first replaced by its tufted cover (Sharp & Crane 2020,
// section 4.2): every face is taken twice, once per side, and around every edge the face sides
// are glued in the cyclic order of the faces' angles about the edge (each wedge between two
// consecutive faces glues the sides that face into it; a boundary edge glues a face's two sides
// to each other). The cover is edge-manifold and consistently oriented, so the flips run on it
// unchanged; its operators count every face twice, so each cover face carries half its area.
// After flipping, a cover face and its mirror copy (same vertices, same lengths) are merged back
// into one face of full area, which leaves a manifold mesh's result exactly as without the cover.
#include <algorithm>
#include <array>
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

inline size_t next_he(size_t h) { return h - h % 3 + (h % 3 + 1) % 3; }
inline size_t prev_he(size_t h) { return h - h % 3 + (h % 3 + 2) % 3; }
inline uint64_t edge_key(int a, int b) { return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b; }

inline double dist3(const double* a, const double* b) {
    return std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
}

// Halfedge twins of a mesh whose every edge has at most two faces, oriented against each other.
// Returns false if some directed edge appears twice (a non-manifold edge, or two neighbours with
// inconsistent orientation).
bool manifold_twins(const std::vector<int>& T, std::vector<int64_t>& twin) {
    const size_t nh = T.size();
    twin.assign(nh, -1);
    std::unordered_map<uint64_t, int64_t> dir;
    dir.reserve(nh * 2);
    for (size_t h = 0; h < nh; ++h)
        if (!dir.emplace(edge_key(T[h], T[next_he(h)]), (int64_t)h).second) return false;
    for (size_t h = 0; h < nh; ++h) {
        auto it = dir.find(edge_key(T[next_he(h)], T[h]));
        if (it != dir.end()) twin[h] = it->second;
    }
    return true;
}

// The tufted cover of (V, F): cover face 2f + s is face f seen from side s (s = 0: the vertex
// order of f; s = 1: reversed, (a, c, b)), and the twins glue the sides around every edge in the
// cyclic order of the faces about it. Returns the number of edges with three or more faces.
int tufted_cover(const double* V, int m, const int* F, std::vector<int>& T, std::vector<int64_t>& twin,
                 std::string& err) {
    T.resize(6 * (size_t)m);
    for (int f = 0; f < m; ++f) {
        const int a = F[3 * (size_t)f], b = F[3 * (size_t)f + 1], c = F[3 * (size_t)f + 2];
        int* t0 = &T[6 * (size_t)f];
        t0[0] = a, t0[1] = b, t0[2] = c;  // side 0
        t0[3] = a, t0[4] = c, t0[5] = b;  // side 1
    }
    twin.assign(T.size(), -1);
    // cover halfedge of face f, side s, on the input halfedge 3f + c (from F[3f + c] to F[3f + c + 1]):
    // side 0 keeps c, side 1 runs the edge the other way in slot 2 - c
    auto cover_he = [](int f, int side, int c) { return (int64_t)(6 * (size_t)f + 3 * side + (side ? 2 - c : c)); };
    std::unordered_map<uint64_t, std::vector<int>> around;  // undirected edge -> input halfedges
    around.reserve(3 * (size_t)m);
    for (size_t h = 0; h < 3 * (size_t)m; ++h) {
        int i = F[h], j = F[next_he(h)];
        if (i > j) std::swap(i, j);
        around[edge_key(i, j)].push_back((int)h);
    }
    int nonmanifold = 0;
    std::vector<std::pair<double, int>> ring;
    for (auto& kv : around) {
        const int i = (int)(kv.first >> 32), j = (int)(kv.first & 0xffffffffu);
        const std::vector<int>& H = kv.second;
        if (H.size() >= 3) ++nonmanifold;
        const double* xi = V + 3 * (size_t)i;
        const double* xj = V + 3 * (size_t)j;
        double e[3] = {xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]};
        const double el = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        for (double& x : e) x /= el;
        // per incident face: its opposite vertex's direction about the edge, and which side of the
        // face looks towards increasing angle
        ring.clear();
        std::vector<char> plus_up(H.size());
        double u[3] = {0, 0, 0}, v[3] = {0, 0, 0};
        for (size_t q = 0; q < H.size(); ++q) {
            const int h = H[q];
            const int f = h / 3;
            const int k = F[3 * (size_t)f + (h % 3 + 2) % 3];
            const double* xk = V + 3 * (size_t)k;
            double w[3] = {xk[0] - xi[0], xk[1] - xi[1], xk[2] - xi[2]};
            const double we = w[0] * e[0] + w[1] * e[1] + w[2] * e[2];
            for (int d = 0; d < 3; ++d) w[d] -= we * e[d];
            if (q == 0) {
                const double wl = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
                for (int d = 0; d < 3; ++d) u[d] = w[d] / wl;
                v[0] = e[1] * u[2] - e[2] * u[1];
                v[1] = e[2] * u[0] - e[0] * u[2];
                v[2] = e[0] * u[1] - e[1] * u[0];
            }
            const double theta = std::atan2(w[0] * v[0] + w[1] * v[1] + w[2] * v[2], w[0] * u[0] + w[1] * u[1] + w[2] * u[2]);
            ring.push_back({theta, (int)q});
            // direction of increasing angle at the face: e x w; the face's normal from its vertex order
            const double t[3] = {e[1] * w[2] - e[2] * w[1], e[2] * w[0] - e[0] * w[2], e[0] * w[1] - e[1] * w[0]};
            const double* p0 = V + 3 * (size_t)F[3 * (size_t)f];
            const double* p1 = V + 3 * (size_t)F[3 * (size_t)f + 1];
            const double* p2 = V + 3 * (size_t)F[3 * (size_t)f + 2];
            const double a1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
            const double a2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
            const double nrm[3] = {a1[1] * a2[2] - a1[2] * a2[1], a1[2] * a2[0] - a1[0] * a2[2], a1[0] * a2[1] - a1[1] * a2[0]};
            plus_up[q] = (nrm[0] * t[0] + nrm[1] * t[1] + nrm[2] * t[2]) > 0.0;
        }
        std::sort(ring.begin(), ring.end(), [&](const std::pair<double, int>& x, const std::pair<double, int>& y) {
            return x.first != y.first ? x.first < y.first : H[x.second] < H[y.second];
        });
        // the wedge after face a (in increasing angle) glues a's side facing up with b's side
        // facing down; a boundary edge's one face glues its two sides
        for (size_t r = 0; r < ring.size(); ++r) {
            const int qa = ring[r].second, qb = ring[(r + 1) % ring.size()].second;
            const int ha = H[qa], hb = H[qb];
            const int64_t ca = cover_he(ha / 3, plus_up[qa] ? 0 : 1, ha % 3);
            const int64_t cb = cover_he(hb / 3, plus_up[qb] ? 1 : 0, hb % 3);
            if (T[(size_t)ca] != T[next_he((size_t)cb)] || T[(size_t)cb] != T[next_he((size_t)ca)] ||
                twin[(size_t)ca] >= 0 || twin[(size_t)cb] >= 0) {
                err = "tufted cover: inconsistent gluing";
                return -1;
            }
            twin[(size_t)ca] = cb;
            twin[(size_t)cb] = ca;
        }
    }
    return nonmanifold;
}

// Merge each cover face with its mirror copy (same vertices, same lengths to 1e-12): one face of
// weight 1; a face left without its copy keeps weight 1/2.
void merge_mirrors(Intrinsic& out) {
    const size_t mf = out.faces.size() / 3;
    std::vector<int> key_order(mf);
    for (size_t f = 0; f < mf; ++f) key_order[f] = (int)f;
    auto sorted3 = [&](size_t f) {
        std::array<int, 3> v = {out.faces[3 * f], out.faces[3 * f + 1], out.faces[3 * f + 2]};
        std::sort(v.begin(), v.end());
        return v;
    };
    auto lens = [&](size_t f) {
        std::array<double, 3> l = {out.len[3 * f], out.len[3 * f + 1], out.len[3 * f + 2]};
        std::sort(l.begin(), l.end());
        return l;
    };
    std::stable_sort(key_order.begin(), key_order.end(), [&](int a, int b) { return sorted3(a) < sorted3(b); });
    std::vector<char> used(mf, 0);
    std::vector<int> faces;
    std::vector<double> len, weight;
    auto emit = [&](size_t f, double wgt) {
        for (int c = 0; c < 3; ++c) {
            faces.push_back(out.faces[3 * f + c]);
            len.push_back(out.len[3 * f + c]);
        }
        weight.push_back(wgt);
    };
    for (size_t g0 = 0; g0 < mf;) {
        size_t g1 = g0;
        const auto k0 = sorted3(key_order[g0]);
        while (g1 < mf && sorted3(key_order[g1]) == k0) ++g1;
        for (size_t a = g0; a < g1; ++a) {
            const size_t fa = key_order[a];
            if (used[fa]) continue;
            used[fa] = 1;
            const auto la = lens(fa);
            size_t match = mf;
            for (size_t b = a + 1; b < g1 && match == mf; ++b) {
                const size_t fb = key_order[b];
                if (used[fb]) continue;
                const auto lb = lens(fb);
                bool same = true;
                for (int c = 0; c < 3; ++c) same = same && std::fabs(la[c] - lb[c]) <= 1e-12 * la[c];
                if (same) match = fb;
            }
            if (match != mf) used[match] = 1;
            emit(fa, match != mf ? 1.0 : 0.5);
        }
        g0 = g1;
    }
    out.faces.swap(faces);
    out.len.swap(len);
    out.weight.swap(weight);
}

}  // namespace

int intrinsic_delaunay(int n, const double* V, int m, const int* F, double mollify, Intrinsic& out, std::string& err) {
    out = Intrinsic();
    for (size_t h = 0; h < 3 * (size_t)m; ++h)
        if (F[h] == F[next_he(h)]) {
            err = "face with a repeated vertex";
            return 1;
        }
    // halfedge h = 3f + c runs from T[h] to T[3f + (c + 1) % 3]
    out.faces.assign(F, F + 3 * (size_t)m);
    std::vector<int64_t> twin;
    if (!manifold_twins(out.faces, twin)) {
        const int nm = tufted_cover(V, m, F, out.faces, twin, err);
        if (nm < 0) return 1;
        out.cover = 1;
        out.nonmanifold_edges = nm;
    }
    std::vector<int>& T = out.faces;
    const size_t nh = T.size();
    const int mf = (int)(nh / 3);
    auto next = next_he;
    auto prev = prev_he;
    std::vector<double> len(nh);
    for (size_t h = 0; h < nh; ++h) len[h] = dist3(V + 3 * (size_t)T[h], V + 3 * (size_t)T[next(h)]);
    // mean length over the input's distinct edges
    double sum = 0.0;
    int64_t ne = 0;
    {
        std::vector<uint64_t> e;
        e.reserve(3 * (size_t)m);
        for (size_t h = 0; h < 3 * (size_t)m; ++h) {
            const int a = std::min(F[h], F[next(h)]), b = std::max(F[h], F[next(h)]);
            e.push_back(edge_key(a, b));
        }
        std::sort(e.begin(), e.end());
        e.erase(std::unique(e.begin(), e.end()), e.end());
        for (uint64_t k : e) sum += dist3(V + 3 * (size_t)(k >> 32), V + 3 * (size_t)(k & 0xffffffffu));
        ne = (int64_t)e.size();
    }
    (void)n;
    // 1. mollification
    const double eps = mollify * (ne ? sum / ne : 0.0);
    double delta = 0.0;
    for (int f = 0; f < mf; ++f) {
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
    for (int f = 0; f < mf; ++f)
        for (int c = 0; c < 3; ++c) out.len[3 * (size_t)f + c] = len[3 * (size_t)f + (c + 1) % 3];
    if (out.cover)
        merge_mirrors(out);
    else
        out.weight.assign(mf, 1.0);
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
