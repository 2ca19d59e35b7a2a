// Host symbolic analysis: nested-dissection ordering of the mesh graph, the
// separator tree used as the supernodal elimination tree, boundary (row)
// structures, level schedule, and the gather maps the device kernels use.
//
// The reference hands the cotan system to CHOLMOD / SuperLU
// (fastcw_lean.py:445-464, 490-500); this replaces their ordering + symbolic
// phase with one designed for level-scheduled dense panel solves on the GPU.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "fcw_internal.h"

namespace fcw {
namespace {

struct Graph {
    std::vector<int> ptr, adj;  // sorted, unique, no self loops
};

// Vertex adjacency (ascending, no duplicates, no self loops) by bucketing both directions of
// every face edge per vertex and sorting each vertex's short list.
Graph build_graph(int n, int m, const int* F) {
    std::vector<int> start(n + 1, 0);
    for (int f = 0; f < m; ++f)
        for (int a = 0; a < 3; ++a) {
            const int u = F[3 * f + a], v = F[3 * f + (a + 1) % 3];
            if (u == v) continue;
            start[u + 1]++;
            start[v + 1]++;
        }
    for (int i = 0; i < n; ++i) start[i + 1] += start[i];
    std::vector<int> nb(start[n]);
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (int f = 0; f < m; ++f)
        for (int a = 0; a < 3; ++a) {
            const int u = F[3 * f + a], v = F[3 * f + (a + 1) % 3];
            if (u == v) continue;
            nb[fill[u]++] = v;
            nb[fill[v]++] = u;
        }
    Graph g;
    g.ptr.assign(n + 1, 0);
    g.adj.resize(nb.size());
    int64_t out = 0;
    for (int u = 0; u < n; ++u) {
        int* lo = nb.data() + start[u];
        int* hi = nb.data() + start[u + 1];
        std::sort(lo, hi);
        hi = std::unique(lo, hi);
        for (int* q = lo; q < hi; ++q) g.adj[out++] = *q;
        g.ptr[u + 1] = (int)out;
    }
    g.adj.resize(out);
    return g;
}

struct TreeNode {
    std::vector<int> verts;  // separator (internal) or subdomain (leaf), input numbering
    std::vector<int> children;
};

class Dissector {
   public:
    Dissector(const Graph& g, int n, int leaf) : g_(g), leaf_(leaf), mark_(n, -1), lev_(n, -1) {}

    int build(std::vector<int> verts) {
        int id = (int)tree.size();
        tree.emplace_back();
        if ((int)verts.size() <= leaf_) {
            tree[id].verts = bfs_order(verts);
            return id;
        }
        std::vector<int> A, B, sep;
        split(verts, A, B, sep);
        if (A.empty() || B.empty()) {  // could not split: keep as one leaf block
            tree[id].verts = bfs_order(verts);
            return id;
        }
        std::vector<int>().swap(verts);
        int ca = build(std::move(A));
        int cb = build(std::move(B));
        tree[id].children = {ca, cb};
        tree[id].verts = std::move(sep);
        return id;
    }

    // As build(), with the two halves of the top `par_depth` dissections built on their own
    // threads, each by a Dissector with its own marks; their trees are appended A then B, which
    // gives every node the id the sequential recursion gives it, so the tree is the same.
    int build_parallel(std::vector<int> verts, int par_depth) {
        if (par_depth <= 0) return build(std::move(verts));
        int id = (int)tree.size();
        tree.emplace_back();
        if ((int)verts.size() <= leaf_) {
            tree[id].verts = bfs_order(verts);
            return id;
        }
        std::vector<int> A, B, sep;
        split(verts, A, B, sep);
        if (A.empty() || B.empty()) {
            tree[id].verts = bfs_order(verts);
            return id;
        }
        std::vector<int>().swap(verts);
        const int n = (int)mark_.size();
        Dissector DA(g_, n, leaf_), DB(g_, n, leaf_);
        std::exception_ptr err_b;
        std::thread tb([&] {
            try {
                DB.build_parallel(std::move(B), par_depth - 1);
            } catch (...) {
                err_b = std::current_exception();
            }
        });
        std::exception_ptr err_a;
        try {
            DA.build_parallel(std::move(A), par_depth - 1);
        } catch (...) {
            err_a = std::current_exception();
        }
        tb.join();
        if (err_a) std::rethrow_exception(err_a);
        if (err_b) std::rethrow_exception(err_b);
        const int ca = append(DA.tree);
        const int cb = append(DB.tree);
        tree[id].children = {ca, cb};
        tree[id].verts = std::move(sep);
        return id;
    }

    std::vector<TreeNode> tree;

   private:
    // Appends a subtree built by another Dissector (its root is its node 0); returns the root's id.
    int append(std::vector<TreeNode>& sub) {
        const int off = (int)tree.size();
        for (TreeNode& t : sub) {
            for (int& c : t.children) c += off;
            tree.push_back(std::move(t));
        }
        return off;
    }

    const Graph& g_;
    int leaf_;
    std::vector<int> mark_, lev_;
    int stamp_ = 0;

    void set_mark(const std::vector<int>& verts) {
        ++stamp_;
        for (int v : verts) mark_[v] = stamp_;
    }

    // BFS over the marked set from `root`; fills lev_, returns visit order.
    std::vector<int> bfs(int root, int& depth) {
        std::vector<int> q;
        q.push_back(root);
        lev_[root] = 0;
        int tag = -stamp_ - 2;  // visited marker distinct from every stamp
        mark_[root] = tag;
        for (size_t h = 0; h < q.size(); ++h) {
            int v = q[h];
            for (int e = g_.ptr[v]; e < g_.ptr[v + 1]; ++e) {
                int u = g_.adj[e];
                if (mark_[u] == stamp_) {
                    mark_[u] = tag;
                    lev_[u] = lev_[v] + 1;
                    q.push_back(u);
                }
            }
        }
        depth = lev_[q.back()];
        for (int v : q) mark_[v] = stamp_;  // restore membership
        return q;
    }

    std::vector<int> bfs_order(const std::vector<int>& verts) {
        set_mark(verts);
        std::vector<int> out;
        out.reserve(verts.size());
        std::vector<int> sorted = verts;
        std::sort(sorted.begin(), sorted.end());
        int done_tag = -1;
        for (int s : sorted) {
            if (mark_[s] != stamp_) continue;
            int d;
            std::vector<int> q = bfs(s, d);
            for (int v : q) {
                out.push_back(v);
                mark_[v] = done_tag;
            }
        }
        return out;
    }

    void split(const std::vector<int>& verts, std::vector<int>& A, std::vector<int>& B, std::vector<int>& sep) {
        set_mark(verts);
        // Connected components of the induced subgraph.
        int depth;
        std::vector<int> comp = bfs(verts[0], depth);
        if (comp.size() < verts.size()) {
            // Disconnected: put whole components on two sides (empty separator).
            std::vector<std::vector<int>> comps;
            std::vector<char> seen_local;  // use mark_ to drop visited comps
            int ctag = -1;
            std::vector<int> sorted = verts;
            std::sort(sorted.begin(), sorted.end());
            for (int s : sorted) {
                if (mark_[s] != stamp_) continue;
                int d;
                std::vector<int> c = bfs(s, d);
                for (int v : c) mark_[v] = ctag;
                comps.push_back(std::move(c));
            }
            std::sort(comps.begin(), comps.end(),
                      [](const std::vector<int>& a, const std::vector<int>& b) { return a.size() > b.size(); });
            for (auto& c : comps) {
                auto& dst = (A.size() <= B.size()) ? A : B;
                dst.insert(dst.end(), c.begin(), c.end());
            }
            return;
        }
        // Pseudo-peripheral root: repeat BFS from a min-degree vertex of the last level.
        int root = verts[0];
        int best_depth = -1;
        std::vector<int> order;
        for (int it = 0; it < 4; ++it) {
            int d;
            std::vector<int> q = bfs(root, d);
            if (d <= best_depth) break;
            best_depth = d;
            order = std::move(q);
            int cand = order.back(), cdeg = 1 << 30;
            for (int i = (int)order.size() - 1; i >= 0 && lev_[order[i]] == d; --i) {
                int v = order[i];
                int dg = g_.ptr[v + 1] - g_.ptr[v];
                if (dg < cdeg) { cdeg = dg; cand = v; }
            }
            root = cand;
        }
        // Recompute levels for the chosen structure (the last BFS may have been from a new root).
        {
            int d;
            int r0 = order.front();
            order = bfs(r0, d);
            best_depth = d;
        }
        const int nl = best_depth + 1;
        std::vector<int> cnt(nl + 1, 0);
        for (int v : order) cnt[lev_[v] + 1]++;
        std::vector<int> cum(nl + 1, 0);  // cum[l] = vertices in levels < l
        for (int l = 0; l < nl; ++l) cum[l + 1] = cum[l] + cnt[l + 1];
        // Thinned separator size per level: level-l vertices with a neighbour at level l+1.
        std::vector<int> thin(nl, 0);
        for (int v : order) {
            int l = lev_[v];
            if (l + 1 >= nl) continue;
            for (int e = g_.ptr[v]; e < g_.ptr[v + 1]; ++e) {
                int u = g_.adj[e];
                if (mark_[u] == stamp_ && lev_[u] == l + 1) { thin[l]++; break; }
            }
        }
        const double N = (double)verts.size();
        int best = -1;
        double best_cost = 1e300;
        for (int l = 1; l + 1 < nl; ++l) {
            double left = cum[l] + (cnt[l + 1] - thin[l]);
            double right = N - cum[l + 1];
            double lo = std::min(left, right), hi = std::max(left, right);
            if (lo < 0.30 * N || hi > 0.70 * N) continue;
            double cost = (double)thin[l] * (1.0 + 0.5 * std::fabs(left - right) / N);
            if (cost < best_cost) { best_cost = cost; best = l; }
        }
        if (best < 0) {
            for (int l = 0; l < nl; ++l)
                if (cum[l + 1] >= N / 2) { best = l; break; }
            if (best < 1 || best + 1 >= nl) return;  // chain-like tiny set: caller keeps it whole
        }
        for (int v : order) {
            int l = lev_[v];
            if (l < best) {
                A.push_back(v);
            } else if (l > best) {
                B.push_back(v);
            } else {
                bool is_sep = false;
                for (int e = g_.ptr[v]; e < g_.ptr[v + 1]; ++e) {
                    int u = g_.adj[e];
                    if (mark_[u] == stamp_ && lev_[u] == l + 1) { is_sep = true; break; }
                }
                (is_sep ? sep : A).push_back(v);
            }
        }
    }
};

}  // namespace

int analyse(int n, const double* V, int m, const int* F, int leaf_size, Symbolic& S, std::string& err, int par_depth) {
    (void)V;
    if (n <= 0 || m <= 0) { err = "empty mesh"; return 1; }
    for (int64_t i = 0; i < 3LL * m; ++i)
        if (F[i] < 0 || F[i] >= n) { err = "face index out of range"; return 1; }
    S = Symbolic();
    S.n = n;
    S.m = m;
    Graph g = build_graph(n, m, F);
    for (int v = 0; v < n; ++v)
        if (g.ptr[v + 1] == g.ptr[v]) { err = "vertex not referenced by any face"; return 1; }

    // ---- nested dissection --------------------------------------------------
    Dissector D(g, n, std::max(4, leaf_size));
    {
        std::vector<int> all(n);
        std::iota(all.begin(), all.end(), 0);
        // the halves of the top two dissections on their own threads on large meshes (same result)
        D.build_parallel(std::move(all), par_depth >= 0 ? par_depth : (n >= 20000 ? 2 : 0));
    }
    std::vector<TreeNode>& T = D.tree;
    // Postorder numbering of tree nodes; positions assigned in postorder.
    std::vector<int> post;  // tree ids in postorder
    {
        std::vector<std::pair<int, int>> st{{0, 0}};
        while (!st.empty()) {
            auto& [id, ci] = st.back();
            if (ci < (int)T[id].children.size()) {
                int c = T[id].children[ci++];
                st.push_back({c, 0});
            } else {
                post.push_back(id);
                st.pop_back();
            }
        }
    }
    const int nn = (int)post.size();
    std::vector<int> tree_to_node(T.size());
    for (int i = 0; i < nn; ++i) tree_to_node[post[i]] = i;
    S.nnodes = nn;
    S.node_start.resize(nn);
    S.node_s.resize(nn);
    S.node_b.assign(nn, 0);
    S.node_parent.assign(nn, -1);
    S.node_level.assign(nn, 0);
    S.perm.resize(n);
    S.iperm.assign(n, -1);
    std::vector<std::vector<int>> children(nn);
    {
        int pos = 0;
        for (int i = 0; i < nn; ++i) {
            const TreeNode& tn = T[post[i]];
            S.node_start[i] = pos;
            S.node_s[i] = (int)tn.verts.size();
            for (int v : tn.verts) {
                S.perm[pos] = v;
                S.iperm[v] = pos;
                ++pos;
            }
            for (int c : tn.children) {
                children[i].push_back(tree_to_node[c]);
                S.node_parent[tree_to_node[c]] = i;
            }
        }
        if (pos != n) { err = "nested dissection lost vertices"; return 1; }
    }
    S.child_ptr.assign(nn + 1, 0);
    for (int i = 0; i < nn; ++i) S.child_ptr[i + 1] = S.child_ptr[i] + (int)children[i].size();
    S.child_list.resize(S.child_ptr[nn]);
    for (int i = 0; i < nn; ++i)
        std::copy(children[i].begin(), children[i].end(), S.child_list.begin() + S.child_ptr[i]);

    // ---- boundary structures (bottom-up) -----------------------------------
    std::vector<std::vector<int>> bnd(nn);
    {
        std::vector<int> seen(n, -1);
        for (int i = 0; i < nn; ++i) {
            const int end = S.node_start[i] + S.node_s[i];
            std::vector<int>& b = bnd[i];
            for (int p = S.node_start[i]; p < end; ++p) {
                int v = S.perm[p];
                for (int e = g.ptr[v]; e < g.ptr[v + 1]; ++e) {
                    int q = S.iperm[g.adj[e]];
                    if (q >= end && seen[q] != i) { seen[q] = i; b.push_back(q); }
                }
            }
            for (int c : children[i]) {
                for (int q : bnd[c])
                    if (q >= end && seen[q] != i) { seen[q] = i; b.push_back(q); }
            }
            std::sort(b.begin(), b.end());
            S.node_b[i] = (int)b.size();
            int lvl = 0;
            for (int c : children[i]) lvl = std::max(lvl, S.node_level[c] + 1);
            S.node_level[i] = lvl;
        }
    }
    S.node_boff.assign(nn + 1, 0);
    for (int i = 0; i < nn; ++i) S.node_boff[i + 1] = S.node_boff[i] + S.node_b[i];
    S.bidx.resize(S.node_boff[nn]);
    S.cmap.assign(S.node_boff[nn], -1);
    for (int i = 0; i < nn; ++i) std::copy(bnd[i].begin(), bnd[i].end(), S.bidx.begin() + S.node_boff[i]);
    for (int c = 0; c < nn; ++c) {
        int p = S.node_parent[c];
        if (p < 0) {
            if (S.node_b[c] != 0) { err = "root has boundary rows"; return 1; }
            continue;
        }
        const int ps = S.node_start[p], pe = ps + S.node_s[p];
        const int* pb = S.bidx.data() + S.node_boff[p];
        for (int r = 0; r < S.node_b[c]; ++r) {
            int q = S.bidx[S.node_boff[c] + r];
            int fr;
            if (q >= ps && q < pe) {
                fr = q - ps;
            } else {
                const int* it = std::lower_bound(pb, pb + S.node_b[p], q);
                if (it == pb + S.node_b[p] || *it != q) { err = "child boundary not in parent front"; return 1; }
                fr = S.node_s[p] + (int)(it - pb);
            }
            S.cmap[S.node_boff[c] + r] = fr;
        }
    }

    // ---- levels, offsets and work counts -----------------------------------
    S.nlevels = 0;
    for (int i = 0; i < nn; ++i) S.nlevels = std::max(S.nlevels, S.node_level[i] + 1);
    S.level_ptr.assign(S.nlevels + 1, 0);
    for (int i = 0; i < nn; ++i) S.level_ptr[S.node_level[i] + 1]++;
    for (int l = 0; l < S.nlevels; ++l) S.level_ptr[l + 1] += S.level_ptr[l];
    S.level_nodes.resize(nn);
    {
        std::vector<int> fill(S.level_ptr.begin(), S.level_ptr.end() - 1);
        for (int i = 0; i < nn; ++i) S.level_nodes[fill[S.node_level[i]]++] = i;
    }
    S.node_poff.assign(nn + 1, 0);
    S.node_uoff.assign(nn + 1, 0);
    S.factor_flops = 0.0;
    S.solve_flops_per_rhs = 0.0;
    S.factor_nnz = 0;
    S.max_front = 0;
    for (int i = 0; i < nn; ++i) {
        const double s = S.node_s[i], b = S.node_b[i];
        S.node_poff[i + 1] = S.node_poff[i] + (int64_t)S.node_s[i] * (S.node_s[i] + S.node_b[i]);
        S.node_uoff[i + 1] = S.node_uoff[i] + S.node_b[i];
        // partial Cholesky + trsm + Schur update, then triangular inverse and W = L21 L11^{-1}
        S.factor_flops += s * s * s / 3.0 + b * s * s + b * b * s + s * s * s / 3.0 + b * s * s;
        S.solve_flops_per_rhs += 2.0 * 2.0 * s * (s + b);
        S.factor_nnz += (int64_t)(s * (s + 1) / 2 + s * b);
        S.max_front = std::max(S.max_front, S.node_s[i] + S.node_b[i]);
    }
    S.panel_elems = S.node_poff[nn];
    S.urows = S.node_uoff[nn];
    S.contrib_own.assign(2 * (size_t)n, -1);
    S.contrib_tail.assign(2 * (size_t)std::max<int64_t>(1, S.urows), -1);
    for (int c = 0; c < nn; ++c) {
        const int p = S.node_parent[c];
        if (p < 0) continue;
        for (int r = 0; r < S.node_b[c]; ++r) {
            const int t = S.cmap[S.node_boff[c] + r];
            const int src_row = (int)(S.node_uoff[c] + r);
            int* slot = (t < S.node_s[p]) ? &S.contrib_own[2 * (size_t)(S.node_start[p] + t)]
                                          : &S.contrib_tail[2 * (size_t)(S.node_uoff[p] + t - S.node_s[p])];
            if (slot[0] < 0) {
                slot[0] = src_row;
            } else if (slot[1] < 0) {
                slot[1] = src_row;
            } else {
                err = "more than two children feed one row";
                return 1;
            }
        }
    }

    // ---- faces in positions, reordered by their smallest position -----------
    {
        std::vector<int> key(m), idx(m);
        for (int f = 0; f < m; ++f) {
            int a = S.iperm[F[3 * f]], b = S.iperm[F[3 * f + 1]], c = S.iperm[F[3 * f + 2]];
            key[f] = std::min(a, std::min(b, c));
            idx[f] = f;
        }
        std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return key[x] < key[y]; });
        S.faces.resize(3 * (size_t)m);
        S.face_perm = idx;
        for (int f = 0; f < m; ++f)
            for (int a = 0; a < 3; ++a) S.faces[3 * (size_t)f + a] = S.iperm[F[3 * (size_t)idx[f] + a]];
    }
    // ---- CSR pattern in positions (symmetric, with diagonal) ----------------
    {
        S.rowptr.assign(n + 1, 0);
        for (int p = 0; p < n; ++p) {
            int v = S.perm[p];
            S.rowptr[p + 1] = S.rowptr[p] + (g.ptr[v + 1] - g.ptr[v]) + 1;
        }
        S.colind.resize(S.rowptr[n]);
        S.diag.resize(n);
        for (int p = 0; p < n; ++p) {
            int v = S.perm[p];
            int* row = S.colind.data() + S.rowptr[p];
            int len = 0;
            row[len++] = p;
            for (int e = g.ptr[v]; e < g.ptr[v + 1]; ++e) row[len++] = S.iperm[g.adj[e]];
            std::sort(row, row + len);
            S.diag[p] = S.rowptr[p] + (int)(std::lower_bound(row, row + len, p) - row);
        }
        // face-corner contributions per CSR entry
        const int nnz = S.rowptr[n];
        std::vector<int> cnt(nnz + 1, 0);
        std::vector<int> ent((size_t)m * 9);
        for (int f = 0; f < m; ++f) {
            for (int a = 0; a < 3; ++a) {
                int i = S.faces[3 * (size_t)f + a];
                const int* row = S.colind.data() + S.rowptr[i];
                int len = S.rowptr[i + 1] - S.rowptr[i];
                for (int b = 0; b < 3; ++b) {
                    int j = S.faces[3 * (size_t)f + b];
                    int e = S.rowptr[i] + (int)(std::lower_bound(row, row + len, j) - row);
                    ent[(size_t)f * 9 + a * 3 + b] = e;
                    cnt[e + 1]++;
                }
            }
        }
        for (int e = 0; e < nnz; ++e) cnt[e + 1] += cnt[e];
        S.contrib_ptr = cnt;
        S.contrib.resize(cnt[nnz]);
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (int f = 0; f < m; ++f)
            for (int ab = 0; ab < 9; ++ab) S.contrib[fill[ent[(size_t)f * 9 + ab]]++] = f * 9 + ab;
        // vertex -> incident face corners
        S.vf_ptr.assign(n + 1, 0);
        for (int f = 0; f < m; ++f)
            for (int a = 0; a < 3; ++a) S.vf_ptr[S.faces[3 * (size_t)f + a] + 1]++;
        for (int p = 0; p < n; ++p) S.vf_ptr[p + 1] += S.vf_ptr[p];
        S.vf.resize(S.vf_ptr[n]);
        std::vector<int> vfill(S.vf_ptr.begin(), S.vf_ptr.end() - 1);
        for (int f = 0; f < m; ++f)
            for (int a = 0; a < 3; ++a) S.vf[vfill[S.faces[3 * (size_t)f + a]]++] = f * 3 + a;
    }
    return 0;
}

}  // namespace fcw
