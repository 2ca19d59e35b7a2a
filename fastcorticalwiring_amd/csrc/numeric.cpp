// Host FP64 multifrontal Cholesky over the separator tree, producing the
// partitioned-inverse panels P_S = [L11^{-1}; L21 L11^{-1}] that turn both
// triangular sweeps into plain dense products on the GPU.
//
// Forward sweep of node S with assembled front f_S (s x k):   [y_S; u_S] = P_S f_S
// Backward sweep with ancestors' solution x_B (b x k):          x_S = P_S^T [y_S; -x_B]
// (y_S = L11^{-1} f_S, u_S = L21 y_S, x_S = L11^{-T}(y_S - L21^T x_B).)
//
// The reference factors the same two matrices, M + tL and L + shift*M, with
// CHOLMOD or SuperLU (fastcw_lean.py:490-500).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "fcw_internal.h"

namespace fcw {
namespace {

// C[0:M, 0:N] -= A[0:M, 0:K] * B[0:N, 0:K]^T   (all column-major)
// If `lower` is set only entries with global row >= global column are touched,
// where global row = i + row0, global column = j + col0.
void gemm_nt_sub(int M, int N, int K, const double* A, int lda, const double* B, int ldb, double* C, int ldc,
                 bool lower, int row0, int col0, bool parallel) {
    if (M <= 0 || N <= 0 || K <= 0) return;
    const int KB = 128;
#pragma omp parallel for schedule(dynamic, 4) if (parallel && (double)M * N * K > 2e6)
    for (int j = 0; j < N; ++j) {
        int ilo = 0;
        if (lower) ilo = std::max(0, (j + col0) - row0);
        if (ilo >= M) continue;
        double* c = C + (size_t)j * ldc;
        for (int k0 = 0; k0 < K; k0 += KB) {
            int k1 = std::min(K, k0 + KB);
            int k = k0;
            for (; k + 3 < k1; k += 4) {
                const double b0 = B[j + (size_t)k * ldb], b1 = B[j + (size_t)(k + 1) * ldb];
                const double b2 = B[j + (size_t)(k + 2) * ldb], b3 = B[j + (size_t)(k + 3) * ldb];
                const double* a0 = A + (size_t)k * lda;
                const double* a1 = a0 + lda;
                const double* a2 = a1 + lda;
                const double* a3 = a2 + lda;
#pragma omp simd
                for (int i = ilo; i < M; ++i) c[i] -= a0[i] * b0 + a1[i] * b1 + a2[i] * b2 + a3[i] * b3;
            }
            for (; k < k1; ++k) {
                const double b0 = B[j + (size_t)k * ldb];
                const double* a0 = A + (size_t)k * lda;
#pragma omp simd
                for (int i = ilo; i < M; ++i) c[i] -= a0[i] * b0;
            }
        }
    }
}

// Unblocked lower Cholesky of an n x n block (column-major). Returns the failing column or -1.
int potrf_unblocked(int n, double* A, int lda) {
    for (int j = 0; j < n; ++j) {
        double* cj = A + (size_t)j * lda;
        for (int k = 0; k < j; ++k) {
            const double* ck = A + (size_t)k * lda;
            const double ljk = ck[j];
            for (int i = j; i < n; ++i) cj[i] -= ck[i] * ljk;
        }
        const double d = cj[j];
        if (!(d > 0.0) || !std::isfinite(d)) return j;
        const double r = std::sqrt(d);
        cj[j] = r;
        const double inv = 1.0 / r;
        for (int i = j + 1; i < n; ++i) cj[i] *= inv;
    }
    return -1;
}

// X[0:M, 0:n] <- X * L^{-T}, L lower n x n (column-major). Right-side transposed triangular solve.
void trsm_right_lt(int M, int n, const double* L, int ldl, double* X, int ldx, bool parallel) {
    if (M <= 0) return;
    const int RB = 256;
#pragma omp parallel for schedule(static) if (parallel && (double)M * n * n > 4e6)
    for (int r0 = 0; r0 < M; r0 += RB) {
        const int r1 = std::min(M, r0 + RB);
        for (int j = 0; j < n; ++j) {
            double* xj = X + (size_t)j * ldx;
            for (int k = 0; k < j; ++k) {
                const double ljk = L[j + (size_t)k * ldl];
                if (ljk == 0.0) continue;
                const double* xk = X + (size_t)k * ldx;
#pragma omp simd
                for (int i = r0; i < r1; ++i) xj[i] -= xk[i] * ljk;
            }
            const double inv = 1.0 / L[j + (size_t)j * ldl];
#pragma omp simd
            for (int i = r0; i < r1; ++i) xj[i] *= inv;
        }
    }
}

// Partial blocked Cholesky of an f x f front: factor the first s columns, leave the Schur
// complement in the trailing (f-s) x (f-s) lower block.
int partial_cholesky(int f, int s, double* A, int lda, bool parallel) {
    const int NB = 64;
    for (int k0 = 0; k0 < s; k0 += NB) {
        const int bs = std::min(NB, s - k0);
        double* Akk = A + k0 + (size_t)k0 * lda;
        int bad = potrf_unblocked(bs, Akk, lda);
        if (bad >= 0) return k0 + bad;
        const int rest = f - k0 - bs;
        if (rest <= 0) continue;
        double* Ark = A + (k0 + bs) + (size_t)k0 * lda;
        trsm_right_lt(rest, bs, Akk, lda, Ark, lda, parallel);
        double* Arr = A + (k0 + bs) + (size_t)(k0 + bs) * lda;
        gemm_nt_sub(rest, rest, bs, Ark, lda, Ark, lda, Arr, lda, true, 0, 0, parallel);
    }
    return -1;
}

// In place lower-triangular inverse (column-major, n x n); upper part zeroed.
void trtri_lower(int n, double* L, int ldl, bool parallel) {
    std::vector<double> diag(n);
    for (int j = 0; j < n; ++j) diag[j] = L[j + (size_t)j * ldl];
    std::vector<double> Lc((size_t)n * n);
    for (int j = 0; j < n; ++j)
        for (int i = j; i < n; ++i) Lc[i + (size_t)j * n] = L[i + (size_t)j * ldl];
#pragma omp parallel for schedule(dynamic, 8) if (parallel && n > 128)
    for (int j = 0; j < n; ++j) {
        double* x = L + (size_t)j * ldl;
        for (int i = 0; i < j; ++i) x[i] = 0.0;
        for (int i = j; i < n; ++i) x[i] = 0.0;
        x[j] = 1.0;
        for (int k = j; k < n; ++k) {
            const double xk = x[k] / diag[k];
            x[k] = xk;
            if (xk == 0.0) continue;
            const double* lk = Lc.data() + (size_t)k * n;
#pragma omp simd
            for (int i = k + 1; i < n; ++i) x[i] -= lk[i] * xk;
        }
    }
}

// W[0:b, 0:s] <- W * Linv (Linv lower s x s), in place via a temporary.
void trmm_right_lower(int b, int s, const double* Linv, int ldl, double* W, int ldw, bool parallel) {
    if (b <= 0 || s <= 0) return;
    std::vector<double> T((size_t)b * s);
    for (int j = 0; j < s; ++j) std::memcpy(T.data() + (size_t)j * b, W + (size_t)j * ldw, sizeof(double) * b);
#pragma omp parallel for schedule(dynamic, 4) if (parallel && (double)b * s * s > 2e6)
    for (int j = 0; j < s; ++j) {
        double* w = W + (size_t)j * ldw;
        for (int i = 0; i < b; ++i) w[i] = 0.0;
        for (int k = j; k < s; ++k) {
            const double lkj = Linv[k + (size_t)j * ldl];
            if (lkj == 0.0) continue;
            const double* t = T.data() + (size_t)k * b;
#pragma omp simd
            for (int i = 0; i < b; ++i) w[i] += t[i] * lkj;
        }
    }
}

}  // namespace

namespace {

// One front of the multifrontal factorisation: assemble node id's columns and its children's
// updates, factor, keep its own update matrix in upd[id] and store its panel. Returns 0, or 2 with
// the first failing column of a front that is not positive definite in *bad_col.
int factor_node(const Symbolic& S, int id, const double* lval, const double* mdiag, double l_scale, double m_scale,
                double* panel, std::vector<std::vector<double>>& upd, std::vector<double>& front, bool par,
                int* bad_col) {
    const int s = S.node_s[id], b = S.node_b[id], f = s + b;
    const int start = S.node_start[id], end = start + s;
    const int* B = S.bidx.data() + S.node_boff[id];
    front.assign((size_t)f * f, 0.0);
    // original entries of the s columns (rows >= column)
    for (int j = 0; j < s; ++j) {
        const int col = start + j;
        for (int e = S.rowptr[col]; e < S.rowptr[col + 1]; ++e) {
            const int row = S.colind[e];
            if (row < col) continue;
            int r;
            if (row < end) {
                r = row - start;
            } else {
                const int* it = std::lower_bound(B, B + b, row);
                if (it == B + b || *it != row) return 1;
                r = s + (int)(it - B);
            }
            double v = l_scale * lval[e];
            if (row == col) v += m_scale * mdiag[col];
            front[r + (size_t)j * f] += v;
        }
    }
    // extend-add children updates
    for (int ci = S.child_ptr[id]; ci < S.child_ptr[id + 1]; ++ci) {
        const int c = S.child_list[ci];
        const int bc = S.node_b[c];
        const int* cm = S.cmap.data() + S.node_boff[c];
        const std::vector<double>& U = upd[c];
        for (int jj = 0; jj < bc; ++jj) {
            double* fc = front.data() + (size_t)cm[jj] * f;
            const double* uc = U.data() + (size_t)jj * bc;
            for (int ii = jj; ii < bc; ++ii) fc[cm[ii]] += uc[ii];
        }
        std::vector<double>().swap(upd[c]);
    }
    const int bad = partial_cholesky(f, s, front.data(), f, par);
    if (bad >= 0) {
        *bad_col = bad;
        return 2;
    }
    if (b > 0) {
        upd[id].resize((size_t)b * b);
        for (int jj = 0; jj < b; ++jj)
            std::memcpy(upd[id].data() + (size_t)jj * b, front.data() + s + (size_t)(s + jj) * f, sizeof(double) * b);
    }
    trtri_lower(s, front.data(), f, par);
    trmm_right_lower(b, s, front.data(), f, front.data() + s, f, par);
    double* P = panel + S.node_poff[id];
    for (int j = 0; j < s; ++j) std::memcpy(P + (size_t)j * f, front.data() + (size_t)j * f, sizeof(double) * f);
    return 0;
}

}  // namespace

// The tree is factored level by level from the leaves (a node's children lie on lower levels).
// A level with many fronts spreads them over the threads, one front per thread; a level with few
// (the separators near the root, whose fronts are large) factors them one after another with the
// threads inside each front's dense kernels. Every front's arithmetic is the same either way, so
// the panels do not depend on the thread count.
int factor(const Symbolic& S, const double* lval, const double* mdiag, double l_scale, double m_scale,
           double* panel, std::string& err, int nthreads) {
    std::vector<std::vector<double>> upd(S.nnodes);
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const int nthr = omp_get_max_threads();
    for (int L = 0; L < S.nlevels; ++L) {
        const int lo = S.level_ptr[L], hi = S.level_ptr[L + 1];
        // first failure of the level in node order (the same whichever thread meets it)
        int fail_id = -1, fail_rc = 0, fail_col = -1;
        auto report = [&](int id, int rc, int col) {
            if (fail_id < 0 || id < fail_id) {
                fail_id = id;
                fail_rc = rc;
                fail_col = col;
            }
        };
        if (hi - lo >= 2 * nthr && nthr > 1) {
#pragma omp parallel
            {
                std::vector<double> front;
#pragma omp for schedule(dynamic, 1)
                for (int i = lo; i < hi; ++i) {
                    const int id = S.level_nodes[i];
                    int col = -1, rc;
                    try {  // no exception may leave the parallel region
                        rc = factor_node(S, id, lval, mdiag, l_scale, m_scale, panel, upd, front, false, &col);
                    } catch (const std::bad_alloc&) {
                        rc = 3;
                    }
                    if (rc) {
#pragma omp critical(fcw_factor_fail)
                        report(id, rc, col);
                    }
                }
            }
        } else {
            std::vector<double> front;
            for (int i = lo; i < hi; ++i) {
                const int id = S.level_nodes[i];
                const int s = S.node_s[id], f = s + S.node_b[id];
                int col = -1;
                const int rc = factor_node(S, id, lval, mdiag, l_scale, m_scale, panel, upd, front,
                                           (double)f * f * s > 4e6, &col);
                if (rc) report(id, rc, col);
            }
        }
        if (fail_id >= 0) {
            if (fail_rc == 1)
                err = "pattern entry outside front";
            else if (fail_rc == 3)
                err = "out of host memory for a front (node " + std::to_string(fail_id) + ")";
            else
                err = "matrix not positive definite (node " + std::to_string(fail_id) + ", column " +
                      std::to_string(fail_col) + ")";
            return fail_rc;
        }
    }
    return 0;
}

void panel_solve_host(const Symbolic& S, const double* panel, double* X, int k) {
    const int nn = S.nnodes;
    std::vector<double> U((size_t)S.urows * k, 0.0);
    std::vector<double> fbuf, obuf;
    // forward (postorder = children before parents)
    for (int id = 0; id < nn; ++id) {
        const int s = S.node_s[id], b = S.node_b[id], f = s + b, start = S.node_start[id];
        double* Ut = U.data() + (size_t)S.node_uoff[id] * k;
        std::fill(Ut, Ut + (size_t)b * k, 0.0);
        for (int ci = S.child_ptr[id]; ci < S.child_ptr[id + 1]; ++ci) {
            const int c = S.child_list[ci];
            const int* cm = S.cmap.data() + S.node_boff[c];
            const double* Uc = U.data() + (size_t)S.node_uoff[c] * k;
            for (int r = 0; r < S.node_b[c]; ++r) {
                const int t = cm[r];
                if (t < s) {
                    for (int j = 0; j < k; ++j) X[(size_t)(start + t) * k + j] -= Uc[(size_t)r * k + j];
                } else {
                    for (int j = 0; j < k; ++j) Ut[(size_t)(t - s) * k + j] += Uc[(size_t)r * k + j];
                }
            }
        }
        const double* P = panel + S.node_poff[id];
        obuf.assign((size_t)f * k, 0.0);
        for (int l = 0; l < s; ++l) {
            const double* xr = X + (size_t)(start + l) * k;
            const double* pc = P + (size_t)l * f;
            for (int i = 0; i < f; ++i) {
                const double pil = pc[i];
                if (pil == 0.0) continue;
                double* o = obuf.data() + (size_t)i * k;
                for (int j = 0; j < k; ++j) o[j] += pil * xr[j];
            }
        }
        for (int i = 0; i < s; ++i)
            std::memcpy(X + (size_t)(start + i) * k, obuf.data() + (size_t)i * k, sizeof(double) * k);
        for (int i = 0; i < b; ++i)
            for (int j = 0; j < k; ++j) Ut[(size_t)i * k + j] += obuf[(size_t)(s + i) * k + j];
    }
    // backward (reverse postorder)
    for (int id = nn - 1; id >= 0; --id) {
        const int s = S.node_s[id], b = S.node_b[id], f = s + b, start = S.node_start[id];
        const int* B = S.bidx.data() + S.node_boff[id];
        const double* P = panel + S.node_poff[id];
        fbuf.assign((size_t)s * k, 0.0);
        for (int l = 0; l < s; ++l) {
            const double* pc = P + (size_t)l * f;
            double* o = fbuf.data() + (size_t)l * k;
            for (int i = 0; i < s; ++i) {
                const double pil = pc[i];
                if (pil == 0.0) continue;
                const double* y = X + (size_t)(start + i) * k;
                for (int j = 0; j < k; ++j) o[j] += pil * y[j];
            }
            for (int i = 0; i < b; ++i) {
                const double pil = pc[s + i];
                const double* xb = X + (size_t)B[i] * k;
                for (int j = 0; j < k; ++j) o[j] -= pil * xb[j];
            }
        }
        std::memcpy(X + (size_t)start * k, fbuf.data(), sizeof(double) * (size_t)s * k);
    }
}

}  // namespace fcw
