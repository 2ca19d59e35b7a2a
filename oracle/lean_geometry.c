/*
 * Oracle (test infrastructure only): plain-C restatement of the reference's
 * per-source geometry kernels and fused normalised divergence. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this; the
 * product path never does.
 *
 * Follows, function by function:
 *   face_cdf                 lean_geometry.py:66-88
 *   grad_norm                lean_geometry.py:91-96
 *   build_distance_buckets   lean_geometry.py:117-180
 *   area_at                  lean_geometry.py:183-212
 *   perimeter_at             lean_geometry.py:215-244
 *   solve_radius             lean_geometry.py:247-299
 *   per_source_geometry      lean_geometry.py:302-340
 *   fused_normalized_div     fastcw_lean.py:416-442
 *
 * Scratch (dmax, order, long list) is caller-owned as in the reference; the
 * bin pointer / prefix arrays are allocated per call (the reference allocates
 * them per call too, lean_geometry.py:152-153).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline void face_cdf(double d0, double d1, double d2, double A, double r,
                            double* out_a, double* out_da) {
    double a = d0, b = d1, c = d2, t;
    if (a > b) { t = a; a = b; b = t; }
    if (b > c) { t = b; b = c; c = t; }
    if (a > b) { t = a; a = b; b = t; }
    if (r >= c) { *out_a = A; *out_da = 0.0; return; }
    if (r <= a) { *out_a = 0.0; *out_da = 0.0; return; }
    if (r < b) {
        double den = (b - a) * (c - a);
        double x = r - a;
        *out_a = A * x * x / den;
        *out_da = 2.0 * A * x / den;
        return;
    }
    double den = (c - a) * (c - b);
    double y = c - r;
    *out_a = A * (1.0 - y * y / den);
    *out_da = 2.0 * A * y / den;
}

static inline double grad_norm(double d0, double d1, double d2, double g11, double g12, double g22) {
    double al = d1 - d0, be = d2 - d0;
    double q = al * al * g11 + 2.0 * al * be * g12 + be * be * g22;
    return q > 0.0 ? sqrt(q) : 0.0;
}

typedef struct {
    int64_t nbins;
    int64_t* ptr;     /* nbins + 1 */
    double* prefix;   /* nbins + 1 */
    int64_t n_long;
} buckets_t;

static int build_buckets(const double* d, const int32_t* F, int64_t nf, const double* area,
                         double width, double* dmax_buf, int32_t* order_buf, int32_t* long_buf,
                         buckets_t* out) {
    double dglob = 0.0;
    int64_t n_long = 0;
    for (int64_t f = 0; f < nf; ++f) {
        double d0 = d[F[3 * f]], d1 = d[F[3 * f + 1]], d2 = d[F[3 * f + 2]];
        if (!(isfinite(d0) && isfinite(d1) && isfinite(d2))) {
            dmax_buf[f] = NAN;
            continue;
        }
        double lo = d0, hi = d0;
        if (d1 < lo) lo = d1;
        if (d1 > hi) hi = d1;
        if (d2 < lo) lo = d2;
        if (d2 > hi) hi = d2;
        dmax_buf[f] = hi;
        if (hi - lo > width) long_buf[n_long++] = (int32_t)f;
        if (hi > dglob) dglob = hi;
    }
    int64_t nbins = (int64_t)(dglob / width) + 2;
    int64_t* counts = (int64_t*)calloc((size_t)(nbins + 1), sizeof(int64_t));
    double* prefix = (double*)calloc((size_t)(nbins + 1), sizeof(double));
    int64_t* fill = (int64_t*)malloc((size_t)nbins * sizeof(int64_t));
    if (!counts || !prefix || !fill) { free(counts); free(prefix); free(fill); return -1; }
    for (int64_t f = 0; f < nf; ++f) {
        double h = dmax_buf[f];
        if (h != h) continue;
        int64_t k = (int64_t)(h / width);
        if (k < 0) k = 0; else if (k >= nbins) k = nbins - 1;
        counts[k + 1] += 1;
        prefix[k + 1] += area[f];
    }
    for (int64_t k = 0; k < nbins; ++k) {
        counts[k + 1] += counts[k];
        prefix[k + 1] += prefix[k];
    }
    memcpy(fill, counts, (size_t)nbins * sizeof(int64_t));
    for (int64_t f = 0; f < nf; ++f) {
        double h = dmax_buf[f];
        if (h != h) continue;
        int64_t k = (int64_t)(h / width);
        if (k < 0) k = 0; else if (k >= nbins) k = nbins - 1;
        order_buf[fill[k]++] = (int32_t)f;
    }
    free(fill);
    out->nbins = nbins;
    out->ptr = counts;
    out->prefix = prefix;
    out->n_long = n_long;
    return 0;
}

static void area_at(double r, const double* d, const int32_t* F, const double* area, double width,
                    const buckets_t* B, const int32_t* order, const double* dmax_buf,
                    const int32_t* long_buf, double* out_a, double* out_da) {
    if (!(r > 0.0)) { *out_a = 0.0; *out_da = 0.0; return; }
    int64_t kr = (int64_t)(r / width);
    if (kr >= B->nbins) { *out_a = B->prefix[B->nbins]; *out_da = 0.0; return; }
    double A = B->prefix[kr], dA = 0.0, a, da;
    int64_t kend = kr + 2;
    if (kend > B->nbins) kend = B->nbins;
    for (int64_t j = B->ptr[kr]; j < B->ptr[kend]; ++j) {
        int32_t f = order[j];
        face_cdf(d[F[3 * f]], d[F[3 * f + 1]], d[F[3 * f + 2]], area[f], r, &a, &da);
        A += a; dA += da;
    }
    for (int64_t j = 0; j < B->n_long; ++j) {
        int32_t f = long_buf[j];
        if ((int64_t)(dmax_buf[f] / width) >= kend) {
            face_cdf(d[F[3 * f]], d[F[3 * f + 1]], d[F[3 * f + 2]], area[f], r, &a, &da);
            A += a; dA += da;
        }
    }
    *out_a = A;
    *out_da = dA;
}

static double perimeter_at(double r, const double* d, const int32_t* F, const double* area,
                           const double* g11, const double* g12, const double* g22, double width,
                           const buckets_t* B, const int32_t* order, const double* dmax_buf,
                           const int32_t* long_buf) {
    if (!(r > 0.0)) return 0.0;
    int64_t kr = (int64_t)(r / width);
    if (kr >= B->nbins) return 0.0;
    int64_t kend = kr + 2;
    if (kend > B->nbins) kend = B->nbins;
    double P = 0.0, a, da;
    for (int64_t j = B->ptr[kr]; j < B->ptr[kend]; ++j) {
        int32_t f = order[j];
        double d0 = d[F[3 * f]], d1 = d[F[3 * f + 1]], d2 = d[F[3 * f + 2]];
        face_cdf(d0, d1, d2, area[f], r, &a, &da);
        if (da > 0.0) P += da * grad_norm(d0, d1, d2, g11[f], g12[f], g22[f]);
    }
    for (int64_t j = 0; j < B->n_long; ++j) {
        int32_t f = long_buf[j];
        if ((int64_t)(dmax_buf[f] / width) >= kend) {
            double d0 = d[F[3 * f]], d1 = d[F[3 * f + 1]], d2 = d[F[3 * f + 2]];
            face_cdf(d0, d1, d2, area[f], r, &a, &da);
            if (da > 0.0) P += da * grad_norm(d0, d1, d2, g11[f], g12[f], g22[f]);
        }
    }
    return P;
}

static double solve_radius(double target, const double* d, const int32_t* F, const double* area,
                           double width, const buckets_t* B, const int32_t* order,
                           const double* dmax_buf, const int32_t* long_buf, double rtol,
                           int max_iter, int64_t* n_eval_out) {
    double total = B->prefix[B->nbins];
    *n_eval_out = 0;
    if (target != target) return NAN;
    if (target <= 0.0) return 0.0;
    if (target > total * (1.0 + 1e-12)) return NAN;
    int64_t lo_i = 0, hi_i = B->nbins;
    while (lo_i < hi_i) {
        int64_t mid = (lo_i + hi_i) / 2;
        if (B->prefix[mid] >= target) hi_i = mid; else lo_i = mid + 1;
    }
    double r_hi = (double)lo_i * width;
    double r_lo = lo_i >= 2 ? (double)(lo_i - 2) * width : 0.0;
    int64_t n_eval = 0;
    double A, dA;
    if (B->n_long > 0) {
        while (r_lo > 0.0) {
            area_at(r_lo, d, F, area, width, B, order, dmax_buf, long_buf, &A, &dA);
            n_eval++;
            if (A < target) break;
            r_hi = r_lo;
            r_lo = r_lo - 2.0 * width;
            if (r_lo < 0.0) r_lo = 0.0;
        }
    }
    double r = 0.5 * (r_lo + r_hi);
    for (int it = 0; it < max_iter; ++it) {
        area_at(r, d, F, area, width, B, order, dmax_buf, long_buf, &A, &dA);
        n_eval++;
        double err = A - target;
        if (fabs(err) <= rtol * target) { *n_eval_out = n_eval; return r; }
        if (err > 0.0) r_hi = r; else r_lo = r;
        double rn = 0.5 * (r_lo + r_hi);
        if (dA > 0.0) {
            double cand = r - err / dA;
            if (r_lo < cand && cand < r_hi) rn = cand;
        }
        if (r_hi - r_lo <= 1e-13 * (1.0 + r_hi)) { *n_eval_out = n_eval; return rn; }
        r = rn;
    }
    *n_eval_out = n_eval;
    return r;
}

/* Evaluate area_at / perimeter_at for a list of radii on one field (test helper). */
int oc_area_perimeter(const double* d, const int32_t* F, int64_t nf, const double* area,
                      const double* g11, const double* g12, const double* g22, double width,
                      const double* radii, int64_t nr, double* out_area, double* out_darea,
                      double* out_perim, int64_t* out_nlong) {
    double* dmax_buf = (double*)malloc((size_t)nf * sizeof(double));
    int32_t* order = (int32_t*)malloc((size_t)nf * sizeof(int32_t));
    int32_t* longb = (int32_t*)malloc((size_t)nf * sizeof(int32_t));
    buckets_t B;
    if (!dmax_buf || !order || !longb || build_buckets(d, F, nf, area, width, dmax_buf, order, longb, &B)) {
        free(dmax_buf); free(order); free(longb);
        return -1;
    }
    for (int64_t i = 0; i < nr; ++i) {
        area_at(radii[i], d, F, area, width, &B, order, dmax_buf, longb, &out_area[i], &out_darea[i]);
        out_perim[i] = perimeter_at(radii[i], d, F, area, g11, g12, g22, width, &B, order, dmax_buf, longb);
    }
    *out_nlong = B.n_long;
    free(B.ptr); free(B.prefix); free(dmax_buf); free(order); free(longb);
    return 0;
}

/* solve_radius for a list of targets on one field (test helper). */
int oc_solve_radius(const double* d, const int32_t* F, int64_t nf, const double* area, double width,
                    const double* targets, int64_t nt, double rtol, int max_iter, double* out_r,
                    int64_t* out_eval) {
    double* dmax_buf = (double*)malloc((size_t)nf * sizeof(double));
    int32_t* order = (int32_t*)malloc((size_t)nf * sizeof(int32_t));
    int32_t* longb = (int32_t*)malloc((size_t)nf * sizeof(int32_t));
    buckets_t B;
    if (!dmax_buf || !order || !longb || build_buckets(d, F, nf, area, width, dmax_buf, order, longb, &B)) {
        free(dmax_buf); free(order); free(longb);
        return -1;
    }
    for (int64_t i = 0; i < nt; ++i)
        out_r[i] = solve_radius(targets[i], d, F, area, width, &B, order, dmax_buf, longb, rtol,
                                max_iter, &out_eval[i]);
    free(B.ptr); free(B.prefix); free(dmax_buf); free(order); free(longb);
    return 0;
}

/* per_source_geometry for `ns` fields stored row-major (ns x n). Returns total band evaluations
 * or -1 on allocation failure. Outputs are (ns x S), (ns x S), (ns x (S-1)q), (ns x (S-1)q). */
int64_t oc_per_source_geometry(const double* D, int64_t ns, int64_t n, const int32_t* F, int64_t nf,
                               const double* area, const double* g11, const double* g12,
                               const double* g22, double width, const double* targets, int64_t S,
                               const double* fracs, int64_t q, double rtol, int max_iter,
                               double* radii_out, double* perim_out, double* extra_r_out,
                               double* extra_a_out, int64_t* n_eval_out) {
    double* dmax_buf = (double*)malloc((size_t)nf * sizeof(double));
    int32_t* order = (int32_t*)malloc((size_t)nf * sizeof(int32_t));
    int32_t* longb = (int32_t*)malloc((size_t)nf * sizeof(int32_t));
    if (!dmax_buf || !order || !longb) { free(dmax_buf); free(order); free(longb); return -1; }
    int64_t total_eval = 0;
    for (int64_t s_i = 0; s_i < ns; ++s_i) {
        const double* d = D + s_i * n;
        double* ro = radii_out + s_i * S;
        double* po = perim_out + s_i * S;
        double* er = extra_r_out + s_i * (S > 0 ? (S - 1) * q : 0);
        double* ea = extra_a_out + s_i * (S > 0 ? (S - 1) * q : 0);
        buckets_t B;
        if (build_buckets(d, F, nf, area, width, dmax_buf, order, longb, &B)) {
            free(dmax_buf); free(order); free(longb);
            return -1;
        }
        int64_t n_eval = 0, ke;
        for (int64_t s = 0; s < S; ++s) {
            double r = solve_radius(targets[s], d, F, area, width, &B, order, dmax_buf, longb, rtol,
                                    max_iter, &ke);
            n_eval += ke;
            ro[s] = r;
            po[s] = (r == r) ? perimeter_at(r, d, F, area, g11, g12, g22, width, &B, order, dmax_buf, longb)
                             : NAN;
        }
        for (int64_t s = 0; s + 1 < S; ++s) {
            double r0 = ro[s], r1 = ro[s + 1];
            for (int64_t j = 0; j < q; ++j) {
                int64_t idx = s * q + j;
                if (r0 == r0 && r1 == r1 && r0 > 0.0 && r1 > r0) {
                    double rr = exp(log(r0) + fracs[j] * (log(r1) - log(r0)));
                    double a, da;
                    area_at(rr, d, F, area, width, &B, order, dmax_buf, longb, &a, &da);
                    er[idx] = rr;
                    ea[idx] = a;
                    n_eval++;
                } else {
                    er[idx] = NAN;
                    ea[idx] = NAN;
                }
            }
        }
        if (n_eval_out) n_eval_out[s_i] = n_eval;
        total_eval += n_eval;
        free(B.ptr);
        free(B.prefix);
    }
    free(dmax_buf); free(order); free(longb);
    return total_eval;
}

/* fastcw_lean.py:416-442: out (n x k, C order) = sum_f A_f (X_f . grad B_i), X_f = -grad u/|grad u|. */
void oc_fused_normalized_divergence(const double* u, int64_t n, int64_t k, const int32_t* F, int64_t nf,
                                    const double* area, const double* gb0, const double* gb1,
                                    const double* gb2, double* out) {
    memset(out, 0, (size_t)(n * k) * sizeof(double));
    for (int64_t f = 0; f < nf; ++f) {
        int64_t i0 = F[3 * f], i1 = F[3 * f + 1], i2 = F[3 * f + 2];
        double A = area[f];
        const double* b0 = gb0 + 3 * f;
        const double* b1 = gb1 + 3 * f;
        const double* b2 = gb2 + 3 * f;
        for (int64_t j = 0; j < k; ++j) {
            double du1 = u[i1 * k + j] - u[i0 * k + j];
            double du2 = u[i2 * k + j] - u[i0 * k + j];
            double gx = du1 * b1[0] + du2 * b2[0];
            double gy = du1 * b1[1] + du2 * b2[1];
            double gz = du1 * b1[2] + du2 * b2[2];
            double nn = sqrt(gx * gx + gy * gy + gz * gz);
            if (!(nn > 1e-300)) continue;
            double s = -A / nn;
            out[i0 * k + j] += s * (gx * b0[0] + gy * b0[1] + gz * b0[2]);
            out[i1 * k + j] += s * (gx * b1[0] + gy * b1[1] + gz * b1[2]);
            out[i2 * k + j] += s * (gx * b2[0] + gy * b2[1] + gz * b2[2]);
        }
    }
}
