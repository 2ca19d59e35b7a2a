// Per-source wiring metrics on gfx950: one 1024-thread workgroup per distance
// field (core_analysis.py:1558-1609, lean_geometry.py:117-357).
//
//   phase 0  vertex pass: health screen, MSD (unweighted / area-weighted),
//            distance to boundary, largest finite distance
//   phase 1  face pass: area histograms of per-face dmax and dmin. For any r on
//            the bin grid, sum(dmax < r) <= A(r) <= sum(dmin < r), which brackets
//            every target radius to a few bins before any exact evaluation.
//   phase 2  face pass: the faces that can change A(r) inside each bracket
//            ("band") are copied, pre-sorted, into contiguous per-scale arrays
//   phase 3  safeguarded Newton on the exact piecewise-quadratic A(r) (one wave
//            per scale), perimeter by the coarea formula at the root
//   phase 4  face pass: exact areas at the log-spaced samples between radii
//   phase 5  outputs and the cleaned (radius, area) samples
//
// Face passes are coalesced (consecutive lanes take consecutive faces); LDS
// histogram and accumulator updates are summed per run of equal keys inside a
// wave first, so one atomic is issued per run.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "fcw_device.h"

namespace fcw {
namespace {

constexpr int kWave = 64;
constexpr int NT = kMetricThreads;
constexpr int NW = NT / kWave;

__device__ inline int lane_id() { return threadIdx.x & (kWave - 1); }

__device__ inline double wave_sum(double v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ inline double wave_max(double v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ inline double wave_min(double v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ inline int wave_sum_i(int v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ inline int wave_min_i(int v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ inline int wave_max_i(int v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, kWave));
    return v;
}

__device__ inline uint64_t lanes_below(int lane) { return (1ull << lane) - 1ull; }
__device__ inline uint64_t lanes_upto(int lane) { return lane == kWave - 1 ? ~0ull : ((1ull << (lane + 1)) - 1ull); }

// H[key] += v for every lane with key >= 0. Runs of equal consecutive keys are summed by a
// segmented scan and the last lane of each run issues the LDS atomic. All 64 lanes must call.
__device__ inline void keyed_add(double* H, int key, double v) {
    const int lane = lane_id();
    const int prev = __shfl_up(key, 1, kWave);
    const bool head = (lane == 0) || (prev != key);
    const uint64_t hm = __ballot(head);
    const int rid = __popcll(hm & lanes_upto(lane));
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const double vv = __shfl_up(v, o, kWave);
        const int rr = __shfl_up(rid, o, kWave);
        if (lane >= o && rr == rid) v += vv;
    }
    const bool tail = (lane == kWave - 1) || ((hm >> (lane + 1)) & 1ull);
    if (tail && key >= 0) atomicAdd(&H[key], v);
}

__device__ inline void sort3(double& a, double& b, double& c) {
    double t;
    if (a > b) { t = a; a = b; b = t; }
    if (b > c) { t = b; b = c; c = t; }
    if (a > b) { t = a; a = b; b = t; }
}

// Area of {d <= r} in a linear face with sorted vertex values a <= b <= c, and its derivative
// (lean_geometry.py:66-88 after its sort).
__device__ inline void cdf_sorted(double a, double b, double c, double A, double r, double& ao, double& dao) {
    if (r >= c) { ao = A; dao = 0.0; return; }
    if (r <= a) { ao = 0.0; dao = 0.0; return; }
    if (r < b) {
        const double den = (b - a) * (c - a);
        const double x = r - a;
        ao = A * x * x / den;
        dao = 2.0 * A * x / den;
        return;
    }
    const double den = (c - a) * (c - b);
    const double y = c - r;
    ao = A * (1.0 - y * y / den);
    dao = 2.0 * A * y / den;
}

__device__ inline double grad_norm(double d0, double d1, double d2, double g11, double g12, double g22) {
    const double al = d1 - d0, be = d2 - d0;
    const double q = al * al * g11 + 2.0 * al * be * g12 + be * be * g22;
    return q > 0.0 ? sqrt(q) : 0.0;
}

__device__ inline int bin_of(double v, double wb, int nbins) {
    int k = (int)(v / wb);
    if (k < 0) k = 0;
    if (k >= nbins) k = nbins - 1;
    return k;
}

// Faces are visited UNR per lane per step: all index loads, then all gathers, then the work, so a
// wave keeps 3 UNR independent gathers in flight instead of one dependent chain per face.
constexpr int UNR = 4;

struct FaceBatch {
    int f[UNR];
    double d0[UNR], d1[UNR], d2[UNR];
};

// CONSEC = false: lane takes faces base + u*64 + lane; true: faces base + UNR*lane + u.
template <bool CONSEC = false>
__device__ inline void load_faces(const int* __restrict__ faces, const double* __restrict__ d, int m, int base,
                                  int lane, FaceBatch& fb) {
    int i0[UNR], i1[UNR], i2[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
        const int f = CONSEC ? base + UNR * lane + u : base + u * kWave + lane;
        fb.f[u] = f;
        const int fs = f < m ? f : 0;  // an in-bounds face; the caller ignores f >= m
        i0[u] = faces[3 * fs];
        i1[u] = faces[3 * fs + 1];
        i2[u] = faces[3 * fs + 2];
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
        fb.d0[u] = d[i0[u]];
        fb.d1[u] = d[i1[u]];
        fb.d2[u] = d[i2[u]];
    }
}

struct Red {
    double d[NW][4];
    int i[NW][4];
};

// Block reduction of 4 doubles + 4 ints; slot `smax` of v is a max, slot `smin` a min, the rest sums.
__device__ void block_reduce(Red& sm, double (&v)[4], int (&iv)[4], int smax, int smin) {
    const int lane = lane_id(), w = threadIdx.x / kWave;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        v[q] = (q == smax) ? wave_max(v[q]) : (q == smin) ? wave_min(v[q]) : wave_sum(v[q]);
        iv[q] = wave_sum_i(iv[q]);
    }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            sm.d[w][q] = v[q];
            sm.i[w][q] = iv[q];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double x = sm.d[0][q];
        int xi = sm.i[0][q];
        for (int ww = 1; ww < NW; ++ww) {
            const double y = sm.d[ww][q];
            x = (q == smax) ? fmax(x, y) : (q == smin) ? fmin(x, y) : x + y;
            xi += sm.i[ww][q];
        }
        v[q] = x;
        iv[q] = xi;
    }
    __syncthreads();
}

__global__ __launch_bounds__(NT) void metrics_kernel(MeshDev mesh, MetricsDev md, const double* __restrict__ Dt,
                                                      const int* __restrict__ src_pos, MetricsOut out,
                                                      int health_mode) {
    const int j = blockIdx.x;
    const int tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
    const int n = mesh.n, m = mesh.m;
    const int* __restrict__ faces = mesh.faces;
    const double* __restrict__ d = Dt + (int64_t)j * n;
    const int S = md.S, q = md.q, NE = (S > 0 ? (S - 1) * q : 0);
    double* scal = out.scal + (int64_t)j * kSlotCount;
    double* radii = out.radii + (int64_t)j * S;
    double* perim = out.perim + (int64_t)j * S;
    double* er = out.extra_r + (int64_t)j * NE;
    double* ea = out.extra_a + (int64_t)j * NE;
    const int cap = md.band_cap;
    double* band = md.band + (int64_t)j * S * 5 * cap;

    __shared__ Red red;
    __shared__ double Hmax[kMaxBins + 1];
    __shared__ double Hmin[kMaxBins + 1];
    __shared__ int lo_k[kMaxScales], hi_k[kMaxScales], cnt[kMaxScales], act[kMaxScales], ev_out[kMaxScales];
    __shared__ double hi_r[kMaxScales], r_out[kMaxScales], p_out[kMaxScales];
    __shared__ double rho[kMaxExtras], vr[kMaxExtras], racc[kMaxExtras];
    __shared__ int rk[kMaxExtras], ridx[kMaxExtras];
    __shared__ double pair_r[kMaxScales + kMaxExtras], pair_a[kMaxScales + kMaxExtras];
    __shared__ int s_nv, s_nbins;
    __shared__ double s_wb;

    // ---- phase 0: health screen, MSD, distance to boundary -------------------------------
    const int sp = src_pos[j];
    const double xs = mesh.V[3 * sp], ys = mesh.V[3 * sp + 1], zs = mesh.V[3 * sp + 2];
    double v4[4] = {0.0, 0.0, 0.0, 0.0};  // sum d, sum d*w, sum w, max finite d
    int i4[4] = {0, 0, 0, 0};             // d < -h, chord violations, valid, non-finite
    const double h = md.health_h;
    for (int p0 = tid; p0 < n; p0 += NT * UNR) {
        double dv[UNR], vx[UNR], vy[UNR], vz[UNR], vw[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const int p = p0 + u * NT;
            const int ps = p < n ? p : 0;
            dv[u] = d[ps];
            vx[u] = mesh.V[3 * ps];
            vy[u] = mesh.V[3 * ps + 1];
            vz[u] = mesh.V[3 * ps + 2];
            vw[u] = md.vw[ps];
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            if (p0 + u * NT >= n) continue;
            const double x = dv[u];
            if (!isfinite(x)) {
                i4[3] += 1;
                continue;
            }
            const double dx = vx[u] - xs, dy = vy[u] - ys, dz = vz[u] - zs;
            const double chord = sqrt(dx * dx + dy * dy + dz * dz);
            if (x < -h) i4[0] += 1;
            if (x < 0.9 * chord - h) i4[1] += 1;
            if (x > md.eps) {
                i4[2] += 1;
                v4[0] += x;
                v4[1] += x * vw[u];
                v4[2] += vw[u];
            }
            v4[3] = fmax(v4[3], x);
        }
    }
    block_reduce(red, v4, i4, 3, -1);
    const double dglob = v4[3];
    double b4[4] = {INFINITY, 0.0, 0.0, 0.0};
    int bi4[4] = {0, 0, 0, 0};
    for (int i = tid; i < md.nb; i += NT) {
        const double dv = d[md.boundary[i]];
        if (isfinite(dv)) {
            b4[0] = fmin(b4[0], dv);
            bi4[0] += 1;
        }
    }
    block_reduce(red, b4, bi4, -1, 0);
    const double neg_frac = i4[3] ? NAN : (double)i4[0] / n;
    const double viol_frac = i4[3] ? NAN : (double)i4[1] / n;
    const bool bad = (i4[3] > 0) || neg_frac > 1e-3 || viol_frac > 1e-3;
    const double dist_b = (md.nb > 0) ? (bi4[0] > 0 ? b4[0] : NAN) : INFINITY;
    if (health_mode && bad) {
        if (tid == 0) {
            scal[kSlotFlag] = 1.0;
            scal[kSlotNeg] = neg_frac;
            scal[kSlotViol] = viol_frac;
            scal[kSlotDistB] = NAN;
            scal[kSlotMsdU] = NAN;
            scal[kSlotMsdW] = NAN;
            scal[kSlotNEval] = 0.0;
            scal[kSlotNSamples] = 0.0;
            scal[kSlotOverflow] = 0.0;
        }
        for (int s = tid; s < S; s += NT) {
            radii[s] = NAN;
            perim[s] = NAN;
        }
        for (int e = tid; e < NE; e += NT) {
            er[e] = NAN;
            ea[e] = NAN;
        }
        return;
    }

    // ---- phase 1: dmax / dmin area histograms ----------------------------------------------
    if (tid == 0) {
        // Bins only steer the work, never the result: a quarter of the reference bucket width,
        // widened when the field is too long for the LDS histogram.
        double wb = 0.25 * md.width;
        if (dglob / wb + 2.0 > (double)kMaxBins) wb = dglob / (double)(kMaxBins - 3);
        int nbins = (int)(dglob / wb) + 2;
        if (nbins > kMaxBins) nbins = kMaxBins;
        s_wb = wb;
        s_nbins = nbins;
    }
    for (int i = tid; i <= kMaxBins; i += NT) {
        Hmax[i] = 0.0;
        Hmin[i] = 0.0;
    }
    __syncthreads();
    const double wb = s_wb;
    const int nbins = s_nbins;
    for (int base = wid * kWave * UNR; base < m; base += NT * UNR) {
        FaceBatch fb;
        load_faces<true>(faces, d, m, base, lane, fb);
        // The lane's UNR faces are consecutive, so mostly share bins: merge equal neighbours
        // first and hand the wave only the lane's distinct (bin, area) runs.
        int kh[UNR], kl[UNR];
        double A[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            kh[u] = -1;
            kl[u] = -1;
            A[u] = 0.0;
            const double d0 = fb.d0[u], d1 = fb.d1[u], d2 = fb.d2[u];
            if (fb.f[u] < m && isfinite(d0) && isfinite(d1) && isfinite(d2)) {
                A[u] = mesh.larea[fb.f[u]];
                kh[u] = bin_of(fmax(d0, fmax(d1, d2)), wb, nbins);
                kl[u] = bin_of(fmin(d0, fmin(d1, d2)), wb, nbins);
            }
        }
#pragma unroll
        for (int which = 0; which < 2; ++which) {
            int* key = which == 0 ? kh : kl;
            double* H = which == 0 ? Hmax : Hmin;
            // runs of equal keys within the lane: run r starts at u = first index of the run
            double run_sum[UNR];
            int run_key[UNR];
            int nrun = 0;
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                if (key[u] < 0) continue;
                if (nrun > 0 && run_key[nrun - 1] == key[u]) {
                    run_sum[nrun - 1] += A[u];
                } else {
                    run_key[nrun] = key[u];
                    run_sum[nrun] = A[u];
                    ++nrun;
                }
            }
            const int rounds = wave_max_i(nrun);
            for (int r = 0; r < rounds; ++r) keyed_add(H, r < nrun ? run_key[r] : -1, r < nrun ? run_sum[r] : 0.0);
        }
    }
    __syncthreads();
    // exclusive prefix sums in place: H[k] = sum over bins < k, k = 0..nbins
    if (wid < 2) {
        double* H = (wid == 0) ? Hmax : Hmin;
        const int per = (nbins + kWave - 1) / kWave;
        const int b0 = lane * per, b1 = min(nbins, b0 + per);
        double loc = 0.0;
        for (int b = b0; b < b1; ++b) loc += H[b];
        double inc = loc;
        for (int o = 1; o < kWave; o <<= 1) {
            const double y = __shfl_up(inc, o, kWave);
            if (lane >= o) inc += y;
        }
        double run = inc - loc;
        for (int b = b0; b < b1; ++b) {
            const double hv = H[b];
            H[b] = run;
            run += hv;
        }
        if (lane == kWave - 1) H[nbins] = inc;
    }
    __syncthreads();
    const double total = Hmax[nbins];
    if (tid < S) {
        const int s = tid;
        const double T = md.targets[s];
        int a;
        double r0v = NAN;
        if (T != T) {
            a = 2;
        } else if (T <= 0.0) {
            a = 0;
            r0v = 0.0;
        } else if (T > total * (1.0 + 1e-12)) {
            a = 2;
        } else {
            a = 1;
        }
        act[s] = a;
        r_out[s] = r0v;
        p_out[s] = NAN;
        ev_out[s] = 0;
        cnt[s] = 0;
        if (a == 1) {
            int lo = 0, hi = nbins;  // first k with Hmax[k] >= T  (A(k wb) >= Hmax[k])
            if (Hmax[nbins] < T) {
                lo = nbins;
            } else {
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (Hmax[mid] >= T) hi = mid; else lo = mid + 1;
                }
            }
            const int khi = lo;
            int l2 = 0, h2 = khi;  // last k <= khi with Hmin[k] < T  (A(k wb) <= Hmin[k])
            while (l2 < h2) {
                const int mid = (l2 + h2 + 1) >> 1;
                if (Hmin[mid] < T) l2 = mid; else h2 = mid - 1;
            }
            lo_k[s] = l2;
            hi_k[s] = khi;
            hi_r[s] = (double)khi * wb;
        } else if (a == 0) {
            lo_k[s] = 0;
            hi_k[s] = 0;
            hi_r[s] = -INFINITY;
        } else {
            lo_k[s] = 0x7fffffff;
            hi_k[s] = 0x7fffffff;
            hi_r[s] = INFINITY;
        }
    }
    __syncthreads();

    // ---- phase 2: band faces per scale, pre-sorted, contiguous -----------------------------
    for (int base0 = wid * kWave * UNR; base0 < m; base0 += NT * UNR) {
        FaceBatch fbat;
        load_faces(faces, d, m, base0, lane, fbat);
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const int f = fbat.f[u];
            int sa = S, sb = -1;
            double fa = 0, fb = 0, fc = 0, A = 0, gn = 0;
            const double d0 = fbat.d0[u], d1 = fbat.d1[u], d2 = fbat.d2[u];
            if (f < m && isfinite(d0) && isfinite(d1) && isfinite(d2)) {
                fa = d0;
                fb = d1;
                fc = d2;
                sort3(fa, fb, fc);
                const int bh = bin_of(fc, wb, nbins);
                int lo = 0, hi = S;  // first s with hi_r[s] > fa
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (hi_r[mid] > fa) hi = mid; else lo = mid + 1;
                }
                sa = lo;
                lo = 0;
                hi = S;  // first s with lo_k[s] > bh
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (lo_k[mid] > bh) hi = mid; else lo = mid + 1;
                }
                sb = lo - 1;
                if (sa <= sb) {
                    A = mesh.larea[f];
                    gn = grad_norm(d0, d1, d2, mesh.g11[f], mesh.g12[f], mesh.g22[f]);
                }
            }
            if (__ballot(sa <= sb) == 0ull) continue;  // no band face in this wave
            const int smin = wave_min_i(sa <= sb ? sa : S);
            const int smax = wave_max_i(sa <= sb ? sb : -1);
            for (int s = smin; s <= smax; ++s) {
                const bool pred = (sa <= s && s <= sb);
                const uint64_t mask = __ballot(pred);
                if (mask == 0ull) continue;
                const int leader = __ffsll((long long)mask) - 1;
                int slot0 = 0;
                if (lane == leader) slot0 = atomicAdd(&cnt[s], __popcll(mask));
                slot0 = __shfl(slot0, leader, kWave);
                if (pred) {
                    const int slot = slot0 + __popcll(mask & lanes_below(lane));
                    if (slot < cap) {
                        double* bs = band + (int64_t)s * 5 * cap;
                        bs[slot] = fa;
                        bs[cap + slot] = fb;
                        bs[2 * cap + slot] = fc;
                        bs[3 * cap + slot] = A;
                        bs[4 * cap + slot] = gn;
                    }
                }
            }
        }
    }
    __syncthreads();

    // ---- phase 3: safeguarded Newton per scale (one wave per scale) ------------------------
    for (int s = wid; s < S; s += NW) {
        if (act[s] != 1) continue;
        const double T = md.targets[s];
        const int count = cnt[s];
        const bool full_scan = count > cap;
        const double* bs = band + (int64_t)s * 5 * cap;
        const int klo = lo_k[s], khi = hi_k[s];
        const double Ain = Hmax[klo];
        const double br_lo = (double)klo * wb, br_hi = hi_r[s];
        // initial guess: interpolate the mid-point of the histogram bounds inside the bracket
        double r = 0.5 * (br_lo + br_hi);
        for (int k = klo; k < khi; ++k) {
            const double e0 = 0.5 * (Hmax[k] + Hmin[k]), e1 = 0.5 * (Hmax[k + 1] + Hmin[k + 1]);
            if (e0 <= T && T < e1) {
                const double cand = ((double)k + (T - e0) / (e1 - e0)) * wb;
                if (cand > br_lo && cand < br_hi) r = cand;
                break;
            }
        }
        double r_lo = br_lo, r_hi = br_hi;
        auto eval = [&](double rr, double& A_out, double& dA_out, double& P_out) {
            double sA = 0.0, sdA = 0.0, sP = 0.0;
            if (!full_scan) {
                for (int i = lane; i < count; i += kWave) {
                    double a, da;
                    cdf_sorted(bs[i], bs[cap + i], bs[2 * cap + i], bs[3 * cap + i], rr, a, da);
                    sA += a;
                    sdA += da;
                    if (da > 0.0) sP += da * bs[4 * cap + i];
                }
            } else {
                for (int f = lane; f < m; f += kWave) {
                    const double d0 = d[faces[3 * f]], d1 = d[faces[3 * f + 1]], d2 = d[faces[3 * f + 2]];
                    if (!(isfinite(d0) && isfinite(d1) && isfinite(d2))) continue;
                    double fa = d0, fb = d1, fc = d2;
                    sort3(fa, fb, fc);
                    if (!(bin_of(fc, wb, nbins) >= klo && fa < br_hi)) continue;
                    double a, da;
                    cdf_sorted(fa, fb, fc, mesh.larea[f], rr, a, da);
                    sA += a;
                    sdA += da;
                    if (da > 0.0) sP += da * grad_norm(d0, d1, d2, mesh.g11[f], mesh.g12[f], mesh.g22[f]);
                }
            }
            A_out = Ain + wave_sum(sA);
            dA_out = wave_sum(sdA);
            P_out = wave_sum(sP);
        };
        int n_eval = 0;
        double result = r, Pres = 0.0;
        bool have_p = false;
        for (int it = 0; it < md.max_iter; ++it) {
            double A, dA, Pv;
            eval(r, A, dA, Pv);
            ++n_eval;
            const double err = A - T;
            if (fabs(err) <= md.rtol * T) {
                result = r;
                Pres = Pv;
                have_p = true;
                break;
            }
            if (err > 0.0) r_hi = r; else r_lo = r;
            double rn = 0.5 * (r_lo + r_hi);
            if (dA > 0.0) {
                const double cand = r - err / dA;
                if (r_lo < cand && cand < r_hi) rn = cand;
            }
            if (r_hi - r_lo <= 1e-13 * (1.0 + r_hi)) {
                result = rn;
                break;
            }
            r = rn;
            result = r;
        }
        if (!have_p) {
            double A, dA;
            eval(result, A, dA, Pres);
            ++n_eval;
        }
        if (lane == 0) {
            r_out[s] = result;
            p_out[s] = Pres;
            ev_out[s] = n_eval;
        }
    }
    __syncthreads();

    // ---- phase 4: supplementary samples between consecutive radii --------------------------
    if (tid < NE) {
        const int s = tid / q, jj = tid - s * q;
        const double r0 = r_out[s], r1 = r_out[s + 1];
        double rr = NAN;
        if (r0 == r0 && r1 == r1 && r0 > 0.0 && r1 > r0) rr = exp(log(r0) + md.fracs[jj] * (log(r1) - log(r0)));
        rho[tid] = rr;
    }
    __syncthreads();
    if (tid == 0) {
        int nv = 0;
        for (int e = 0; e < NE; ++e) {
            const double rr = rho[e];
            if (rr == rr) {
                ridx[nv] = e;
                vr[nv] = rr;
                rk[nv] = (int)(rr / wb);
                ++nv;
            }
        }
        s_nv = nv;
    }
    __syncthreads();
    const int nv = s_nv;
    for (int v = tid; v < nv; v += NT) racc[v] = (rk[v] >= nbins) ? total : Hmax[rk[v]];
    __syncthreads();
    if (nv > 0) {
        for (int base0 = wid * kWave * UNR; base0 < m; base0 += NT * UNR) {
            FaceBatch fbat;
            load_faces(faces, d, m, base0, lane, fbat);
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                const int f = fbat.f[u];
                int va = nv, vb = -1;
                double fa = fbat.d0[u], fb = fbat.d1[u], fc = fbat.d2[u], A = 0;
                if (f < m && isfinite(fa) && isfinite(fb) && isfinite(fc)) {
                    sort3(fa, fb, fc);
                    const int bh = bin_of(fc, wb, nbins);
                    int lo = 0, hi = nv;  // first v with vr > fa
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (vr[mid] > fa) hi = mid; else lo = mid + 1;
                    }
                    va = lo;
                    lo = 0;
                    hi = nv;  // first v with rk > bh
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (rk[mid] > bh) hi = mid; else lo = mid + 1;
                    }
                    vb = lo - 1;
                    A = mesh.larea[f];
                }
                if (__ballot(va <= vb) == 0ull) continue;  // no face straddles a sample radius
                const int vmin = wave_min_i(va <= vb ? va : nv);
                const int vmax = wave_max_i(va <= vb ? vb : -1);
                for (int v = vmin; v <= vmax; ++v) {
                    double a = 0.0, da;
                    if (va <= v && v <= vb && rk[v] < nbins) cdf_sorted(fa, fb, fc, A, vr[v], a, da);
                    const double sum = wave_sum(a);
                    if (lane == 0 && sum != 0.0) atomicAdd(&racc[v], sum);
                }
            }
        }
    }
    __syncthreads();

    // ---- phase 5: outputs and cleaned (radius, area) samples -------------------------------
    for (int s = tid; s < S; s += NT) {
        radii[s] = r_out[s];
        perim[s] = (r_out[s] == r_out[s]) ? (act[s] == 1 ? p_out[s] : 0.0) : NAN;
    }
    for (int e = tid; e < NE; e += NT) {
        er[e] = rho[e];
        ea[e] = NAN;
    }
    __syncthreads();
    for (int v = tid; v < nv; v += NT) ea[ridx[v]] = racc[v];
    __syncthreads();
    if (tid == 0) {
        int ne = 0;
        for (int s = 0; s < S; ++s) ne += ev_out[s];
        ne += nv;
        const int maxp = S + NE;
        float* sr = out.samp_r + (int64_t)j * maxp;
        float* sa = out.samp_a + (int64_t)j * maxp;
        double* pr = pair_r;
        double* pa = pair_a;
        int np = 0;
        for (int s = 0; s < S; ++s)
            if (isfinite(r_out[s])) {
                pr[np] = r_out[s];
                pa[np] = md.targets[s];
                ++np;
            }
        const double db32 = (double)(float)dist_b;  // the reference reads the float32-stored value back
        for (int v = 0; v < nv; ++v) {
            const double rr = vr[v], aa = racc[v];
            if (!(isfinite(rr) && isfinite(aa))) continue;
            if (md.cap >= 0.0) {
                double loss;
                if (rr <= 0.0) {
                    loss = NAN;
                } else if (!isfinite(db32) || db32 >= rr) {
                    loss = 0.0;
                } else {
                    const double x = fmin(1.0, fmax(0.0, db32 / rr));
                    loss = (acos(x) - x * sqrt(fmax(0.0, 1.0 - x * x))) / M_PI;
                }
                if (loss > md.cap) continue;
            }
            pr[np] = rr;
            pa[np] = aa;
            ++np;
        }
        for (int a = 1; a < np; ++a) {  // stable insertion sort by radius
            const double kr = pr[a], ka = pa[a];
            int b = a - 1;
            while (b >= 0 && pr[b] > kr) {
                pr[b + 1] = pr[b];
                pa[b + 1] = pa[b];
                --b;
            }
            pr[b + 1] = kr;
            pa[b + 1] = ka;
        }
        int nout = 0;
        if (np > 0) {
            const double tol = fmax(1e-12, 1e-9 * (pr[np - 1] - pr[0]));
            for (int a = 0; a < np; ++a) {
                if (a == 0 || (pr[a] - pr[a - 1]) > tol) {
                    sr[nout] = (float)pr[a];
                    sa[nout] = (float)pa[a];
                    ++nout;
                }
            }
        }
        double msd_u = NAN, msd_w = NAN;
        if (i4[2] > 0) {
            msd_u = v4[0] / (double)i4[2];
            msd_w = (isfinite(v4[2]) && v4[2] > 0.0) ? v4[1] / v4[2] : NAN;
        }
        int ovf = 0;
        for (int s = 0; s < S; ++s) ovf |= (act[s] == 1 && cnt[s] > cap);
        scal[kSlotFlag] = bad ? 1.0 : 0.0;
        scal[kSlotNeg] = neg_frac;
        scal[kSlotViol] = viol_frac;
        scal[kSlotDistB] = dist_b;
        scal[kSlotMsdU] = msd_u;
        scal[kSlotMsdW] = msd_w;
        scal[kSlotNEval] = (double)ne;
        scal[kSlotNSamples] = (double)nout;
        scal[kSlotOverflow] = (double)ovf;
    }
}

}  // namespace

void launch_metrics(hipStream_t st, int k, const MeshDev& mesh, const MetricsDev& md, const double* Dt,
                    const int* src_pos, MetricsOut out, bool health_screen) {
    if (k <= 0) return;
    metrics_kernel<<<dim3(k), NT, 0, st>>>(mesh, md, Dt, src_pos, out, health_screen ? 1 : 0);
}

}  // namespace fcw
