// Per-source wiring metrics on gfx950: one 512-thread workgroup per distance
// field (core_analysis.py:1558-1609, lean_geometry.py:117-357).
//
//   phase 0  vertex pass: health screen, MSD (unweighted / area-weighted),
//            distance to boundary, largest finite distance
//   phase 1  face pass: per-bin face area and face count of dmax and of dmin.
//            For r on the bin grid, sum(dmax < r) <= A(r) <= sum(dmin < r), which
//            brackets every target radius to a few bins before any exact
//            evaluation; the counts bound how many faces a bracket can hold.
//   phase 2  brackets: a band [kL wb, H] per target radius; each band gets an
//            exact-size slot range in the per-source scratch.
//   phase 3  face pass: every face whose [dmin, dmax] meets a band is copied,
//            pre-sorted (with its area and index), into that band. Phase 1
//            records the distance range of every 256-face chunk, so chunks that
//            meet no band (most of the surface: the bands lie within the largest
//            radius) are skipped without loading them. For r in [kL wb, H],
//            A(r) = (area of faces with dmax bin < kL) + sum over the band of the
//            face's exact area inside r.
//   phase 4  safeguarded Newton on the exact piecewise-quadratic A(r) (one wave
//            per scale), perimeter by the coarea formula at the root
//   phase 5  exact areas at the supplementary sample radii, now known: one face
//            pass over the chunks that meet any sample's shell (faces with dmax
//            bin >= the sample's bin and dmin below its radius), each face adding
//            its exact area inside each such radius (no band lists: every sample
//            is evaluated once)
//   phase 6  outputs and the cleaned (radius, area) samples
//
// Face passes are coalesced and issue all loads of a step before any is used;
// histogram updates are LDS atomics (lanes cover consecutive faces, so they
// rarely collide). Sums whose order depends on scheduling (histogram bins,
// band contents) only steer the work or change results in the last bits.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdint>

#include "fcw_device.h"

namespace fcw {
namespace {

constexpr int kWave = 64;
constexpr int NT = kMetricThreads;
constexpr int NW = NT / kWave;
constexpr int kMaxBands = kMaxScales + kMaxExtras;

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

__device__ inline uint64_t lanes_below(int lane) { return (1ull << lane) - 1ull; }

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

// Faces are visited UNR per lane per step: all index loads, then all gathers (and the face
// area), so a wave keeps several independent gathers in flight.
constexpr int UNR = 4;

struct FaceBatch {
    int f[UNR];
    double d0[UNR], d1[UNR], d2[UNR], A[UNR];
};

// A wave-step covers the 256 consecutive faces from base. CONSEC: lane takes base + UNR*lane + u
// (runs of consecutive faces per lane: neighbouring faces share histogram bins, few collisions
// across lanes); otherwise base + 64 u + lane (each step u spans a 64-face patch: few bands).
template <bool CONSEC>
__device__ inline void load_faces(const int* __restrict__ faces, const double* __restrict__ larea,
                                  const double* __restrict__ d, double sh, int m, int base, int lane, FaceBatch& fb) {
    int i0[UNR], i1[UNR], i2[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
        const int f = CONSEC ? base + UNR * lane + u : base + kWave * u + lane;
        fb.f[u] = f;
        const int fs = f < m ? f : m - 1;  // an in-bounds face; the caller ignores f >= m
        i0[u] = faces[3 * fs];
        i1[u] = faces[3 * fs + 1];
        i2[u] = faces[3 * fs + 2];
        fb.A[u] = larea[fs];
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
        fb.d0[u] = d[i0[u]] - sh;
        fb.d1[u] = d[i1[u]] - sh;
        fb.d2[u] = d[i2[u]] - sh;
    }
    __builtin_amdgcn_sched_barrier(0);  // every load of the step is issued before any use
}

// The chunks c < nchunks with keep(c) written in increasing order to list (block-wide, every thread
// takes part); returns their number.
template <typename Keep>
__device__ int chunk_list(int nchunks, Keep keep, int* list, int* wcnt, int& nlist) {
    const int tid = threadIdx.x, lane = lane_id(), wid = __builtin_amdgcn_readfirstlane(tid / kWave);
    if (tid == 0) nlist = 0;
    __syncthreads();
    for (int c0 = 0; c0 < nchunks; c0 += NT) {
        const int c = c0 + tid;
        const bool k = c < nchunks && keep(c);
        const uint64_t mk = __ballot(k);
        if (lane == 0) wcnt[wid] = __popcll(mk);
        __syncthreads();
        int before = nlist;
        for (int w = 0; w < wid; ++w) before += wcnt[w];
        if (k) list[before + __popcll(mk & lanes_below(lane))] = c;
        __syncthreads();
        if (tid == 0)
            for (int w = 0; w < NW; ++w) nlist += wcnt[w];
        __syncthreads();
    }
    return nlist;
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

// Log-spaced supplementary radius between r0 < r1 (lean_geometry.py:320-330).
__device__ inline double log_mix(double r0, double r1, double fr) { return exp(log(r0) + fr * (log(r1) - log(r0))); }

// One band's sums at radius r over lanes of one wave: area inside r beyond the fully-inside
// part and its derivative, and (with want_p) the coarea perimeter density. `pool` holds the
// band's sorted (a, b, c) in three arrays of stride `pstride` and, after them, the face indices
// (ints); each face's area comes from the mesh (an array every source shares); |grad d| is
// recomputed from the face for the faces that cross r. With `full`, the band is instead
// recomputed from every face (the band did not fit in the scratch).
struct BandRef {
    int kL;
    double H;
    int off, count;
    bool full;
};

// U band entries per lane in flight per step; each lane adds its entries in increasing index order
// whatever U is, so the sums do not depend on it.
template <int U>
__device__ void band_sums_u(const BandRef& B, double r, bool want_p, const double* __restrict__ pool, int64_t pstride,
                            const int* __restrict__ faces, const double* __restrict__ d, double sh,
                            const double* __restrict__ larea, const double* __restrict__ g11,
                            const double* __restrict__ g12, const double* __restrict__ g22, int m, double wb,
                            int nbins, double& sA, double& sdA, double& sP) {
    const int lane = lane_id();
    sA = sdA = sP = 0.0;
    if (!B.full) {
        const double* pa = pool + B.off;
        const double* pb = pa + pstride;
        const double* pc = pb + pstride;
        const int* pf = reinterpret_cast<const int*>(pool + 3 * pstride) + B.off;
        for (int i0 = lane; i0 < B.count; i0 += U * kWave) {
            double a[U], b[U], c[U], A[U];
            int f[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = min(i0 + u * kWave, B.count - 1);
                a[u] = pa[i];
                b[u] = pb[i];
                c[u] = pc[i];
                f[u] = pf[i];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) A[u] = larea[f[u]];
            __builtin_amdgcn_sched_barrier(0);
            double dx[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                double x;
                cdf_sorted(a[u], b[u], c[u], A[u], r, x, dx[u]);
                const bool ok = i0 + u * kWave < B.count;
                sA += ok ? x : 0.0;
                sdA += ok ? dx[u] : 0.0;
                dx[u] = ok ? dx[u] : 0.0;
            }
            if (want_p) {
                int i1[U], i2[U], i3[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    i1[u] = faces[3 * f[u]];
                    i2[u] = faces[3 * f[u] + 1];
                    i3[u] = faces[3 * f[u] + 2];
                }
                double d0[U], d1[U], d2[U], e11[U], e12[U], e22[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    d0[u] = d[i1[u]] - sh;
                    d1[u] = d[i2[u]] - sh;
                    d2[u] = d[i3[u]] - sh;
                    e11[u] = g11[f[u]];
                    e12[u] = g12[f[u]];
                    e22[u] = g22[f[u]];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int u = 0; u < U; ++u)
                    sP += (dx[u] > 0.0) ? dx[u] * grad_norm(d0[u], d1[u], d2[u], e11[u], e12[u], e22[u]) : 0.0;
            }
        }
    } else {
        for (int f = lane; f < m; f += kWave) {
            const double d0 = d[faces[3 * f]] - sh, d1 = d[faces[3 * f + 1]] - sh, d2 = d[faces[3 * f + 2]] - sh;
            if (!(isfinite(d0) && isfinite(d1) && isfinite(d2))) continue;
            double fa = d0, fb = d1, fc = d2;
            sort3(fa, fb, fc);
            if (!(bin_of(fc, wb, nbins) >= B.kL && fa < B.H)) continue;
            double x, dx;
            cdf_sorted(fa, fb, fc, larea[f], r, x, dx);
            sA += x;
            sdA += dx;
            if (want_p && dx > 0.0) sP += dx * grad_norm(d0, d1, d2, g11[f], g12[f], g22[f]);
        }
    }
}

// The perimeter evaluations (once or twice per scale) hold three more values per face in flight,
// so they take two entries per lane per step instead of four: with four the kernel needs 137
// registers per lane, one workgroup per CU; with two, 124, and two sources share a CU.
__device__ void band_sums(const BandRef& B, double r, bool want_p, const double* __restrict__ pool, int64_t pstride,
                          const int* __restrict__ faces, const double* __restrict__ d, double sh,
                          const double* __restrict__ larea, const double* __restrict__ g11,
                          const double* __restrict__ g12, const double* __restrict__ g22, int m, double wb,
                          int nbins, double& sA, double& sdA, double& sP) {
    if (want_p)
        band_sums_u<2>(B, r, true, pool, pstride, faces, d, sh, larea, g11, g12, g22, m, wb, nbins, sA, sdA, sP);
    else
        band_sums_u<4>(B, r, false, pool, pstride, faces, d, sh, larea, g11, g12, g22, m, wb, nbins, sA, sdA, sP);
}

// Four waves per SIMD, so that two workgroups (two sources) share a CU (the registers allow it
// without spilling: see band_sums); a 512-source launch takes 0.39 ms instead of 0.54 with one.
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void metrics_kernel(
    MeshDev mesh, MetricsDev md, const double* __restrict__ Dt, const int* __restrict__ src_pos, MetricsOut out,
    int health_mode) {
    const int j = blockIdx.x;
    const int tid = threadIdx.x, lane = lane_id(), wid = __builtin_amdgcn_readfirstlane(tid / kWave);
    const int n = mesh.n, m = mesh.m;
    const int* __restrict__ faces = mesh.faces;
    const double* __restrict__ d = Dt + (int64_t)j * n;
    // the field's value at its own source, subtracted from every value read (0 for fields stored
    // shifted already)
    const double sh = md.shift_src ? md.shift_src[(int64_t)src_pos[j] * md.shift_ld + j] : 0.0;
    const int S = md.S, q = md.q, NE = (S > 0 ? (S - 1) * q : 0);
    const int NB = S;  // one band per scale (the supplementary samples are summed by a face pass in phase 5)
    double* scal = out.scal + (int64_t)j * kSlotCount;
    double* radii = out.radii + (int64_t)j * S;
    double* perim = out.perim + (int64_t)j * S;
    double* er = out.extra_r + (int64_t)j * NE;
    double* ea = out.extra_a + (int64_t)j * NE;
    const int64_t pstride = md.pool_cap;
    double* pool = md.band + (int64_t)j * band_pool_doubles(pstride);

    __shared__ Red red;
    __shared__ double Hmax[kMaxBins + 1];
    __shared__ double Hmin[kMaxBins + 1];
    // Face counts per bin (phases 1-2); afterwards the same storage holds the sample arrays.
    __shared__ union {
        struct {
            unsigned cmax[kMaxBins + 1];
            unsigned cmin[kMaxBins + 1];
        } c;
        struct {
            int lk[kMaxBins + 1];  // per dmax bin: last band whose kL <= bin
            int fa[kMaxBins + 1];  // per dmin bin: first band whose H > bin start
        } t;
        struct {
            double rho[kMaxExtras], acc[kMaxExtras];
            double pair_r[kMaxBands], pair_a[kMaxBands];
            int key[kMaxBands], pos[kMaxBands];
        } x;
    } U;
    __shared__ int b_kL[kMaxBands], b_cnt[kMaxBands], b_off[kMaxBands], b_cap[kMaxBands];
    __shared__ double b_H[kMaxBands];
    __shared__ int lo_k[kMaxScales], act[kMaxScales], ev_out[kMaxScales];
    __shared__ double hi_r[kMaxScales], r_out[kMaxScales], p_out[kMaxScales];
    __shared__ float ch_lo[kMaxChunks], ch_hi[kMaxChunks];  // per 256-face chunk: min dmin (rounded
                                                             // down), max dmax (rounded up)
    __shared__ int s_nbins, s_ovf, s_nsamp;
    __shared__ int s_list[kMaxChunks], s_wcnt[NW], s_nlist;  // face chunks a pass visits, in order
    __shared__ double s_skip_lo, s_skip_hi, s_cut;
    __shared__ double s_wb;
    __shared__ double s_scal[kSlotCount];

    // ---- phase 0: health screen, MSD, distance to boundary -------------------------------
    const int sp = src_pos[j];
    const double xs = mesh.V[3 * sp], ys = mesh.V[3 * sp + 1], zs = mesh.V[3 * sp + 2];
    double v4[4] = {0.0, 0.0, 0.0, 0.0};  // sum d, sum d*w, sum w, max finite d
    int i4[4] = {0, 0, 0, 0};             // d < -h, chord violations, valid, non-finite
    double b4[4] = {INFINITY, 0.0, 0.0, 0.0};
    int bi4[4] = {0, 0, 0, 0};
    const double h = md.health_h;
    if (md.part) {
        // the vertex pass was made for the whole block by chunk_stats: add up its chunk records
        for (int c = tid; c < md.nch; c += NT) {
            const double* pr = md.part + chunk_rec_index(j, c, 0, md.nch);  // record values 8 apart
            v4[0] += pr[0];
            v4[1] += pr[8];
            v4[2] += pr[16];
            v4[3] = fmax(v4[3], pr[24]);
            b4[0] = fmin(b4[0], pr[32]);
            i4[0] += (int)pr[40];
            i4[1] += (int)pr[48];
            i4[2] += (int)pr[56];
            i4[3] += (int)pr[64];
            bi4[0] += (int)pr[72];
        }
    }
    for (int p0 = md.part ? n : tid; p0 < n; p0 += NT * UNR) {
        double dv[UNR], vx[UNR], vy[UNR], vz[UNR], vw[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const int p = p0 + u * NT;
            const int ps = p < n ? p : 0;
            dv[u] = d[ps] - sh;
            vx[u] = mesh.V[3 * ps];
            vy[u] = mesh.V[3 * ps + 1];
            vz[u] = mesh.V[3 * ps + 2];
            vw[u] = md.vw[ps];
        }
        __builtin_amdgcn_sched_barrier(0);
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
    for (int i = md.part ? md.nb : tid; i < md.nb; i += NT) {
        const double dv = d[md.boundary[i]] - sh;
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
    if (tid == 0) {  // parked in LDS until the end (frees registers for the face passes)
        s_scal[kSlotFlag] = bad ? 1.0 : 0.0;
        s_scal[kSlotNeg] = neg_frac;
        s_scal[kSlotViol] = viol_frac;
        s_scal[kSlotDistB] = dist_b;
        double msd_u = NAN, msd_w = NAN;
        if (i4[2] > 0) {
            msd_u = v4[0] / (double)i4[2];
            msd_w = (isfinite(v4[2]) && v4[2] > 0.0) ? v4[1] / v4[2] : NAN;
        }
        s_scal[kSlotMsdU] = msd_u;
        s_scal[kSlotMsdW] = msd_w;
    }
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

    // ---- phase 1: dmax / dmin area and count histograms ------------------------------------
    if (tid == 0) {
        // Bins only steer the work, never the result: an eighth of the reference bucket width
        // (narrow brackets, small bands), widened when the field is too long for the histogram.
        double wb = 0.125 * md.width;
        if (dglob / wb + 2.0 > (double)kMaxBins) wb = dglob / (double)(kMaxBins - 3);
        int nbins = (int)(dglob / wb) + 2;
        if (nbins > kMaxBins) nbins = kMaxBins;
        s_wb = wb;
        s_nbins = nbins;
        s_ovf = 0;
    }
    for (int i = tid; i <= kMaxBins; i += NT) {
        Hmax[i] = 0.0;
        Hmin[i] = 0.0;
        U.c.cmax[i] = 0u;
        U.c.cmin[i] = 0u;
    }
    __syncthreads();
    const double wb = s_wb;
    const int nbins = s_nbins;
    // Chunk distance ranges: from chunk_stats when given (global), else recorded below (LDS).
    const bool have_ch = md.ch_lo != nullptr;
    auto gch_lo = [&](int c) { return md.ch_lo[chunk_range_index(j, c, md.nch)]; };
    auto gch_hi = [&](int c) { return md.ch_hi[chunk_range_index(j, c, md.nch)]; };
    const bool chunked = have_ch || m <= kMaxChunks * kWave * UNR;
    // The face passes test every chunk's range before loading its faces, one chunk per wave step:
    // from LDS that test costs a few cycles, from global memory a full load latency per step.
    const bool ch_in_lds = !have_ch || md.nch <= kMaxChunks;
    if (have_ch && ch_in_lds)
        for (int c = tid; c < md.nch; c += NT) {
            ch_lo[c] = gch_lo(c);
            ch_hi[c] = gch_hi(c);
        }
    // With chunk ranges and a finite field, only faces that start below `cut` are histogrammed:
    // the chunks entirely below bin K already hold the largest target's area, so every radius,
    // bracket, band and band capacity lies below bin K + 3 (eight bins of slack are kept).
    if (tid == 0) s_cut = INFINITY;
    if (have_ch && i4[3] == 0) {
        double need = 0.0;
        for (int s = S - 1; s >= 0; --s)
            if (md.targets[s] == md.targets[s]) {
                need = md.targets[s] * (1.0 + 1e-9);
                break;
            }
        for (int c = tid; c < md.nch; c += NT) {
            const double hi = (double)gch_hi(c);
            if (hi > -INFINITY) atomicAdd(&Hmin[bin_of(hi, wb, nbins)], md.chunk_area[c]);
        }
        __syncthreads();
        if (wid == 0) {
            const int per = (nbins + kWave - 1) / kWave;
            const int b0 = lane * per, b1 = min(nbins, b0 + per);
            double loc = 0.0;
            for (int b = b0; b < b1; ++b) loc += Hmin[b];
            double inc = loc;
            for (int o = 1; o < kWave; o <<= 1) {
                const double y = __shfl_up(inc, o, kWave);
                if (lane >= o) inc += y;
            }
            double run = inc - loc;  // sum over bins < b0
            int kfound = INT_MAX;     // first K with sum over bins < K >= need
            if (run >= need) kfound = b0;
            for (int b = b0; b < b1 && kfound == INT_MAX; ++b) {
                run += Hmin[b];
                if (run >= need) kfound = b + 1;
            }
            kfound = wave_min_i(kfound);
            if (lane == 0 && kfound != INT_MAX) s_cut = (double)(kfound + 8) * wb;
        }
        __syncthreads();
        for (int i = tid; i <= kMaxBins; i += NT) Hmin[i] = 0.0;
        __syncthreads();
    }
    __syncthreads();
    const double cut = s_cut;
    // one wave step of the histogram pass (CONSEC face order); returns the step's distance range
    auto hist_step = [&](const FaceBatch& fb, double& cmin, double& cmax) {
        cmin = INFINITY;
        cmax = -INFINITY;
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const double d0 = fb.d0[u], d1 = fb.d1[u], d2 = fb.d2[u];
            if (fb.f[u] < m && isfinite(d0) && isfinite(d1) && isfinite(d2)) {
                const double hi = fmax(d0, fmax(d1, d2)), lo = fmin(d0, fmin(d1, d2));
                const int kh = bin_of(hi, wb, nbins);
                const int kl = bin_of(lo, wb, nbins);
                atomicAdd(&Hmax[kh], fb.A[u]);
                atomicAdd(&Hmin[kl], fb.A[u]);
                atomicAdd(&U.c.cmax[kh], 1u);
                atomicAdd(&U.c.cmin[kl], 1u);
                cmin = fmin(cmin, lo);
                cmax = fmax(cmax, hi);
            }
        }
    };
    const int nchunks1 = (m + kWave * UNR - 1) / (kWave * UNR);
    if (have_ch && ch_in_lds) {
        // the chunks that start below the cut, listed in order; each wave loads its next chunk's
        // faces while it adds up the current one
        auto below_cut = [&](int c) { return (double)ch_lo[c] < cut; };
        const int nlist = chunk_list(nchunks1, below_cut, s_list, s_wcnt, s_nlist);
        FaceBatch cur, nxt;
        int k = wid;
        if (k < nlist) load_faces<true>(faces, mesh.larea, d, sh, m, s_list[k] * (kWave * UNR), lane, cur);
        nxt = cur;
        for (; k < nlist; k += NW) {
            if (k + NW < nlist) load_faces<true>(faces, mesh.larea, d, sh, m, s_list[k + NW] * (kWave * UNR), lane, nxt);
            double cmin, cmax;
            hist_step(cur, cmin, cmax);
            cur = nxt;
        }
    } else {
        for (int base = wid * kWave * UNR; base < m; base += NT * UNR) {
            if (have_ch && (double)gch_lo(base / (kWave * UNR)) >= cut) continue;  // beyond every band
            FaceBatch fb;
            load_faces<true>(faces, mesh.larea, d, sh, m, base, lane, fb);
            double cmin, cmax;
            hist_step(fb, cmin, cmax);
            if (!have_ch && chunked) {
                // float bounds rounded outwards, reduced over the wave
                float flo = __double2float_rd(cmin), fhi = __double2float_ru(cmax);
                for (int o = kWave / 2; o > 0; o >>= 1) {
                    flo = fminf(flo, __shfl_xor(flo, o, kWave));
                    fhi = fmaxf(fhi, __shfl_xor(fhi, o, kWave));
                }
                if (lane == 0) {
                    ch_lo[base / (kWave * UNR)] = flo;
                    ch_hi[base / (kWave * UNR)] = fhi;
                }
            }
        }
    }
    __syncthreads();
    // exclusive prefix sums in place: X[k] = sum over bins < k, k = 0..nbins
    if (wid < 4) {
        const int per = (nbins + kWave - 1) / kWave;
        const int b0 = lane * per, b1 = min(nbins, b0 + per);
        if (wid < 2) {
            double* H = (wid == 0) ? Hmax : Hmin;
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
        } else {
            unsigned* C = (wid == 2) ? U.c.cmax : U.c.cmin;
            unsigned loc = 0;
            for (int b = b0; b < b1; ++b) loc += C[b];
            unsigned inc = loc;
            for (int o = 1; o < kWave; o <<= 1) {
                const unsigned y = __shfl_up(inc, o, kWave);
                if (lane >= o) inc += y;
            }
            unsigned run = inc - loc;
            for (int b = b0; b < b1; ++b) {
                const unsigned cv = C[b];
                C[b] = run;
                run += cv;
            }
            if (lane == kWave - 1) C[nbins] = inc;
        }
    }
    __syncthreads();
    const double total = Hmax[nbins];

    // ---- phase 2: brackets and bands ------------------------------------------------------
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
            hi_r[s] = (double)khi * wb;
        } else {
            lo_k[s] = 0;
            hi_r[s] = 0.0;
        }
    }
    __syncthreads();
    // Band brackets. An empty band has kL = INT_MAX (no dmax bin reaches it) and H = -inf
    // before the first active scale or +inf after the last.
    for (int i = tid; i < NB; i += NT) {
        const int s = i;
        int kL = INT_MAX;
        double H = act[s] == 2 ? INFINITY : -INFINITY;
        if (act[s] == 1) {
            kL = lo_k[s];
            H = hi_r[s];
        }
        b_kL[i] = kL;
        b_H[i] = H;
    }
    __syncthreads();
    if (tid == 0) {
        // Monotone in band order (a band may only widen): then each face's bands form one
        // contiguous range, found by two binary searches.
        double hm = -INFINITY;
        for (int i = 0; i < NB; ++i) {
            hm = fmax(hm, b_H[i]);
            b_H[i] = hm;
        }
        int km = INT_MAX;
        for (int i = NB - 1; i >= 0; --i) {
            km = min(km, b_kL[i]);
            b_kL[i] = km;
        }
        // capacity: faces with dmin < H minus faces with dmax bin < kL, exact-size offsets
        int64_t off = 0;
        int ovf = 0;
        for (int i = 0; i < NB; ++i) {
            int cap = 0;
            if (b_kL[i] < nbins && b_H[i] > -INFINITY) {
                const unsigned below_h = (b_H[i] == INFINITY) ? U.c.cmin[nbins]
                                                              : U.c.cmin[bin_of(b_H[i], wb, nbins) + 1];
                const unsigned inside = U.c.cmax[b_kL[i]];
                cap = below_h > inside ? (int)(below_h - inside) : 0;
            }
            if (off + cap > md.pool_cap) {
                b_off[i] = -1;  // recomputed from every face when evaluated
                b_cap[i] = 0;
                ovf = 1;
            } else {
                b_off[i] = (int)off;
                b_cap[i] = cap;
                off += cap;
            }
            b_cnt[i] = 0;
        }
        s_ovf = ovf;
        // a chunk can hold a band face only if some face has dmax bin >= min kL and dmin < max H
        // over the bands that can hold faces (one bin of slack on the kL side)
        double lo = INFINITY, hi = -INFINITY;
        for (int i = 0; i < NB; ++i) {
            if (b_kL[i] >= nbins || !(b_H[i] > -INFINITY)) continue;
            lo = fmin(lo, (double)(b_kL[i] - 1) * wb);
            hi = fmax(hi, b_H[i]);
        }
        s_skip_lo = lo;
        s_skip_hi = hi;
    }
    __syncthreads();
    // Band lookup by bin (the count histograms are no longer needed): a face's bands are
    // [first band with H > dmin, last band with kL <= dmax bin].
    for (int k = tid; k < nbins; k += NT) {
        int lo = 0, hi = NB;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (b_kL[mid] > k) hi = mid; else lo = mid + 1;
        }
        U.t.lk[k] = lo - 1;
        const double x = (double)k * wb;
        lo = 0;
        hi = NB;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (b_H[mid] > x) hi = mid; else lo = mid + 1;
        }
        U.t.fa[k] = lo;
    }
    __syncthreads();

    // ---- phase 3: band faces, pre-sorted, contiguous per band -----------------------------
    const double skip_lo = s_skip_lo, skip_hi = s_skip_hi;
    auto band_step = [&](const FaceBatch& fb) {
        int ia[UNR], ib[UNR];
        double sa_[UNR], sb_[UNR], sc_[UNR];
        bool anyf = false;
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            ia[u] = NB;
            ib[u] = -1;
            double fa = fb.d0[u], fbv = fb.d1[u], fc = fb.d2[u];
            if (fb.f[u] < m && isfinite(fa) && isfinite(fbv) && isfinite(fc)) {
                sort3(fa, fbv, fc);
                ib[u] = U.t.lk[bin_of(fc, wb, nbins)];
                int i = U.t.fa[bin_of(fa, wb, nbins)];  // H > bin start; step past H <= dmin
                while (i < NB && b_H[i] <= fa) ++i;
                ia[u] = i;
            }
            sa_[u] = fa;
            sb_[u] = fbv;
            sc_[u] = fc;
            anyf |= ia[u] <= ib[u];
        }
        if (__ballot(anyf) == 0ull) return;  // no band face in this wave's step
        // Each lane takes a slot in each of its faces' bands with its own LDS atomic: the lanes'
        // (few) bands are walked together, so the wave waits for one atomic per band a face holds
        // rather than for one per band the whole step spans.
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            for (int i = ia[u]; i <= ib[u]; ++i) {
                if (b_off[i] < 0) continue;  // a band recomputed from every face
                const int slot = atomicAdd(&b_cnt[i], 1);
                if (slot < b_cap[i]) {
                    double* bs = pool + b_off[i] + slot;
                    bs[0] = sa_[u];
                    bs[pstride] = sb_[u];
                    bs[2 * pstride] = sc_[u];
                    reinterpret_cast<int*>(pool + 3 * pstride)[b_off[i] + slot] = fb.f[u];
                }
            }
        }
    };
    auto meets_bands = [&](int c) {
        const double clo = ch_in_lds ? (double)ch_lo[c] : (double)gch_lo(c);
        const double chi = ch_in_lds ? (double)ch_hi[c] : (double)gch_hi(c);
        return !(chi < skip_lo || clo >= skip_hi);
    };
    const int nchunks = (m + kWave * UNR - 1) / (kWave * UNR);
    if (chunked && nchunks <= kMaxChunks) {
        // The chunks that meet the bands, listed in order; the waves take them in turn, each
        // loading its next chunk's faces while it files the current one.
        const int nlist = chunk_list(nchunks, meets_bands, s_list, s_wcnt, s_nlist);
        FaceBatch cur, nxt;
        int k = wid;
        if (k < nlist) load_faces<false>(faces, mesh.larea, d, sh, m, s_list[k] * (kWave * UNR), lane, cur);
        nxt = cur;
        for (; k < nlist; k += NW) {
            if (k + NW < nlist) load_faces<false>(faces, mesh.larea, d, sh, m, s_list[k + NW] * (kWave * UNR), lane, nxt);
            band_step(cur);
            cur = nxt;
        }
    } else {
        for (int base = wid * kWave * UNR; base < m; base += NT * UNR) {
            if (chunked && !meets_bands(base / (kWave * UNR))) continue;
            FaceBatch fb;
            load_faces<false>(faces, mesh.larea, d, sh, m, base, lane, fb);
            band_step(fb);
        }
    }
    __syncthreads();
    if (tid < NB && b_off[tid] >= 0 && b_cnt[tid] > b_cap[tid]) atomicOr(&s_ovf, 2);  // cannot happen: caps are bounds
    __syncthreads();

    auto band_ref = [&](int i) {
        BandRef B;
        B.kL = b_kL[i];
        B.H = b_H[i];
        B.full = b_off[i] < 0 || b_cnt[i] > b_cap[i];
        B.off = B.full ? 0 : b_off[i];
        B.count = B.full ? 0 : b_cnt[i];
        return B;
    };

    // ---- phase 4: safeguarded Newton per scale (one wave per scale) ------------------------
    auto newton = [&](int s) {
        if (act[s] != 1) return;
        const double T = md.targets[s];
        const BandRef B = band_ref(s);
        const double Ain = Hmax[min(B.kL, nbins)];
        const int klo = lo_k[s];
        const double br_lo = (double)klo * wb, br_hi = hi_r[s];
        const int khi = (int)(br_hi / wb + 0.5);
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
        auto eval = [&](double rr, bool want_p, double& A_out, double& dA_out, double& P_out) {
            double sA, sdA, sP;
            band_sums(B, rr, want_p, pool, pstride, faces, d, sh, mesh.larea, mesh.g11, mesh.g12, mesh.g22, m, wb,
                      nbins, sA, sdA, sP);
            A_out = Ain + wave_sum(sA);
            dA_out = wave_sum(sdA);
            P_out = wave_sum(sP);
        };
        int n_eval = 0;
        double result = r, Pres = 0.0;
        bool converged = false, have_p = false;
        // When a Newton step starts within 1e-5 T of the target, the next iterate almost always
        // meets the tolerance (quadratic convergence), so that evaluation also sums the perimeter
        // and the separate perimeter pass at the root is skipped.
        bool want_p = false;
        for (int it = 0; it < md.max_iter; ++it) {
            double A, dA, Pv;
            eval(r, want_p, A, dA, Pv);
            ++n_eval;
            const double err = A - T;
            if (fabs(err) <= md.rtol * T) {
                result = r;
                converged = true;
                have_p = want_p;
                Pres = Pv;
                break;
            }
            if (err > 0.0) r_hi = r; else r_lo = r;
            double rn = 0.5 * (r_lo + r_hi);
            bool newton_step = false;
            if (dA > 0.0) {
                const double cand = r - err / dA;
                if (r_lo < cand && cand < r_hi) {
                    rn = cand;
                    newton_step = true;
                }
            }
            if (r_hi - r_lo <= 1e-13 * (1.0 + r_hi)) {
                result = rn;
                break;
            }
            want_p = newton_step && fabs(err) <= 1e-5 * T;
            r = rn;
            result = r;
        }
        if (!have_p) {
            // perimeter at the root (counted as an evaluation when the root was not the last one)
            double A, dA;
            eval(result, true, A, dA, Pres);
            if (!converged) ++n_eval;
        }
        if (lane == 0) {
            r_out[s] = result;
            p_out[s] = Pres;
            ev_out[s] = n_eval;
        }
    };
    // Scales by wave. The targets increase with the scale index and so do the bands, the cost of
    // each evaluation: with NW < S <= 2 NW the 2 NW - S largest scales run alone and the others in
    // pairs, smallest with largest (12 scales on 8 waves: about the largest scale's time instead of
    // the 4th and 12th together); otherwise round robin. Each scale's result does not depend on the
    // wave that computes it.
    if (S > NW && S <= 2 * NW) {
        const int alone = 2 * NW - S;
        if (wid < alone) {
            newton(S - 1 - wid);
        } else {
            const int p = wid - alone, rest = S - alone;
            newton(p);
            newton(rest - 1 - p);
        }
    } else {
        for (int s = wid; s < S; s += NW) newton(s);
    }
    __syncthreads();

    // ---- phase 5: supplementary samples between consecutive radii --------------------------
    // A(r) = (area of faces with dmax bin < bin(r)) + the exact area inside r of every face with
    // dmax bin >= bin(r) and dmin < r. The sample radii increase with their index, so the valid
    // ones are listed in order (U.x.key: sample index, U.x.pos: its bin, U.x.pair_r: its radius)
    // and a face's samples form one run of that list.
    if (tid == 0) {
        int nv = 0;
        for (int e = 0; e < NE; ++e) {
            const int s = e / q, jj = e - s * q;
            const double r0 = r_out[s], r1 = r_out[s + 1];
            double rr = NAN;
            if (r0 == r0 && r1 == r1 && r0 > 0.0 && r1 > r0) rr = log_mix(r0, r1, md.fracs[jj]);
            U.x.rho[e] = rr;
            U.x.acc[e] = rr == rr ? 0.0 : NAN;
            if (rr == rr) {
                U.x.key[nv] = e;
                U.x.pos[nv] = min(bin_of(rr, wb, nbins), (int)(rr / wb));
                U.x.pair_r[nv] = rr;
                ++nv;
            }
        }
        s_nsamp = nv;
    }
    __syncthreads();
    const int nsv = s_nsamp;
    if (nsv > 0) {
        // first listed sample whose radius exceeds x
        auto first_above = [&](double x) {
            int lo = 0, hi = nsv;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (U.x.pair_r[mid] > x) hi = mid; else lo = mid + 1;
            }
            return lo;
        };
        auto sample_step = [&](const FaceBatch& fb) {
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                double fa = fb.d0[u], fbv = fb.d1[u], fc = fb.d2[u];
                if (!(fb.f[u] < m && isfinite(fa) && isfinite(fbv) && isfinite(fc))) continue;
                sort3(fa, fbv, fc);
                const int kmax = bin_of(fc, wb, nbins);
                for (int i = first_above(fa); i < nsv && U.x.pos[i] <= kmax; ++i) {
                    double x, dx;
                    cdf_sorted(fa, fbv, fc, fb.A[u], U.x.pair_r[i], x, dx);
                    atomicAdd(&U.x.acc[U.x.key[i]], x);
                }
            }
        };
        // a chunk can hold a sample's face only if its range reaches that sample's bin (one bin
        // of slack) and starts below its radius
        auto meets_samples = [&](int c) {
            const double clo = ch_in_lds ? (double)ch_lo[c] : (double)gch_lo(c);
            const double chi = ch_in_lds ? (double)ch_hi[c] : (double)gch_hi(c);
            const int i = first_above(clo);
            return i < nsv && (double)(U.x.pos[i] - 1) * wb <= chi;
        };
        const int nchunks5 = (m + kWave * UNR - 1) / (kWave * UNR);
        if (chunked && nchunks5 <= kMaxChunks) {
            const int nlist = chunk_list(nchunks5, meets_samples, s_list, s_wcnt, s_nlist);
            FaceBatch cur, nxt;
            int k = wid;
            if (k < nlist) load_faces<false>(faces, mesh.larea, d, sh, m, s_list[k] * (kWave * UNR), lane, cur);
            nxt = cur;
            for (; k < nlist; k += NW) {
                if (k + NW < nlist) load_faces<false>(faces, mesh.larea, d, sh, m, s_list[k + NW] * (kWave * UNR), lane, nxt);
                sample_step(cur);
                cur = nxt;
            }
        } else {
            for (int base = wid * kWave * UNR; base < m; base += NT * UNR) {
                if (chunked && !meets_samples(base / (kWave * UNR))) continue;
                FaceBatch fb;
                load_faces<false>(faces, mesh.larea, d, sh, m, base, lane, fb);
                sample_step(fb);
            }
        }
        __syncthreads();
        for (int i = tid; i < nsv; i += NT) {
            const int e = U.x.key[i];
            U.x.acc[e] += Hmax[min(U.x.pos[i], nbins)];
        }
    }
    __syncthreads();

    // ---- phase 6: outputs and cleaned (radius, area) samples -------------------------------
    for (int s = tid; s < S; s += NT) {
        radii[s] = r_out[s];
        perim[s] = (r_out[s] == r_out[s]) ? (act[s] == 1 ? p_out[s] : 0.0) : NAN;
    }
    for (int e = tid; e < NE; e += NT) {
        er[e] = U.x.rho[e];
        ea[e] = U.x.acc[e];
    }
    // Cleaned samples (core_analysis.py:1599-1609, 1424-1439) by wave 0: candidates are the
    // solved (radius, target) pairs then the kept supplementary pairs, stably sorted by radius
    // (rank = smaller radii + equal radii earlier), then radii within tol of their predecessor
    // dropped.
    const int maxp = S + NE;
    if (wid == 0) {
        const double db32 = (double)(float)s_scal[kSlotDistB];  // the reference reads the float32-stored value back
        int np = 0;
        for (int c0 = 0; c0 < maxp; c0 += kWave) {
            const int c = c0 + lane;
            bool keep = false;
            double rr = 0.0, aa = 0.0;
            if (c < S) {
                rr = r_out[c];
                aa = md.targets[c];
                keep = isfinite(rr);
            } else if (c < maxp) {
                rr = U.x.rho[c - S];
                aa = U.x.acc[c - S];
                keep = isfinite(rr) && isfinite(aa);
                if (keep && md.cap >= 0.0) {
                    double loss;
                    if (rr <= 0.0) {
                        loss = NAN;
                    } else if (!isfinite(db32) || db32 >= rr) {
                        loss = 0.0;
                    } else {
                        const double x = fmin(1.0, fmax(0.0, db32 / rr));
                        loss = (acos(x) - x * sqrt(fmax(0.0, 1.0 - x * x))) / M_PI;
                    }
                    keep = !(loss > md.cap);
                }
            }
            const uint64_t km = __ballot(keep);
            if (keep) {
                const int at = np + __popcll(km & lanes_below(lane));
                U.x.pair_r[at] = rr;
                U.x.pair_a[at] = aa;
            }
            np += __popcll(km);
        }
        __builtin_amdgcn_wave_barrier();
        for (int a = lane; a < np; a += kWave) {
            const double ra = U.x.pair_r[a];
            int rank = 0;
            for (int b = 0; b < np; ++b) {
                const double rb = U.x.pair_r[b];
                rank += (rb < ra || (rb == ra && b < a)) ? 1 : 0;
            }
            U.x.pos[rank] = a;
        }
        __builtin_amdgcn_wave_barrier();
        float* sr = out.samp_r + (int64_t)j * maxp;
        float* sa = out.samp_a + (int64_t)j * maxp;
        const double tol = np > 0 ? fmax(1e-12, 1e-9 * (U.x.pair_r[U.x.pos[np - 1]] - U.x.pair_r[U.x.pos[0]])) : 0.0;
        int nout = 0;
        for (int c0 = 0; c0 < np; c0 += kWave) {
            const int c = c0 + lane;
            bool keep = false;
            double rr = 0.0, aa = 0.0;
            if (c < np) {
                rr = U.x.pair_r[U.x.pos[c]];
                aa = U.x.pair_a[U.x.pos[c]];
                keep = (c == 0) || (rr - U.x.pair_r[U.x.pos[c - 1]]) > tol;
            }
            const uint64_t km = __ballot(keep);
            if (keep) {
                const int at = nout + __popcll(km & lanes_below(lane));
                sr[at] = (float)rr;
                sa[at] = (float)aa;
            }
            nout += __popcll(km);
        }
        if (lane == 0) U.x.key[0] = nout;
    }
    __syncthreads();
    if (tid == 0) {
        int ne = 0, nv = 0;
        for (int s = 0; s < S; ++s) ne += ev_out[s];
        for (int e = 0; e < NE; ++e) nv += (U.x.rho[e] == U.x.rho[e]) ? 1 : 0;
        ne += nv;
        const int nout = U.x.key[0];
        for (int c = kSlotFlag; c <= kSlotMsdW; ++c) scal[c] = s_scal[c];
        scal[kSlotNEval] = (double)ne;
        scal[kSlotNSamples] = (double)nout;
        scal[kSlotOverflow] = (double)s_ovf;
    }
}

// Area of every 256-face chunk (the unit of the distance-range records), summed in face order.
__global__ void chunk_area_kernel(int m, const double* __restrict__ larea, double* __restrict__ chunk_area) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int f0 = c * kWave * UNR;
    if (f0 >= m) return;
    double a = 0.0;
    for (int f = f0; f < min(m, f0 + kWave * UNR); ++f) a += larea[f];
    chunk_area[c] = a;
}

// The far-field part of the per-source metrics for all fields of a block at once, vertex-major
// (core_analysis.py:1558-1581, lean_geometry.py:343-357): one workgroup per face chunk and 256
// sources, one lane per source, each wave walking all of the chunk's vertices (their indices
// staged in LDS) for its 64 sources, 8 vertices per step with every load of the step in flight
// together. For each source: the distance range over the chunk's vertices (bounds every face of
// the chunk), and over the vertices the chunk owns (each vertex is owned by exactly one chunk)
// the partial record of kChunkStats values. d = phi[p][j] - phi[src_j][j], as the shifted fields.
constexpr int kStatsVerts = 1024;  // chunk vertices staged in LDS per page (face vertices: at most 768)

__global__ __launch_bounds__(256) void chunk_stats_kernel(int ldk, int k, int nch, const double* __restrict__ Z,
                                                          const int* __restrict__ src_pos, const double* __restrict__ V,
                                                          const double* __restrict__ vw, const int* __restrict__ isb,
                                                          const int* __restrict__ cvptr, const int* __restrict__ cverts,
                                                          const int* __restrict__ cnown, double eps, double h,
                                                          float* __restrict__ ch_lo, float* __restrict__ ch_hi,
                                                          double* __restrict__ part) {
    constexpr int U = 8;
    __shared__ int verts[kStatsVerts];
    const int c = blockIdx.x;
    const int v0 = cvptr[c], ntot = cvptr[c + 1] - v0, nown_all = cnown[c];
    const int jl = blockIdx.y * 256 + threadIdx.x;
    const bool wave_on = blockIdx.y * 256 + (int)(threadIdx.x & ~(kWave - 1)) < k;  // wave holds a source
    const int jc = min(jl, k - 1);
    const int sp = src_pos[jc];
    const double base = Z[(int64_t)sp * ldk + jc];
    const double xs = V[3 * sp], ys = V[3 * sp + 1], zs = V[3 * sp + 2];
    double lo = INFINITY, hi = -INFINITY, sd = 0.0, sdw = 0.0, sw = 0.0, dmax = 0.0, bmin = INFINITY;
    int nneg = 0, nviol = 0, nval = 0, nnf = 0, nbf = 0;
    for (int pg = 0; pg < ntot; pg += kStatsVerts) {  // one page unless the chunk owns faceless vertices
        const int nv = min(ntot - pg, kStatsVerts), nown = nown_all - pg;
        __syncthreads();
        for (int i = threadIdx.x; i < nv; i += blockDim.x) verts[i] = cverts[v0 + pg + i];
        __syncthreads();
        for (int vb = 0; wave_on && vb < nv; vb += U) {
            int p[U];
            double z[U], px[U], py[U], pz[U], pw[U];
            int pb[U];
#pragma unroll
            for (int u = 0; u < U; ++u) p[u] = verts[min(vb + u, nv - 1)];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                z[u] = Z[(int64_t)p[u] * ldk + jc];
                px[u] = V[3 * p[u]];
                py[u] = V[3 * p[u] + 1];
                pz[u] = V[3 * p[u] + 2];
                pw[u] = vw[p[u]];
                pb[u] = isb[p[u]];
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int v = vb + u;  // uniform
                const double x = z[u] - base;
                const bool fin = isfinite(x);
                const bool in = v < nv;
                if (in && fin) {
                    lo = fmin(lo, x);
                    hi = fmax(hi, x);
                }
                if (!in || v >= nown) continue;
                if (!fin) {
                    nnf += 1;
                    continue;
                }
                const double dx = px[u] - xs, dy = py[u] - ys, dz = pz[u] - zs;
                const double chord = sqrt(dx * dx + dy * dy + dz * dz);
                if (x < -h) nneg += 1;
                if (x < 0.9 * chord - h) nviol += 1;
                if (x > eps) {
                    nval += 1;
                    sd += x;
                    sdw += x * pw[u];
                    sw += pw[u];
                }
                dmax = fmax(dmax, x);
                if (pb[u]) {
                    bmin = fmin(bmin, x);
                    nbf += 1;
                }
            }
        }
    }
    if (jl >= k) return;
    const double r[kChunkStats] = {sd, sdw, sw, dmax, bmin, (double)nneg, (double)nviol, (double)nval, (double)nnf,
                                   (double)nbf};
    double* out = part + chunk_rec_index(jl, c, 0, nch);  // the record's kChunkStats values, 8 apart
#pragma unroll
    for (int q = 0; q < kChunkStats; ++q) out[8 * q] = r[q];
    ch_lo[chunk_range_index(jl, c, nch)] = __double2float_rd(lo);
    ch_hi[chunk_range_index(jl, c, nch)] = __double2float_ru(hi);
}

}  // namespace

void launch_chunk_area(hipStream_t st, int m, const double* larea, double* chunk_area) {
    const int nch = (m + kWave * UNR - 1) / (kWave * UNR);
    if (nch <= 0) return;
    fcw::launch(chunk_area_kernel, dim3(dim3((nch + 255) / 256)), dim3(256), 0, st, m, larea, chunk_area);
}

void launch_chunk_stats(hipStream_t st, int k, int ldk, int nch, const double* Z, const int* src_pos, const double* V,
                        const double* vw, const int* isb, const int* cvptr, const int* cverts, const int* cnown,
                        double eps, double h, float* ch_lo, float* ch_hi, double* part) {
    if (k <= 0 || nch <= 0) return;
    fcw::launch(chunk_stats_kernel, dim3(dim3(nch, (k + 255) / 256)), dim3(256), 0, st, ldk, k, nch, Z, src_pos, V, vw, isb, cvptr, cverts,
                                                                   cnown, eps, h, ch_lo, ch_hi, part);
}

void launch_metrics(hipStream_t st, int k, const MeshDev& mesh, const MetricsDev& md, const double* Dt,
                    const int* src_pos, MetricsOut out, bool health_screen) {
    if (k <= 0) return;
    fcw::launch(metrics_kernel, dim3(dim3(k)), dim3(NT), 0, st, mesh, md, Dt, src_pos, out, health_screen ? 1 : 0);
}

}  // namespace fcw
