// CDNA4 (gfx950) kernels for the heat-method geodesic pipeline and the
// per-source wiring metrics. All arrays live in "position" numbering (the
// nested-dissection order of the host analysis); RHS blocks are row-major
// n x ldk (one vertex per row, one source per column).
#include <hip/hip_runtime.h>

#include <type_traits>

#include <climits>
#include <cmath>
#include <cstdint>

#include "fcw_device.h"

namespace fcw {
namespace {

// ---------------------------------------------------------------------------
// Operators (fastcw_lean.py:375-413, lean_geometry.py:44-63)
// ---------------------------------------------------------------------------

__global__ void face_geometry_kernel(int m, const double* __restrict__ V, const int* __restrict__ faces,
                                     double* __restrict__ area, double* __restrict__ gram,
                                     double* __restrict__ larea, double* __restrict__ g11,
                                     double* __restrict__ g12, double* __restrict__ g22) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m) return;
    const int i0 = faces[3 * f], i1 = faces[3 * f + 1], i2 = faces[3 * f + 2];
    double p0[3], e1[3], e2[3];
    for (int c = 0; c < 3; ++c) {
        p0[c] = V[3 * i0 + c];
        e1[c] = V[3 * i1 + c] - p0[c];
        e2[c] = V[3 * i2 + c] - p0[c];
    }
    double nr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const double nl = sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
    const double A = 0.5 * nl;
    double nh[3] = {nr[0] / nl, nr[1] / nl, nr[2] / nl};
    const double inv2a = 1.0 / (2.0 * A);
    double b1[3] = {(e2[1] * nh[2] - e2[2] * nh[1]) * inv2a, (e2[2] * nh[0] - e2[0] * nh[2]) * inv2a,
                    (e2[0] * nh[1] - e2[1] * nh[0]) * inv2a};
    double b2[3] = {(nh[1] * e1[2] - nh[2] * e1[1]) * inv2a, (nh[2] * e1[0] - nh[0] * e1[2]) * inv2a,
                    (nh[0] * e1[1] - nh[1] * e1[0]) * inv2a};
    area[f] = A;
    // Gram matrix of the basis gradients grad B0 = -b1 - b2, grad B1 = b1, grad B2 = b2
    double g[3][3];
    for (int c = 0; c < 3; ++c) {
        g[0][c] = -b1[c] - b2[c];
        g[1][c] = b1[c];
        g[2][c] = b2[c];
    }
    double* G = gram + 6 * (int64_t)f;
    int q = 0;
    for (int a = 0; a < 3; ++a)
        for (int b = a; b < 3; ++b) G[q++] = g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2];
    const double a11 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
    const double a22 = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
    const double a12 = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
    const double det = a11 * a22 - a12 * a12;
    larea[f] = 0.5 * sqrt(fmax(det, 0.0));
    const double inv = det > 0.0 ? 1.0 / det : 0.0;
    g11[f] = a22 * inv;
    g12[f] = -a12 * inv;
    g22[f] = a11 * inv;
}

// One thread per CSR entry: L_e = sum over (face, a, b) of A_f gradB_a . gradB_b (deterministic order).
// Entry (a, b) of a face's symmetric 3 x 3 Gram matrix, stored G00 G01 G02 G11 G12 G22.
__device__ inline int gram_index(int a, int b) {
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    return (lo * (5 - lo)) / 2 + hi;
}

// L = sum over faces of A_f grad B_a . grad B_b (fastcw_lean.py:394-413), one CSR entry per thread,
// its (face, corner pair) contributions in a fixed order.
__global__ void assemble_l_kernel(int nnz, const int* __restrict__ cptr, const int* __restrict__ contrib,
                                  const double* __restrict__ area, const double* __restrict__ gram,
                                  double* __restrict__ lval) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    double s = 0.0;
    for (int c = cptr[e]; c < cptr[e + 1]; ++c) {
        const int code = contrib[c];
        const int f = code / 9, ab = code - 9 * f, a = ab / 3, b = ab - 3 * a;
        s += area[f] * gram[6 * (int64_t)f + gram_index(a, b)];
    }
    lval[e] = s;
}

__global__ void lumped_mass_kernel(int n, const int* __restrict__ vf_ptr, const int* __restrict__ vf,
                                   const double* __restrict__ area, double* __restrict__ mass) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double s = 0.0;
    for (int e = vf_ptr[p]; e < vf_ptr[p + 1]; ++e) s += area[vf[e] / 3] / 3.0;
    mass[p] = s;
}

__global__ void fill_kernel(double* __restrict__ p, int64_t count, double value) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

__global__ void calibration_copy_kernel(const double4* __restrict__ src, double4* __restrict__ dst, int64_t count) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void delta_rhs_kernel(double* __restrict__ X, int ldk, const int* __restrict__ src_pos, int k) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < k) X[(int64_t)src_pos[j] * ldk + j] = 1.0;
}

// ---------------------------------------------------------------------------
// Level-scheduled supernodal sweeps with partitioned-inverse panels
// ---------------------------------------------------------------------------

// Zero the forward-sweep rows (Y rows, and with U set the update rows) of the listed supernodes.
__global__ __launch_bounds__(256) void clear_rows_kernel(const int* __restrict__ nodes, int ldk,
                                                         const int* __restrict__ node_start,
                                                         const int* __restrict__ node_s, const int* __restrict__ node_b,
                                                         const int64_t* __restrict__ node_uoff, double* __restrict__ Y,
                                                         double* __restrict__ U) {
    const int id = nodes[blockIdx.x];
    const int col = blockIdx.y * 64 + (threadIdx.x & 63);
    const int rg = threadIdx.x >> 6;
    const int s = node_s[id], b = node_b[id], start = node_start[id];
    for (int r = rg; r < s; r += 4) Y[(int64_t)(start + r) * ldk + col] = 0.0;
    if (!U) return;
    double* Ut = U + node_uoff[id] * ldk;
    for (int r = rg; r < b; r += 4) Ut[(int64_t)r * ldk + col] = 0.0;
}

// ---------------------------------------------------------------------------
// Fused gradient -> normalise -> divergence (fastcw_lean.py:416-442), gathered per vertex
// ---------------------------------------------------------------------------
//
// Per incidence (face f, corner a) of a vertex row the term needs five coefficients of the Gram
// matrix of the face's gradient basis (g1 = grad B1, g2 = grad B2, ga = grad Ba) and its area A
// (extrinsic or intrinsic geometry alike):
//     |grad u|^2        = du1^2 |g1|^2 + du1 du2 (2 g1.g2) + du2^2 |g2|^2
//     A (grad u . ga)   = du1 (A g1.ga) + du2 (A g2.ga)
// with du1 = u1 - u0, du2 = u2 - u0, so the per-column work is two differences, two short dot
// products and one reciprocal square root. The coefficients are computed once per mesh
// (incidence_geometry_kernel), in the incidence (CSR) order of the rows.
//
// The rows are cut, once per mesh, into groups of consecutive rows whose incidences touch at most
// kDivVerts distinct vertices (at most kDivInc incidences, kDivGroupRows rows); with the
// nested-dissection order a group of about 20 rows of a degree-6 mesh touches about 57 vertices.
// A workgroup takes one (group, 64-column slice): it loads each distinct vertex's 64 values once
// into LDS (one 512-byte run per vertex, all loads in flight together) with each incidence's three
// local vertex indices, then each wave sums rows of the group, one lane per column, reading the
// values from LDS and each incidence's five coefficients (the same for every lane) by scalar loads
// into scalar registers: as LDS broadcasts they took about two thirds of the LDS cycles. A vertex
// row is read from L2 about 2.7 times per slice instead of 18 times (three per incidence). Each
// row sums its incidences in CSR order, so the result does not depend on scheduling. A row that
// alone exceeds the limits (a vertex of degree above 63) forms a group of its own that reads its
// values from global memory.
//
// Workgroups are handed to the 8 XCDs round robin; the grid maps them so that each XCD walks its
// own contiguous run of (group, slice) items, the slices of a group one after another, and the
// neighbour rows that consecutive groups share are read from that XCD's L2.
// 1/sqrt(q) for q > 0: the hardware estimate (relative error below 6e-8, measured on the device
// over 2^20 values spanning 2^-700..2^701) refined by one Newton step, which squares the error
// (below 5e-15 measured), with q scaled out of the range where the estimate loses accuracy. It
// replaces a correctly rounded square root followed by a division.
__device__ inline double rsqrt_refined(double q) {
    const bool tiny = q < 0x1p-767;
    const double qs = tiny ? q * 0x1p+256 : q;
    double y = __builtin_amdgcn_rsq(qs);
    y = fma(0.5 * y, fma(-qs * y, y, 1.0), y);
    return tiny ? y * 0x1p+128 : y;
}

// One incidence's term -A (grad u . ga) / |grad u| for one column, added to the row sum as
// fma(-A (grad u . ga), 1 / |grad u|, acc) (one rounding per term, the same in every path); no term
// where grad u vanishes (|grad u| > 1e-300, fastcw_lean.py:433, holds for every positive q).
template <typename GP>  // pointer to the five coefficients (generic or constant address space)
__device__ inline double divergence_add(double acc, double u0, double u1, double u2, GP G) {
    const double du1 = u1 - u0;
    const double du2 = u2 - u0;
    const double num = fma(du2, G[1], du1 * G[0]);
    const double qq = fma(du1, fma(du2, G[3], du1 * G[2]), du2 * (du2 * G[4]));
    return qq > 0.0 ? fma(-num, rsqrt_refined(qq), acc) : acc;
}

// 1/sqrt(q) for q well inside the range where the hardware estimate is accurate (q >= 2^-767):
// the estimate refined by one Newton step, as rsqrt_refined.
__device__ inline double rsqrt_newton(double q) {
    const double y = __builtin_amdgcn_rsq(q);
    return fma(0.5 * y, fma(-q * y, y, 1.0), y);
}

constexpr int kXcds = 8;
constexpr int kDivSlice = 64;  // columns per workgroup (one per lane)
constexpr int kDivC = 6;       // incidences per row step (interior vertices have about 6)

// geo[5 e + 0..4] = A g1.ga, A g2.ga, |g1|^2, 2 g1.g2, |g2|^2 of incidence e = vf[e] (face*3 + a).
__global__ void incidence_geometry_kernel(int ninc, const int* __restrict__ vf, const double* __restrict__ area,
                                          const double* __restrict__ gram, double* __restrict__ geo) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ninc) return;
    const int code = vf[e];
    const int f = code / 3, a = code - 3 * f;
    const double* Gp = gram + 6 * (int64_t)f;  // G00 G01 G02 G11 G12 G22
    const double A = area[f];
    const double g1a = a == 0 ? Gp[1] : (a == 1 ? Gp[3] : Gp[4]);
    const double g2a = a == 0 ? Gp[2] : (a == 1 ? Gp[4] : Gp[5]);
    double* o = geo + 5 * (int64_t)e;
    o[0] = A * g1a;
    o[1] = A * g2a;
    o[2] = Gp[3];
    o[3] = 2.0 * Gp[4];
    o[4] = Gp[5];
}

__global__ __launch_bounds__(256) void divergence_kernel(int nitems, int ncs, int run, DivDev d, const double* __restrict__ u,
                                                         double* __restrict__ out, int ldk, float* __restrict__ out32) {
    __shared__ double vals[kDivVerts][kDivSlice];  // the group's distinct vertex rows, this slice
    // per incidence: the LDS byte offsets of its three vertex rows in vals, (o0 | o1 << 16, o2)
    __shared__ uint2 soff[kDivInc];
    __shared__ int rptr[kDivGroupRows + 1];
    __shared__ int sexp[4][kDivSlice];   // per wave and column: largest exponent of its staged values

    const int item = (blockIdx.x % kXcds) * run + blockIdx.x / kXcds;
    if (item >= nitems) return;
    const int g = item / ncs, cs = item - g * ncs;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int4 G = d.grp[g];
    const int r0 = G.x, e0 = G.y, v0 = G.z, nv = G.w;
    const int nr = d.grp[g + 1].x - r0;
    const int col = cs * kDivSlice + lane;

    if (nv < 0) {
        // a row beyond the limits, alone: its values straight from global memory
        if (w != 0) return;
        const int e1 = d.vf_ptr[r0 + 1];
        double acc = 0.0;
        for (int e = e0; e < e1; ++e) {
            const int f = d.vf[e] / 3;
            const int* fv = d.faces + 3 * (int64_t)f;
            const double u0 = u[(int64_t)fv[0] * ldk + col];
            const double u1 = u[(int64_t)fv[1] * ldk + col];
            const double u2 = u[(int64_t)fv[2] * ldk + col];
            acc = divergence_add(acc, u0, u1, u2, d.geo + 5 * (int64_t)e);
        }
        out[(int64_t)r0 * ldk + col] = acc;
        if (out32) out32[(int64_t)r0 * ldk + col] = (float)acc;
        return;
    }

    if (tid <= nr) rptr[tid] = d.vf_ptr[r0 + tid] - e0;
    // Each term is unchanged when the values are scaled by a power of two (exactly: the
    // differences, the quadratic form and the reciprocal square root scale exactly), so each
    // column's values are stored scaled by its largest magnitude in the group: the group spans a few
    // mesh spacings, over which u changes by a few decades at most, so |grad u|^2 stays far above
    // the range where the hardware estimate needs the rescaling of rsqrt_refined. A wave that meets
    // such a value anyway (a vanishing gradient) sums the row again the general way.
    // the wave index as a uniform value: vertex indices, row bounds and incidence offsets below are
    // then scalar values and the loops over them scalar branches
    const int wu = __builtin_amdgcn_readfirstlane(w);
    if (nv > 0) {
        // wave w loads vertices w, w + 4, ...: all its loads first (in-range repeats of the last
        // vertex pad the count), then the LDS stores
        constexpr int Q = kDivVerts / 4;
        double x[Q];
        const int* gv = d.verts + v0;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int v = __builtin_amdgcn_readfirstlane(gv[min(wu + 4 * q, nv - 1)]);
            x[q] = u[(int64_t)v * ldk + col];
        }
        // largest exponent of the lane's values: that of the largest high word of their magnitudes
        // (the padding repeats a staged vertex, so it changes nothing); none when every high word
        // is zero (values of zero, or below 2^-1054)
        unsigned hmax = 0;
#pragma unroll
        for (int q = 0; q < Q; ++q) hmax = max(hmax, (unsigned)__double2hiint(x[q]) & 0x7fffffffu);
        int emax = INT_MIN;
        if (hmax != 0) (void)frexp(__hiloint2double((int)hmax, 0), &emax);
        sexp[wu][lane] = emax;
        __syncthreads();
        emax = max(max(sexp[0][lane], sexp[1][lane]), max(sexp[2][lane], sexp[3][lane]));
        const double sc = ldexp(1.0, emax == INT_MIN ? 0 : -emax);
#pragma unroll
        for (int q = 0; q < Q; ++q)
            if (wu + 4 * q < nv) vals[wu + 4 * q][lane] = x[q] * sc;
    }
    {
        const int ni = d.vf_ptr[r0 + nr] - e0;
        for (int i = tid; i < ni; i += blockDim.x) {
            const unsigned L = d.loc[e0 + i];
            constexpr unsigned kRow = kDivSlice * sizeof(double);
            soff[i] = make_uint2((L & 63u) * kRow | (((L >> 6) & 63u) * kRow) << 16, ((L >> 12) & 63u) * kRow);
        }
    }
    __syncthreads();
    // One incidence of this lane's column: -A (grad u . ga) and 1 / |grad u|. The
    // fast path assumes |grad u|^2 >= 2^-767 (the scaled values keep it far above that) and only
    // records whether an incidence breaks that; a row with such an incidence in any lane is summed
    // again the general way (no term where grad u vanishes, rescaled estimate where it is tiny),
    // which adds every other term with the same bits.
    const char* const vrow = reinterpret_cast<const char*>(&vals[0][0]) + lane * sizeof(double);
    auto val = [&](unsigned off) { return *reinterpret_cast<const double*>(vrow + off); };
    // The coefficients are not written while the kernel runs, so they are read through the
    // constant address space, whose wave-uniform loads are scalar loads.
    typedef const __attribute__((address_space(4))) double cf64;
    cf64* const ggeo = (cf64*)(d.geo + 5 * (int64_t)e0);
    // C incidences from e: the step's coefficients (scalar loads) and offsets first, then every
    // value read, then the arithmetic; acc += each term in order
    auto step = [&](auto width, int e, double acc, uint64_t& odd) -> double {
        constexpr int C = decltype(width)::value;
        double gc[C][5];
#pragma unroll
        for (int q = 0; q < C; ++q)
#pragma unroll
            for (int c = 0; c < 5; ++c) gc[q][c] = ggeo[5 * (e + q) + c];
        __builtin_amdgcn_sched_barrier(0);
        uint2 P[C];
#pragma unroll
        for (int q = 0; q < C; ++q) P[q] = soff[e + q];
        double u0[C], u1[C], u2[C];
#pragma unroll
        for (int q = 0; q < C; ++q) {
            u0[q] = val(P[q].x & 0xffffu);
            u1[q] = val(P[q].x >> 16);
            u2[q] = val(P[q].y);
        }
        __builtin_amdgcn_sched_barrier(0);  // every value read is issued before any is used
        double nn[C], rr[C];
#pragma unroll
        for (int q = 0; q < C; ++q) {
            const double* G = gc[q];
            const double du1 = u1[q] - u0[q], du2 = u2[q] - u0[q];
            nn[q] = -fma(du2, G[1], du1 * G[0]);
            const double qq = fma(du1, fma(du2, G[3], du1 * G[2]), du2 * (du2 * G[4]));
            odd |= __ballot(!(qq >= 0x1p-767));
            rr[q] = rsqrt_newton(qq);
        }
#pragma unroll
        for (int q = 0; q < C; ++q) acc = fma(nn[q], rr[q], acc);
        return acc;
    };
    for (int r = wu; r < nr; r += 4) {
        const int rs = __builtin_amdgcn_readfirstlane(rptr[r]), re = __builtin_amdgcn_readfirstlane(rptr[r + 1]);
        double acc = 0.0;
        uint64_t odd = 0;  // lanes with an incidence off the fast path
        // kDivC incidences per step, then the rest one by one (in CSR order throughout)
        int e = rs;
        for (; e + kDivC <= re; e += kDivC) acc = step(std::integral_constant<int, kDivC>{}, e, acc, odd);
        for (; e < re; ++e) acc = step(std::integral_constant<int, 1>{}, e, acc, odd);
        if (odd) {
            acc = 0.0;
            for (int ec = rs; ec < re; ++ec) {
                const uint2 P = soff[ec];
                acc = divergence_add(acc, *reinterpret_cast<const double*>(vrow + (P.x & 0xffffu)),
                                     *reinterpret_cast<const double*>(vrow + (P.x >> 16)),
                                     *reinterpret_cast<const double*>(vrow + P.y), ggeo + 5 * ec);
            }
        }
        out[(int64_t)(r0 + r) * ldk + col] = acc;
        if (out32) out32[(int64_t)(r0 + r) * ldk + col] = (float)acc;
    }
}

// ---------------------------------------------------------------------------
// Refined (FP32 + FP64 residual) solve support
// ---------------------------------------------------------------------------
// r[p] = b[p] - sum_q A[p][q] x[q] with A = L + shift M (CSR of L in positions, lumped M), in
// FP64 and in CSR order, then rounded to FP32. One wave per row, one lane per RHS column.
__global__ __launch_bounds__(256) void residual32_kernel(int n, int ldk, const int* __restrict__ rowptr,
                                                         const int* __restrict__ colind, const double* __restrict__ lval,
                                                         const double* __restrict__ mass, double shift,
                                                         const double* __restrict__ b, const double* __restrict__ x,
                                                         float* __restrict__ r) {
    const int col = blockIdx.y * 64 + (threadIdx.x & 63);
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    double v = b[(int64_t)row * ldk + col];
    if (x) {
        const int e0 = rowptr[row], e1 = rowptr[row + 1];
        double ax = 0.0;
        constexpr int kC = 8;
        for (int e = e0; e < e1; e += kC) {
            double xv[kC], a[kC];
#pragma unroll
            for (int q = 0; q < kC; ++q) {
                const int ec = min(e + q, e1 - 1);
                const int cq = colind[ec];
                a[q] = lval[ec] + (cq == row ? shift * mass[row] : 0.0);
                xv[q] = x[(int64_t)cq * ldk + col];
            }
#pragma unroll
            for (int q = 0; q < kC; ++q)
                if (e + q < e1) ax = fma(a[q], xv[q], ax);
        }
        v -= ax;
    }
    r[(int64_t)row * ldk + col] = (float)v;
}

// The same residual over the divergence's row groups: a workgroup stages the values of x at the
// group's distinct vertices for one 64-column slice in LDS (every row's matrix neighbours are
// vertices of its faces, so they are among them), with the group's row pointers and local column
// indices (when the group has at most kResEntries entries), then each wave takes rows of the
// group, one lane per column, reading x from LDS and the matrix values (the same for every lane)
// by scalar loads. aval = the CSR values of L + shift M (shift on the diagonal); cloc = each CSR
// entry's local vertex index in its row's group. Each row is summed in CSR order (fused
// multiply-adds), so the result does not depend on scheduling.
constexpr int kResEntries = 512;
__global__ __launch_bounds__(256) void residual_grouped_kernel(int nitems, int ncs, int run, DivDev d,
                                                               const int* __restrict__ rowptr,
                                                               const int* __restrict__ colind,
                                                               const unsigned char* __restrict__ cloc,
                                                               const double* __restrict__ aval,
                                                               const double* __restrict__ b,
                                                               const double* __restrict__ x, float* __restrict__ r,
                                                               int ldk) {
    __shared__ double vals[kDivVerts][kDivSlice];
    __shared__ unsigned char sc[kResEntries];
    __shared__ int rptr[kDivGroupRows + 1];
    const int item = (blockIdx.x % kXcds) * run + blockIdx.x / kXcds;
    if (item >= nitems) return;
    const int g = item / ncs, cs = item - g * ncs;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int4 G = d.grp[g];
    const int r0 = G.x, v0 = G.z, nv = G.w;
    const int nr = d.grp[g + 1].x - r0;
    const int col = cs * kDivSlice + lane;
    if (nv < 0) {
        // a row beyond the group limits, alone: x straight from global memory
        if (w != 0) return;
        double ax = 0.0;
        for (int e = rowptr[r0]; e < rowptr[r0 + 1]; ++e) ax = fma(aval[e], x[(int64_t)colind[e] * ldk + col], ax);
        r[(int64_t)r0 * ldk + col] = (float)(b[(int64_t)r0 * ldk + col] - ax);
        return;
    }
    const int eg = rowptr[r0], ne = rowptr[r0 + nr] - eg;
    const bool staged = ne <= kResEntries;  // the same for the whole workgroup
    if (tid <= nr) rptr[tid] = rowptr[r0 + tid] - eg;
    if (staged)
        for (int i = tid; i < ne; i += blockDim.x) sc[i] = cloc[eg + i];
    const int wu = __builtin_amdgcn_readfirstlane(w);  // vertex indices and row bounds as scalars
    if (nv > 0) {
        constexpr int Q = kDivVerts / 4;
        double xv[Q];
        const int* gv = d.verts + v0;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int v = __builtin_amdgcn_readfirstlane(gv[min(wu + 4 * q, nv - 1)]);
            xv[q] = x[(int64_t)v * ldk + col];
        }
#pragma unroll
        for (int q = 0; q < Q; ++q)
            if (wu + 4 * q < nv) vals[wu + 4 * q][lane] = xv[q];
    }
    __syncthreads();
    // The matrix values (the same for every lane, not written while the kernel runs) are read
    // through the constant address space: scalar loads into scalar registers, once per step.
    typedef const __attribute__((address_space(4))) double cf64;
    cf64* const ga = (cf64*)(aval + eg);
    // Steps of kC entries; past the row's end a step repeats its last entry with the value 0, whose
    // product adds nothing (exactly: the values are finite and the sum starts at +0).
    auto rows = [&](auto local) {
        constexpr bool kStaged = decltype(local)::value;
        for (int rr = wu; rr < nr; rr += 4) {
            const int p = r0 + rr;
            const double bv = b[(int64_t)p * ldk + col];
            const int e0 = __builtin_amdgcn_readfirstlane(rptr[rr]), e1 = __builtin_amdgcn_readfirstlane(rptr[rr + 1]);
            double ax = 0.0;
            constexpr int kC = 8;
            for (int e = e0; e < e1; e += kC) {
                double a[kC], xq[kC];
                unsigned c[kC];
#pragma unroll
                for (int q = 0; q < kC; ++q) a[q] = ga[min(e + q, e1 - 1)];
                __builtin_amdgcn_sched_barrier(0);  // the step's loads issued together
#pragma unroll
                for (int q = 0; q < kC; ++q) a[q] = e + q < e1 ? a[q] : 0.0;
#pragma unroll
                for (int q = 0; q < kC; ++q) {
                    const int ec = min(e + q, e1 - 1);
                    c[q] = kStaged ? sc[ec] : cloc[eg + ec];
                }
#pragma unroll
                for (int q = 0; q < kC; ++q) xq[q] = vals[c[q]][lane];
#pragma unroll
                for (int q = 0; q < kC; ++q) ax = fma(a[q], xq[q], ax);
            }
            r[(int64_t)p * ldk + col] = (float)(bv - ax);
        }
    };
    if (staged)
        rows(std::true_type{});
    else
        rows(std::false_type{});
}

// aval[e] = lval[e] plus shift * mass[p] on row p's diagonal entry (the residual's matrix values).
__global__ void shifted_values_kernel(int n, const int* __restrict__ rowptr, const int* __restrict__ colind,
                                      const double* __restrict__ lval, const double* __restrict__ mass, double shift,
                                      double* __restrict__ aval) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    for (int e = rowptr[p]; e < rowptr[p + 1]; ++e) aval[e] = lval[e] + (colind[e] == p ? shift * mass[p] : 0.0);
}

__global__ void to_float_kernel(const double* __restrict__ src, float* __restrict__ dst, int64_t count) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}

// out[j][v] = Dt[j][iperm[v]] - phi[src_pos[j]][j] (fastcw_lean.py:526: each field minus its value at
// the source), rows of Dt in positions, rows of out in the caller's vertex order.
__global__ void gather_rows_kernel(int n, int ldk, int k, const double* __restrict__ Dt, const double* __restrict__ phi,
                                   const int* __restrict__ src_pos, const int* __restrict__ iperm,
                                   double* __restrict__ out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (v < n && j < k) out[(int64_t)j * n + v] = Dt[(int64_t)j * n + iperm[v]] - phi[(int64_t)src_pos[j] * ldk + j];
}


inline dim3 grid1(int64_t n, int bs) { return dim3((unsigned)((n + bs - 1) / bs)); }

}  // namespace

void launch_face_geometry(hipStream_t st, int m, const double* V, const int* faces, double* area, double* gram,
                          double* larea, double* g11, double* g12, double* g22) {
    fcw::launch(face_geometry_kernel, dim3(grid1(m, 256)), dim3(256), 0, st, m, V, faces, area, gram, larea, g11, g12, g22);
}

void launch_assemble(hipStream_t st, int n, int nnz, const int* contrib_ptr, const int* contrib, const double* area,
                     const double* gram, const int* vf_ptr, const int* vf, double* lval, double* mass) {
    fcw::launch(assemble_l_kernel, dim3(grid1(nnz, 256)), dim3(256), 0, st, nnz, contrib_ptr, contrib, area, gram, lval);
    fcw::launch(lumped_mass_kernel, dim3(grid1(n, 256)), dim3(256), 0, st, n, vf_ptr, vf, area, mass);
}

void launch_fill(hipStream_t st, double* p, int64_t count, double value) {
    if (count <= 0) return;
    int64_t blocks = (count + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    fcw::launch(fill_kernel, dim3(dim3((unsigned)blocks)), dim3(256), 0, st, p, count, value);
}

void launch_calibration_copy(hipStream_t st, const void* src, void* dst, int64_t bytes) {
    const int64_t count = bytes / (int64_t)sizeof(double4);
    fcw::launch(calibration_copy_kernel, dim3(dim3(4096)), dim3(256), 0, st, (const double4*)src, (double4*)dst, count);
}

void launch_set_delta(hipStream_t st, double* X, int ldk, const int* src_pos, int k) {
    fcw::launch(delta_rhs_kernel, dim3(grid1(k, 256)), dim3(256), 0, st, X, ldk, src_pos, k);
}

void launch_delta_rhs(hipStream_t st, double* X, int n, int ldk, const int* src_pos, int k) {
    launch_fill(st, X, (int64_t)n * ldk, 0.0);
    fcw::launch(delta_rhs_kernel, dim3(grid1(k, 256)), dim3(256), 0, st, X, ldk, src_pos, k);
}

void launch_clear_rows(hipStream_t st, const int* nodes, int count, int ldk, const int* node_start,
                       const int* node_s, const int* node_b, const int64_t* node_uoff, double* Y, double* U) {
    if (count <= 0) return;
    fcw::launch(clear_rows_kernel, dim3(dim3(count, ldk / 64)), dim3(256), 0, st, nodes, ldk, node_start, node_s, node_b, node_uoff, Y, U);
}

void launch_incidence_geometry(hipStream_t st, int ninc, const int* vf, const double* area, const double* gram,
                               double* geo) {
    if (ninc <= 0) return;
    fcw::launch(incidence_geometry_kernel, dim3(grid1(ninc, 256)), dim3(256), 0, st, ninc, vf, area, gram, geo);
}

void launch_divergence(hipStream_t st, const DivDev& d, int ldk, const double* u, double* out, float* out32) {
    if (d.ngroups <= 0) return;
    const int ncs = ldk / kDivSlice;
    const int nitems = d.ngroups * ncs;
    const int run = (nitems + kXcds - 1) / kXcds;  // items per XCD
    fcw::launch(divergence_kernel, dim3(dim3(run * kXcds)), dim3(256), 0, st, nitems, ncs, run, d, u, out, ldk, out32);
}

void launch_residual32(hipStream_t st, int n, int ldk, const int* rowptr, const int* colind, const double* lval,
                       const double* mass, double shift, const double* b, const double* x, float* r) {
    fcw::launch(residual32_kernel, dim3(dim3((n + 3) / 4, ldk / 64)), dim3(256), 0, st, n, ldk, rowptr, colind, lval, mass, shift, b, x, r);
}

void launch_residual_grouped(hipStream_t st, const DivDev& d, int ldk, const int* rowptr, const int* colind,
                             const unsigned char* cloc, const double* aval, const double* b, const double* x,
                             float* r) {
    if (d.ngroups <= 0) return;
    const int ncs = ldk / kDivSlice;
    const int nitems = d.ngroups * ncs;
    const int run = (nitems + kXcds - 1) / kXcds;  // items per XCD
    fcw::launch(residual_grouped_kernel, dim3(dim3(run * kXcds)), dim3(256), 0, st, nitems, ncs, run, d, rowptr,
                colind, cloc, aval, b, x, r, ldk);
}

void launch_shifted_values(hipStream_t st, int n, const int* rowptr, const int* colind, const double* lval,
                           const double* mass, double shift, double* aval) {
    if (n <= 0) return;
    fcw::launch(shifted_values_kernel, dim3(grid1(n, 256)), dim3(256), 0, st, n, rowptr, colind, lval, mass, shift,
                aval);
}

void launch_to_float(hipStream_t st, const double* src, float* dst, int64_t count) {
    if (count <= 0) return;
    fcw::launch(to_float_kernel, dim3(grid1(count, 256)), dim3(256), 0, st, src, dst, count);
}

void launch_gather_rows(hipStream_t st, int n, int ldk, int k, const double* Dt, const double* phi, const int* src_pos,
                        const int* iperm, double* out) {
    fcw::launch(gather_rows_kernel, dim3(dim3((n + 255) / 256, k)), dim3(256), 0, st, n, ldk, k, Dt, phi, src_pos, iperm, out);
}


}  // namespace fcw
