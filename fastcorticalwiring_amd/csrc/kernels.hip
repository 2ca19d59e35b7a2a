// CDNA4 (gfx950) kernels for the heat-method geodesic pipeline and the
// per-source wiring metrics. All arrays live in "position" numbering (the
// nested-dissection order of the host analysis); RHS blocks are row-major
// n x ldk (one vertex per row, one source per column).
#include <hip/hip_runtime.h>

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

// Zero the forward-sweep rows (Y rows and update rows) of the listed supernodes.
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
    double* Ut = U + node_uoff[id] * ldk;
    for (int r = rg; r < b; r += 4) Ut[(int64_t)r * ldk + col] = 0.0;
}

// ---------------------------------------------------------------------------
// Fused gradient -> normalise -> divergence (fastcw_lean.py:416-442), gathered per vertex
// ---------------------------------------------------------------------------
//
// A workgroup owns kDivRows consecutive vertex rows and every column of the RHS block. It first
// stages, cooperatively and in one pass, every incidence (face f, corner a) of its rows in LDS:
// the byte offsets of the face's three vertex rows and five coefficients from the Gram matrix of
// the face's gradient basis (g1 = grad B1, g2 = grad B2, ga = grad Ba) and its area A (extrinsic
// or intrinsic geometry alike):
//     |grad u|^2        = du1^2 |g1|^2 + du1 du2 (2 g1.g2) + du2^2 |g2|^2
//     A (grad u . ga)   = du1 (A g1.ga) + du2 (A g2.ga)
// with du1 = u1 - u0, du2 = u2 - u0, so the per-column work is two differences, two short dot
// products and one reciprocal square root. The work is then (row, 64-column slice) pairs dealt
// to the waves round robin (with 256 columns, wave w takes slice w of every row), one lane per
// column: a row's field values are loaded kDivC incidences at a time with all loads in flight
// together (32-bit offsets through a buffer resource, no address arithmetic), and every face
// quantity comes from LDS (a broadcast read). Each row sums its incidences in CSR order, so the
// result does not depend on scheduling.
//
// Workgroups are handed to the 8 XCDs round robin; the grid maps them so that each XCD walks
// its own contiguous run of row groups, and the neighbour rows that consecutive groups share
// are read from that XCD's L2.
// 1/sqrt(q) for q > 0: the hardware estimate refined by two Newton steps (relative error a few
// units in the last place), with q scaled out of the range where the estimate loses accuracy.
// It replaces a correctly rounded square root followed by a division.
__device__ inline double rsqrt_refined(double q) {
    const bool tiny = q < 0x1p-767;
    const double qs = tiny ? q * 0x1p+256 : q;
    double y = __builtin_amdgcn_rsq(qs);
    y = fma(0.5 * y, fma(-qs * y, y, 1.0), y);
    y = fma(0.5 * y, fma(-qs * y, y, 1.0), y);
    return tiny ? y * 0x1p+128 : y;
}

constexpr int kDivRows = 16;   // rows per workgroup
constexpr int kDivInc = 256;   // incidences staged per pass (a pass per 256 when rows are long)
constexpr int kDivC = 6;       // incidences in flight per row step (interior vertices have about 6)
constexpr int kXcds = 8;
constexpr unsigned kBufWord3 = 0x00020000u;  // gfx9 buffer descriptor word 3 (raw 32-bit loads)

// BUF: the field block is below 2 GiB, so loads go through a buffer resource with 32-bit
// offsets; otherwise through 64-bit pointers (offsets still below 4 GiB).
template <bool BUF>
__global__ __launch_bounds__(256) void divergence_kernel(int n, int ldk, int ngroups, int run,
                                                         const int* __restrict__ vf_ptr, const int* __restrict__ vf,
                                                         const int* __restrict__ faces, const double* __restrict__ area,
                                                         const double* __restrict__ gram, const double* __restrict__ u,
                                                         double* __restrict__ out) {
    __shared__ uint4 rec[kDivInc];       // byte offsets of the face's vertex rows 0, 1, 2
    __shared__ double geo[kDivInc][5];   // A g1.ga, A g2.ga, |g1|^2, 2 g1.g2, |g2|^2
    __shared__ int rptr[kDivRows + 1];

    const int ncs = ldk / 64;
    const int grp = (blockIdx.x % kXcds) * run + blockIdx.x / kXcds;
    if (grp >= ngroups) return;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int r0 = grp * kDivRows, r1 = min(n, r0 + kDivRows);
    if (tid <= r1 - r0) rptr[tid] = vf_ptr[r0 + tid];
    __syncthreads();
    const int eb = rptr[0], ee = rptr[r1 - r0];
    const int npairs = (r1 - r0) * ncs;
    const unsigned row_bytes = (unsigned)ldk * 8u;
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void*)u, (short)0, BUF ? (int)(row_bytes * (unsigned)n) : 0, kBufWord3);

    for (int base = eb; base < ee; base += kDivInc) {
        const int cnt = min(kDivInc, ee - base);
        if (base != eb) __syncthreads();  // the previous pass is done with the staged data
        if (tid < cnt) {
            const int code = vf[base + tid];
            const int f = code / 3, a = code - 3 * f;
            const int i0 = faces[3 * (int64_t)f], i1 = faces[3 * (int64_t)f + 1], i2 = faces[3 * (int64_t)f + 2];
            const double* Gp = gram + 6 * (int64_t)f;  // G00 G01 G02 G11 G12 G22
            double G[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) G[c] = Gp[c];
            const double A = area[f];
            const double g1a = a == 0 ? G[1] : (a == 1 ? G[3] : G[4]);
            const double g2a = a == 0 ? G[2] : (a == 1 ? G[4] : G[5]);
            rec[tid] = make_uint4((unsigned)i0 * row_bytes, (unsigned)i1 * row_bytes, (unsigned)i2 * row_bytes, 0u);
            geo[tid][0] = A * g1a;
            geo[tid][1] = A * g2a;
            geo[tid][2] = G[3];
            geo[tid][3] = 2.0 * G[4];
            geo[tid][4] = G[5];
        }
        __syncthreads();
        for (int pr = w; pr < npairs; pr += 4) {
            const int r = pr / ncs, cs = pr - r * ncs;
            const unsigned cb = (unsigned)(cs * 64 + lane) * 8u;  // this lane's column, in bytes
            const int rs = max(rptr[r], base), re = min(rptr[r + 1], base + cnt);
            if (rs >= re) continue;  // uniform over the wave
            // a row split over staging passes continues the running sum the lane stored before
            double* o = out + (int64_t)(r0 + r) * ldk + cs * 64 + lane;
            double acc = rs == rptr[r] ? 0.0 : *o;
            for (int e = rs; e < re; e += kDivC) {
                double u0[kDivC], u1[kDivC], u2[kDivC];
#pragma unroll
                for (int q = 0; q < kDivC; ++q) {
                    const uint4 R = rec[min(e + q, re - 1) - base];
                    if (BUF) {
                        u0[q] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc, R.x + cb, 0, 0));
                        u1[q] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc, R.y + cb, 0, 0));
                        u2[q] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc, R.z + cb, 0, 0));
                    } else {
                        const char* ub = reinterpret_cast<const char*>(u) + cb;
                        u0[q] = *reinterpret_cast<const double*>(ub + R.x);
                        u1[q] = *reinterpret_cast<const double*>(ub + R.y);
                        u2[q] = *reinterpret_cast<const double*>(ub + R.z);
                    }
                }
                __builtin_amdgcn_sched_barrier(0);  // keep every load of the step ahead of the arithmetic
#pragma unroll
                for (int q = 0; q < kDivC; ++q) {
                    // every step computes (a skip would pull its loads down to it), padding masked
                    const double* G = geo[min(e + q, re - 1) - base];
                    const double du1 = u1[q] - u0[q];
                    const double du2 = u2[q] - u0[q];
                    const double num = fma(du2, G[1], du1 * G[0]);
                    const double qq = fma(du1, fma(du2, G[3], du1 * G[2]), du2 * (du2 * G[4]));
                    // |grad u| > 1e-300 (fastcw_lean.py:433) holds for every positive qq
                    const double term = -num * rsqrt_refined(qq);
                    acc += (e + q < re && qq > 0.0) ? term : 0.0;
                }
            }
            *o = acc;
        }
    }
    // rows without incidences (isolated vertices) still get their zero
    for (int pr = w; pr < npairs; pr += 4) {
        const int r = pr / ncs, col = (pr - r * ncs) * 64 + lane;
        if (rptr[r] == rptr[r + 1]) out[(int64_t)(r0 + r) * ldk + col] = 0.0;
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
            for (int q = 0; q < kC; ++q) ax += (e + q < e1) ? a[q] * xv[q] : 0.0;
        }
        v -= ax;
    }
    r[(int64_t)row * ldk + col] = (float)v;
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
    face_geometry_kernel<<<grid1(m, 256), 256, 0, st>>>(m, V, faces, area, gram, larea, g11, g12, g22);
}

void launch_assemble(hipStream_t st, int n, int nnz, const int* contrib_ptr, const int* contrib, const double* area,
                     const double* gram, const int* vf_ptr, const int* vf, double* lval, double* mass) {
    assemble_l_kernel<<<grid1(nnz, 256), 256, 0, st>>>(nnz, contrib_ptr, contrib, area, gram, lval);
    lumped_mass_kernel<<<grid1(n, 256), 256, 0, st>>>(n, vf_ptr, vf, area, mass);
}

void launch_fill(hipStream_t st, double* p, int64_t count, double value) {
    if (count <= 0) return;
    int64_t blocks = (count + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    fill_kernel<<<dim3((unsigned)blocks), 256, 0, st>>>(p, count, value);
}

void launch_calibration_copy(hipStream_t st, const void* src, void* dst, int64_t bytes) {
    const int64_t count = bytes / (int64_t)sizeof(double4);
    calibration_copy_kernel<<<dim3(4096), 256, 0, st>>>((const double4*)src, (double4*)dst, count);
}

void launch_delta_rhs(hipStream_t st, double* X, int n, int ldk, const int* src_pos, int k) {
    launch_fill(st, X, (int64_t)n * ldk, 0.0);
    delta_rhs_kernel<<<grid1(k, 256), 256, 0, st>>>(X, ldk, src_pos, k);
}

void launch_clear_rows(hipStream_t st, const int* nodes, int count, int ldk, const int* node_start,
                       const int* node_s, const int* node_b, const int64_t* node_uoff, double* Y, double* U) {
    if (count <= 0) return;
    clear_rows_kernel<<<dim3(count, ldk / 64), 256, 0, st>>>(nodes, ldk, node_start, node_s, node_b, node_uoff, Y, U);
}

void launch_divergence(hipStream_t st, const MeshDev& mesh, int ldk, const double* u, double* out) {
    const int ngroups = (mesh.n + kDivRows - 1) / kDivRows;
    const int run = (ngroups + kXcds - 1) / kXcds;  // row groups per XCD
    const uint64_t bytes = (uint64_t)mesh.n * (uint64_t)ldk * 8u;  // below 4 GiB (checked at creation)
    if (bytes < (1ull << 31))
        divergence_kernel<true><<<dim3(run * kXcds), 256, 0, st>>>(mesh.n, ldk, ngroups, run, mesh.op_vf_ptr, mesh.op_vf,
                                                                   mesh.op_faces, mesh.op_area, mesh.op_gram, u, out);
    else
        divergence_kernel<false><<<dim3(run * kXcds), 256, 0, st>>>(mesh.n, ldk, ngroups, run, mesh.op_vf_ptr, mesh.op_vf,
                                                                    mesh.op_faces, mesh.op_area, mesh.op_gram, u, out);
}

void launch_residual32(hipStream_t st, int n, int ldk, const int* rowptr, const int* colind, const double* lval,
                       const double* mass, double shift, const double* b, const double* x, float* r) {
    residual32_kernel<<<dim3((n + 3) / 4, ldk / 64), 256, 0, st>>>(n, ldk, rowptr, colind, lval, mass, shift, b, x, r);
}

void launch_to_float(hipStream_t st, const double* src, float* dst, int64_t count) {
    if (count <= 0) return;
    to_float_kernel<<<grid1(count, 256), 256, 0, st>>>(src, dst, count);
}

void launch_gather_rows(hipStream_t st, int n, int ldk, int k, const double* Dt, const double* phi, const int* src_pos,
                        const int* iperm, double* out) {
    gather_rows_kernel<<<dim3((n + 255) / 256, k), 256, 0, st>>>(n, ldk, k, Dt, phi, src_pos, iperm, out);
}


}  // namespace fcw
