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

constexpr int kWave = 64;

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

// ---------------------------------------------------------------------------
// Operators (fastcw_lean.py:375-413, lean_geometry.py:44-63)
// ---------------------------------------------------------------------------

__global__ void face_geometry_kernel(int m, const double* __restrict__ V, const int* __restrict__ faces,
                                     double* __restrict__ area, double* __restrict__ gb,
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
    double* g = gb + 9 * (int64_t)f;
    for (int c = 0; c < 3; ++c) {
        g[c] = -b1[c] - b2[c];
        g[3 + c] = b1[c];
        g[6 + c] = b2[c];
    }
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
__global__ void assemble_l_kernel(int nnz, const int* __restrict__ cptr, const int* __restrict__ contrib,
                                  const double* __restrict__ area, const double* __restrict__ gb,
                                  double* __restrict__ lval) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    double s = 0.0;
    for (int c = cptr[e]; c < cptr[e + 1]; ++c) {
        const int code = contrib[c];
        const int f = code / 9, ab = code - 9 * f, a = ab / 3, b = ab - 3 * a;
        const double* g = gb + 9 * (int64_t)f;
        s += area[f] * (g[3 * a] * g[3 * b] + g[3 * a + 1] * g[3 * b + 1] + g[3 * a + 2] * g[3 * b + 2]);
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

__global__ void delta_rhs_kernel(double* __restrict__ X, int ldk, const int* __restrict__ src_pos, int k) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < k) X[(int64_t)src_pos[j] * ldk + j] = 1.0;
}

// ---------------------------------------------------------------------------
// Level-scheduled supernodal sweeps with partitioned-inverse panels
// ---------------------------------------------------------------------------

// Children's update rows are folded into the node's own RHS rows (subtract) and
// into its update rows (add). Children run one after another, so no atomics.
__global__ __launch_bounds__(256) void fwd_assemble_kernel(
    const int* __restrict__ nodes, int ldk, const int* __restrict__ node_start, const int* __restrict__ node_s,
    const int* __restrict__ node_b, const int64_t* __restrict__ node_boff, const int64_t* __restrict__ node_uoff,
    const int* __restrict__ child_ptr, const int* __restrict__ child_list, const int* __restrict__ cmap,
    double* __restrict__ X, double* __restrict__ U) {
    const int id = nodes[blockIdx.x];
    const int col = blockIdx.y * 64 + (threadIdx.x & 63);
    const int rg = threadIdx.x >> 6;
    const int s = node_s[id], b = node_b[id], start = node_start[id];
    double* Ut = U + node_uoff[id] * ldk;
    for (int r = rg; r < b; r += 4) Ut[(int64_t)r * ldk + col] = 0.0;
    __syncthreads();
    for (int ci = child_ptr[id]; ci < child_ptr[id + 1]; ++ci) {
        const int c = child_list[ci];
        const int bc = node_b[c];
        const int* cm = cmap + node_boff[c];
        const double* Uc = U + node_uoff[c] * ldk;
        for (int r = rg; r < bc; r += 4) {
            const int t = cm[r];
            const double v = Uc[(int64_t)r * ldk + col];
            if (t < s)
                X[(int64_t)(start + t) * ldk + col] -= v;
            else
                Ut[(int64_t)(t - s) * ldk + col] += v;
        }
        __syncthreads();
    }
}

// [Y_S; U_S] (+)= P_S X_S : P_S is (s+b) x s column-major, X_S the node's s RHS rows.
__global__ __launch_bounds__(256) void fwd_gemm_kernel(
    const LevelTile* __restrict__ tiles, int ldk, const int* __restrict__ node_start, const int* __restrict__ node_s,
    const int* __restrict__ node_b, const int64_t* __restrict__ node_poff, const int64_t* __restrict__ node_uoff,
    const double* __restrict__ panel, const double* __restrict__ X, double* __restrict__ Y, double* __restrict__ U,
    int accumulate) {
    const LevelTile t = tiles[blockIdx.x];
    const int s = node_s[t.node], b = node_b[t.node], f = s + b, start = node_start[t.node];
    const double* P = panel + node_poff[t.node];
    const int n0 = blockIdx.y * kTileN;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    __shared__ double As[kTileK][kTileM + 1];
    __shared__ double Bs[kTileK][kTileN];
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
    for (int k0 = 0; k0 < s; k0 += kTileK) {
#pragma unroll
        for (int e = tid; e < kTileM * kTileK; e += 256) {
            const int r = e % kTileM, kk = e / kTileM;
            const int row = t.m0 + r, col = k0 + kk;
            As[kk][r] = (row < f && col < s) ? P[row + (int64_t)col * f] : 0.0;
        }
#pragma unroll
        for (int e = tid; e < kTileK * kTileN; e += 256) {
            const int c = e % kTileN, kk = e / kTileN;
            const int row = k0 + kk;
            Bs[kk][c] = (row < s) ? X[(int64_t)(start + row) * ldk + n0 + c] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kTileK; ++kk) {
            double a[4], bb[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) bb[j] = Bs[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], bb[j], acc[i][j]);
        }
        __syncthreads();
    }
    double* Ut = U + node_uoff[t.node] * ldk;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int row = t.m0 + ty + 16 * i;
        if (row >= f) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = n0 + tx + 16 * j;
            if (row < s) {
                Y[(int64_t)(start + row) * ldk + col] = acc[i][j];
            } else {
                double* u = Ut + (int64_t)(row - s) * ldk + col;
                *u = accumulate ? (*u + acc[i][j]) : acc[i][j];
            }
        }
    }
}

// Z_S = P_S^T [Y_S; -Z_B] : B = the node's boundary rows (ancestors, already solved).
__global__ __launch_bounds__(256) void bwd_gemm_kernel(
    const LevelTile* __restrict__ tiles, int ldk, const int* __restrict__ node_start, const int* __restrict__ node_s,
    const int* __restrict__ node_b, const int64_t* __restrict__ node_boff, const int* __restrict__ bidx,
    const int64_t* __restrict__ node_poff, const double* __restrict__ panel, const double* __restrict__ Y,
    double* __restrict__ Z) {
    const LevelTile t = tiles[blockIdx.x];
    const int s = node_s[t.node], b = node_b[t.node], f = s + b, start = node_start[t.node];
    const double* P = panel + node_poff[t.node];
    const int* B = bidx + node_boff[t.node];
    const int n0 = blockIdx.y * kTileN;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    __shared__ double As[kTileK][kTileM + 1];
    __shared__ double Bs[kTileK][kTileN];
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
    for (int k0 = 0; k0 < f; k0 += kTileK) {
#pragma unroll
        for (int e = tid; e < kTileM * kTileK; e += 256) {
            const int kk = e % kTileK, r = e / kTileK;
            const int orow = t.m0 + r, l = k0 + kk;
            As[kk][r] = (orow < s && l < f) ? P[l + (int64_t)orow * f] : 0.0;
        }
#pragma unroll
        for (int e = tid; e < kTileK * kTileN; e += 256) {
            const int c = e % kTileN, kk = e / kTileN;
            const int l = k0 + kk;
            double v = 0.0;
            if (l < s)
                v = Y[(int64_t)(start + l) * ldk + n0 + c];
            else if (l < f)
                v = -Z[(int64_t)B[l - s] * ldk + n0 + c];
            Bs[kk][c] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kTileK; ++kk) {
            double a[4], bb[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) bb[j] = Bs[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], bb[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int row = t.m0 + ty + 16 * i;
        if (row >= s) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) Z[(int64_t)(start + row) * ldk + n0 + tx + 16 * j] = acc[i][j];
    }
}

// ---------------------------------------------------------------------------
// Fused gradient -> normalise -> divergence (fastcw_lean.py:416-442), gathered per vertex
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void divergence_kernel(int n, int ldk, const int* __restrict__ vf_ptr,
                                                         const int* __restrict__ vf, const int* __restrict__ faces,
                                                         const double* __restrict__ area,
                                                         const double* __restrict__ gb, const double* __restrict__ u,
                                                         double* __restrict__ out) {
    const int col = blockIdx.y * 64 + (threadIdx.x & 63);
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    double acc = 0.0;
    for (int e = vf_ptr[row]; e < vf_ptr[row + 1]; ++e) {
        const int code = vf[e];
        const int f = code / 3, a = code - 3 * f;
        const int i0 = faces[3 * f], i1 = faces[3 * f + 1], i2 = faces[3 * f + 2];
        const double u0 = u[(int64_t)i0 * ldk + col];
        const double du1 = u[(int64_t)i1 * ldk + col] - u0;
        const double du2 = u[(int64_t)i2 * ldk + col] - u0;
        const double* g = gb + 9 * (int64_t)f;
        const double gx = du1 * g[3] + du2 * g[6];
        const double gy = du1 * g[4] + du2 * g[7];
        const double gz = du1 * g[5] + du2 * g[8];
        const double nn = sqrt(gx * gx + gy * gy + gz * gz);
        if (!(nn > 1e-300)) continue;
        const double sc = -area[f] / nn;
        acc += sc * (gx * g[3 * a] + gy * g[3 * a + 1] + gz * g[3 * a + 2]);
    }
    out[(int64_t)row * ldk + col] = acc;
}

// Dt[j][p] = phi[p][j] - phi[src_j][j]   (source-major distance fields)
__global__ void shift_transpose_kernel(int n, int ldk, int k, const double* __restrict__ phi,
                                       const int* __restrict__ src_pos, double* __restrict__ Dt) {
    __shared__ double tile[32][33];
    __shared__ double base[32];
    const int p0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
    const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
    if (ty == 0) {
        const int j = j0 + tx;
        base[tx] = (j < k) ? phi[(int64_t)src_pos[j] * ldk + j] : 0.0;
    }
    __syncthreads();
    for (int yy = ty; yy < 32; yy += 8) {
        const int p = p0 + yy, j = j0 + tx;
        tile[yy][tx] = (p < n && j < k) ? phi[(int64_t)p * ldk + j] - base[tx] : 0.0;
    }
    __syncthreads();
    for (int yy = ty; yy < 32; yy += 8) {
        const int j = j0 + yy, p = p0 + tx;
        if (j < k && p < n) Dt[(int64_t)j * n + p] = tile[tx][yy];
    }
}

__global__ void gather_rows_kernel(int n, int k, const double* __restrict__ Dt, const int* __restrict__ iperm,
                                   double* __restrict__ out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (v < n && j < k) out[(int64_t)j * n + v] = Dt[(int64_t)j * n + iperm[v]];
}

// ---------------------------------------------------------------------------
// Per-source wiring metrics: one workgroup per source field
// (core_analysis.py:1558-1609, lean_geometry.py:117-357)
// ---------------------------------------------------------------------------

__device__ inline void face_cdf(double d0, double d1, double d2, double A, double r, double& a_out, double& da_out) {
    double a = d0, b = d1, c = d2, tmp;
    if (a > b) { tmp = a; a = b; b = tmp; }
    if (b > c) { tmp = b; b = c; c = tmp; }
    if (a > b) { tmp = a; a = b; b = tmp; }
    if (r >= c) { a_out = A; da_out = 0.0; return; }
    if (r <= a) { a_out = 0.0; da_out = 0.0; return; }
    if (r < b) {
        const double den = (b - a) * (c - a);
        const double x = r - a;
        a_out = A * x * x / den;
        da_out = 2.0 * A * x / den;
        return;
    }
    const double den = (c - a) * (c - b);
    const double y = c - r;
    a_out = A * (1.0 - y * y / den);
    da_out = 2.0 * A * y / den;
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

struct BlockRed {
    double d[kMetricThreads / kWave][4];
    int i[kMetricThreads / kWave][4];
};

template <int NT>
__device__ void block_reduce(BlockRed& sm, double (&v)[4], int (&iv)[4], int mode_max_slot, int mode_min_slot) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    constexpr int NW = NT / kWave;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double x = v[q];
        if (q == mode_max_slot)
            x = wave_max(x);
        else if (q == mode_min_slot)
            x = wave_min(x);
        else
            x = wave_sum(x);
        v[q] = x;
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
            if (q == mode_max_slot)
                x = fmax(x, y);
            else if (q == mode_min_slot)
                x = fmin(x, y);
            else
                x += y;
            xi += sm.i[ww][q];
        }
        v[q] = x;
        iv[q] = xi;
    }
    __syncthreads();
}

__global__ __launch_bounds__(kMetricThreads) void metrics_kernel(MeshDev mesh, MetricsDev md,
                                                                  const double* __restrict__ Dt,
                                                                  const int* __restrict__ src_pos, MetricsOut out,
                                                                  int health_mode) {
    constexpr int NT = kMetricThreads;
    const int j = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int n = mesh.n, m = mesh.m;
    const double* __restrict__ d = Dt + (int64_t)j * n;
    const int S = md.S, q = md.q, NE = (S > 0 ? (S - 1) * q : 0);
    double* scal = out.scal + (int64_t)j * kSlotCount;
    double* radii = out.radii + (int64_t)j * S;
    double* perim = out.perim + (int64_t)j * S;
    double* er = out.extra_r + (int64_t)j * NE;
    double* ea = out.extra_a + (int64_t)j * NE;

    __shared__ BlockRed red;
    __shared__ double Hmax[kMaxBins + 1];
    __shared__ double Hmin[kMaxBins + 1];
    __shared__ int lo_k[kMaxScales], hi_k[kMaxScales], cnt[kMaxScales], act[kMaxScales];
    __shared__ double hi_r[kMaxScales], r_out[kMaxScales], p_out[kMaxScales];
    __shared__ int ev_out[kMaxScales];
    __shared__ double rho[kMaxExtras], racc[kMaxExtras];
    __shared__ int rk[kMaxExtras], ridx[kMaxExtras];
    __shared__ int s_nv;
    __shared__ double s_wb, s_total;
    __shared__ int s_nbins;
    __shared__ double pair_r[kMaxScales + kMaxExtras], pair_a[kMaxScales + kMaxExtras];

    // ---- phase 0: health screen, MSD, distance to boundary (one vertex pass) ----
    const int sp = src_pos[j];
    const double xs = mesh.V[3 * sp], ys = mesh.V[3 * sp + 1], zs = mesh.V[3 * sp + 2];
    double v4[4] = {0.0, 0.0, 0.0, 0.0};  // sum_d, sum_dw, sum_w, dglob(max)
    int i4[4] = {0, 0, 0, 0};             // neg, viol, valid, nonfinite
    const double h = md.health_h;
    for (int p = tid; p < n; p += NT) {
        const double dv = d[p];
        if (!isfinite(dv)) {
            i4[3] += 1;
            continue;
        }
        const double dx = mesh.V[3 * p] - xs, dy = mesh.V[3 * p + 1] - ys, dz = mesh.V[3 * p + 2] - zs;
        const double chord = sqrt(dx * dx + dy * dy + dz * dz);
        if (dv < -h) i4[0] += 1;
        if (dv < 0.9 * chord - h) i4[1] += 1;
        if (dv > md.eps) {
            const double w = md.vw[p];
            i4[2] += 1;
            v4[0] += dv;
            v4[1] += dv * w;
            v4[2] += w;
        }
        v4[3] = fmax(v4[3], dv);
    }
    block_reduce<NT>(red, v4, i4, 3, -1);
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
    block_reduce<NT>(red, b4, bi4, -1, 0);
    const double neg_frac = i4[3] ? NAN : (double)i4[0] / n;
    const double viol_frac = i4[3] ? NAN : (double)i4[1] / n;
    const bool bad = (i4[3] > 0) || neg_frac > 1e-3 || viol_frac > 1e-3;
    double dist_b;
    if (md.nb > 0)
        dist_b = bi4[0] > 0 ? b4[0] : NAN;
    else
        dist_b = INFINITY;
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

    // ---- phase 1: dmax / dmin area histograms -> brackets -------------------
    if (tid == 0) {
        double wb = md.width;
        double nbd = floor(dglob / wb) + 2.0;
        if (nbd > (double)kMaxBins) {
            const double fac = ceil(nbd / (double)(kMaxBins - 2));
            wb = md.width * fac;
        }
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
    const int chunk = (m + NT - 1) / NT;
    const int f_begin = tid * chunk, f_end = min(m, f_begin + chunk);
    {
        int cur_hi = -1, cur_lo = -1;
        double acc_hi = 0.0, acc_lo = 0.0;
        for (int f = f_begin; f < f_end; ++f) {
            const double d0 = d[mesh.faces[3 * f]], d1 = d[mesh.faces[3 * f + 1]], d2 = d[mesh.faces[3 * f + 2]];
            if (!(isfinite(d0) && isfinite(d1) && isfinite(d2))) continue;
            const double lo = fmin(d0, fmin(d1, d2)), hi = fmax(d0, fmax(d1, d2));
            const double A = mesh.larea[f];
            const int bh = bin_of(hi, wb, nbins), bl = bin_of(lo, wb, nbins);
            if (bh != cur_hi) {
                if (cur_hi >= 0) atomicAdd(&Hmax[cur_hi], acc_hi);
                cur_hi = bh;
                acc_hi = 0.0;
            }
            acc_hi += A;
            if (bl != cur_lo) {
                if (cur_lo >= 0) atomicAdd(&Hmin[cur_lo], acc_lo);
                cur_lo = bl;
                acc_lo = 0.0;
            }
            acc_lo += A;
        }
        if (cur_hi >= 0) atomicAdd(&Hmax[cur_hi], acc_hi);
        if (cur_lo >= 0) atomicAdd(&Hmin[cur_lo], acc_lo);
    }
    __syncthreads();
    // exclusive prefix sums in place: H[k] = sum_{b<k} hist[b], k = 0..nbins (wave 0 / wave 1)
    if (wid < 2) {
        double* H = (wid == 0) ? Hmax : Hmin;
        const int per = (nbins + kWave - 1) / kWave;
        const int b0 = lane * per, b1 = min(nbins, b0 + per);
        double loc = 0.0;
        for (int b = b0; b < b1; ++b) loc += H[b];
        // inclusive scan of loc across the wave
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
        int a = 0;  // 0 = leading inactive (r = 0 or NaN handled below), 1 = active, 2 = trailing inactive
        double r0v = NAN;
        if (T != T) {
            a = 2;
            r0v = NAN;
        } else if (T <= 0.0) {
            a = 0;
            r0v = 0.0;
        } else if (T > total * (1.0 + 1e-12)) {
            a = 2;
            r0v = NAN;
        } else {
            a = 1;
        }
        act[s] = a;
        r_out[s] = r0v;
        p_out[s] = NAN;
        ev_out[s] = 0;
        if (a == 1) {
            int lo = 0, hi = nbins;  // first k with Hmax[k] >= T
            if (Hmax[nbins] < T) {
                lo = nbins;
            } else {
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (Hmax[mid] >= T) hi = mid; else lo = mid + 1;
                }
            }
            const int khi = lo;
            int l2 = 0, h2 = khi;  // last k <= khi with Hmin[k] < T
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
        cnt[s] = 0;
    }
    if (tid == 0) s_total = total;
    __syncthreads();

    // ---- phase 2: band lists per scale --------------------------------------
    int* band = md.band + (int64_t)j * S * md.band_cap;
    for (int f = f_begin; f < f_end; ++f) {
        const double d0 = d[mesh.faces[3 * f]], d1 = d[mesh.faces[3 * f + 1]], d2 = d[mesh.faces[3 * f + 2]];
        if (!(isfinite(d0) && isfinite(d1) && isfinite(d2))) continue;
        const double lo = fmin(d0, fmin(d1, d2)), hi = fmax(d0, fmax(d1, d2));
        const int bh = bin_of(hi, wb, nbins);
        int a = 0, bnd = S;  // first s with hi_r[s] > lo
        while (a < bnd) {
            const int mid = (a + bnd) >> 1;
            if (hi_r[mid] > lo) bnd = mid; else a = mid + 1;
        }
        for (int s = a; s < S && lo_k[s] <= bh; ++s) {
            const int slot = atomicAdd(&cnt[s], 1);
            if (slot < md.band_cap) band[(int64_t)s * md.band_cap + slot] = f;
        }
    }
    __syncthreads();

    // ---- phase 3: safeguarded Newton per scale (one wave per scale) ---------
    constexpr int NW = NT / kWave;
    for (int s = wid; s < S; s += NW) {
        if (act[s] != 1) continue;
        const double T = md.targets[s];
        const int count = cnt[s];
        const bool full_scan = count > md.band_cap;
        const int* list = band + (int64_t)s * md.band_cap;
        const int klo = lo_k[s];
        const double Ain = Hmax[klo];
        const double bl_r = (double)klo * wb, bh_r = hi_r[s];
        double r_lo = bl_r, r_hi = bh_r;
        double r = 0.5 * (r_lo + r_hi);
        int n_eval = 0;
        double result = r;
        auto eval = [&](double rr, double& A_out, double& dA_out, double& P_out, bool want_p) {
            double A = 0.0, dA = 0.0, P = 0.0;
            const int lim = full_scan ? m : count;
            for (int i = lane; i < lim; i += kWave) {
                int f;
                if (full_scan) {
                    f = i;
                    const double e0 = d[mesh.faces[3 * f]], e1 = d[mesh.faces[3 * f + 1]], e2 = d[mesh.faces[3 * f + 2]];
                    if (!(isfinite(e0) && isfinite(e1) && isfinite(e2))) continue;
                    const double lo = fmin(e0, fmin(e1, e2)), hi = fmax(e0, fmax(e1, e2));
                    if (!(bin_of(hi, wb, nbins) >= klo && lo < bh_r)) continue;
                } else {
                    f = list[i];
                }
                const double d0 = d[mesh.faces[3 * f]], d1 = d[mesh.faces[3 * f + 1]], d2 = d[mesh.faces[3 * f + 2]];
                double a, da;
                face_cdf(d0, d1, d2, mesh.larea[f], rr, a, da);
                A += a;
                dA += da;
                if (want_p && da > 0.0) P += da * grad_norm(d0, d1, d2, mesh.g11[f], mesh.g12[f], mesh.g22[f]);
            }
            A_out = Ain + wave_sum(A);
            dA_out = wave_sum(dA);
            P_out = want_p ? wave_sum(P) : 0.0;
        };
        bool done = false;
        for (int it = 0; it < md.max_iter && !done; ++it) {
            double A, dA, Pd;
            eval(r, A, dA, Pd, false);
            ++n_eval;
            const double err = A - T;
            if (fabs(err) <= md.rtol * T) {
                result = r;
                done = true;
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
                done = true;
                break;
            }
            r = rn;
            result = r;
        }
        double A, dA, Pv;
        if (result > 0.0) {
            eval(result, A, dA, Pv, true);
        } else {
            Pv = 0.0;
        }
        if (lane == 0) {
            r_out[s] = result;
            p_out[s] = Pv;
            ev_out[s] = n_eval + 1;
        }
    }
    __syncthreads();

    // ---- phase 4: supplementary samples between consecutive radii -----------
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
                rk[nv] = (int)(rr / wb);
                ++nv;
            }
        }
        s_nv = nv;
    }
    __syncthreads();
    const int nv = s_nv;
    if (tid < nv) {
        const int kr = rk[tid];
        racc[tid] = (kr >= nbins) ? s_total : Hmax[kr];
    }
    __syncthreads();
    if (nv > 0) {
        int cur = -1;
        double accv = 0.0;
        for (int f = f_begin; f < f_end; ++f) {
            const double d0 = d[mesh.faces[3 * f]], d1 = d[mesh.faces[3 * f + 1]], d2 = d[mesh.faces[3 * f + 2]];
            if (!(isfinite(d0) && isfinite(d1) && isfinite(d2))) continue;
            const double lo = fmin(d0, fmin(d1, d2)), hi = fmax(d0, fmax(d1, d2));
            const int bh = bin_of(hi, wb, nbins);
            int a = 0, bnd = nv;  // first v with rho > lo
            while (a < bnd) {
                const int mid = (a + bnd) >> 1;
                if (rho[ridx[mid]] > lo) bnd = mid; else a = mid + 1;
            }
            for (int v = a; v < nv && rk[v] <= bh; ++v) {
                if (rk[v] >= nbins) break;
                double av, dav;
                face_cdf(d0, d1, d2, mesh.larea[f], rho[ridx[v]], av, dav);
                if (v != cur) {
                    if (cur >= 0) atomicAdd(&racc[cur], accv);
                    cur = v;
                    accv = 0.0;
                }
                accv += av;
            }
        }
        if (cur >= 0) atomicAdd(&racc[cur], accv);
    }
    __syncthreads();

    // ---- phase 5: outputs and cleaned (radius, area) samples ----------------
    for (int s = tid; s < S; s += NT) {
        radii[s] = r_out[s];
        perim[s] = (r_out[s] == r_out[s]) ? p_out[s] : NAN;
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
        // Pairs: solved (r_s, T_s) for finite r_s, then finite extras passing the boundary cap.
        double* pr = pair_r;
        double* pa = pair_a;
        int np = 0;
        for (int s = 0; s < S; ++s)
            if (isfinite(r_out[s])) {
                pr[np] = r_out[s];
                pa[np] = md.targets[s];
                ++np;
            }
        const double db32 = (double)(float)dist_b;
        for (int v = 0; v < nv; ++v) {
            const double rr = rho[ridx[v]], aa = racc[v];
            if (!(isfinite(rr) && isfinite(aa))) continue;
            if (md.cap >= 0.0) {
                double loss;
                if (!isfinite(rr) || rr <= 0.0) {
                    loss = NAN;
                } else if (!isfinite(db32) || db32 >= rr) {
                    loss = 0.0;
                } else {
                    double x = fmin(1.0, fmax(0.0, db32 / rr));
                    loss = (acos(x) - x * sqrt(fmax(0.0, 1.0 - x * x))) / M_PI;
                }
                if (loss > md.cap) continue;
            }
            pr[np] = rr;
            pa[np] = aa;
            ++np;
        }
        // stable insertion sort by radius
        for (int a = 1; a < np; ++a) {
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
        for (int s = 0; s < S; ++s) ovf |= (act[s] == 1 && cnt[s] > md.band_cap);
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

inline dim3 grid1(int64_t n, int bs) { return dim3((unsigned)((n + bs - 1) / bs)); }

}  // namespace

void launch_face_geometry(hipStream_t st, int m, const double* V, const int* faces, double* area, double* gb,
                          double* larea, double* g11, double* g12, double* g22) {
    face_geometry_kernel<<<grid1(m, 256), 256, 0, st>>>(m, V, faces, area, gb, larea, g11, g12, g22);
}

void launch_assemble(hipStream_t st, int n, int nnz, const int* contrib_ptr, const int* contrib, const double* area,
                     const double* gb, const int* vf_ptr, const int* vf, double* lval, double* mass) {
    assemble_l_kernel<<<grid1(nnz, 256), 256, 0, st>>>(nnz, contrib_ptr, contrib, area, gb, lval);
    lumped_mass_kernel<<<grid1(n, 256), 256, 0, st>>>(n, vf_ptr, vf, area, mass);
}

void launch_fill(hipStream_t st, double* p, int64_t count, double value) {
    if (count <= 0) return;
    int64_t blocks = (count + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    fill_kernel<<<dim3((unsigned)blocks), 256, 0, st>>>(p, count, value);
}

void launch_delta_rhs(hipStream_t st, double* X, int n, int ldk, const int* src_pos, int k) {
    launch_fill(st, X, (int64_t)n * ldk, 0.0);
    delta_rhs_kernel<<<grid1(k, 256), 256, 0, st>>>(X, ldk, src_pos, k);
}

void launch_fwd_assemble(hipStream_t st, const int* nodes, int count, int ldk, const int* node_start,
                         const int* node_s, const int* node_b, const int64_t* node_boff, const int64_t* node_uoff,
                         const int* child_ptr, const int* child_list, const int* cmap, double* X, double* U) {
    if (count <= 0) return;
    fwd_assemble_kernel<<<dim3(count, ldk / 64), 256, 0, st>>>(nodes, ldk, node_start, node_s, node_b, node_boff,
                                                                node_uoff, child_ptr, child_list, cmap, X, U);
}

void launch_fwd_gemm(hipStream_t st, const LevelTile* tiles, int ntiles, int ldk, const int* node_start,
                     const int* node_s, const int* node_b, const int64_t* node_poff, const int64_t* node_uoff,
                     const double* panel, const double* X, double* Y, double* U, bool accumulate) {
    if (ntiles <= 0) return;
    fwd_gemm_kernel<<<dim3(ntiles, ldk / kTileN), 256, 0, st>>>(tiles, ldk, node_start, node_s, node_b, node_poff,
                                                                 node_uoff, panel, X, Y, U, accumulate ? 1 : 0);
}

void launch_bwd_gemm(hipStream_t st, const LevelTile* tiles, int ntiles, int ldk, const int* node_start,
                     const int* node_s, const int* node_b, const int64_t* node_boff, const int* bidx,
                     const int64_t* node_poff, const double* panel, const double* Y, double* Z) {
    if (ntiles <= 0) return;
    bwd_gemm_kernel<<<dim3(ntiles, ldk / kTileN), 256, 0, st>>>(tiles, ldk, node_start, node_s, node_b, node_boff,
                                                                 bidx, node_poff, panel, Y, Z);
}

void launch_divergence(hipStream_t st, const MeshDev& mesh, int ldk, const double* u, double* out) {
    divergence_kernel<<<dim3((mesh.n + 3) / 4, ldk / 64), 256, 0, st>>>(mesh.n, ldk, mesh.vf_ptr, mesh.vf,
                                                                          mesh.faces, mesh.area, mesh.gb, u, out);
}

void launch_shift_transpose(hipStream_t st, int n, int ldk, int k, const double* phi, const int* src_pos,
                            double* Dt) {
    shift_transpose_kernel<<<dim3((n + 31) / 32, (k + 31) / 32), dim3(32, 8), 0, st>>>(n, ldk, k, phi, src_pos, Dt);
}

void launch_gather_rows(hipStream_t st, int n, int k, const double* Dt, const int* iperm, double* out) {
    gather_rows_kernel<<<dim3((n + 255) / 256, k), 256, 0, st>>>(n, k, Dt, iperm, out);
}

void launch_metrics(hipStream_t st, int k, const MeshDev& mesh, const MetricsDev& md, const double* Dt,
                    const int* src_pos, MetricsOut out, bool health_screen) {
    if (k <= 0) return;
    metrics_kernel<<<dim3(k), kMetricThreads, 0, st>>>(mesh, md, Dt, src_pos, out, health_screen ? 1 : 0);
}

}  // namespace fcw
