// Level-scheduled supernodal sweeps as dense panel products on the FP64
// matrix cores of gfx950 (v_mfma_f64_16x16x4f64).
//
// Forward, node S with RHS rows X_S (s x k):              [Y_S ; U_S] = P_S F_S + [0 ; T_S]
//   F_S = X_S minus the children's update rows that land on S's own rows,
//   T_S = the sum of the children's update rows that land on S's boundary rows
//   (both gathered through per-row contributor maps while the tiles load).
// Backward, with the ancestors' solution rows Z_B (b x k):  Z_S = P_S^T [Y_S ; -Z_B]
// where P_S = [L11^{-1} ; L21 L11^{-1}] is the node's (s+b) x s column-major panel.
//
// Workgroup tile 64 x 64 (4 waves in 2 x 2, each wave 32 x 32 = 2 x 2 MFMA
// blocks), depth 32 per step staged through LDS, next step prefetched into
// registers while the current one runs on the matrix cores. The node's
// boundary-row indices (backward) or child-contributor maps (forward) are staged
// in LDS once per workgroup, so no tile load waits on an index load.
//
// Operand layout of v_mfma_f64_16x16x4f64 (one wave): A[i][k] in lane i + 16k, B[k][j] in lane
// j + 16k, and the 16 x 16 result D[i][j] in lane j + 16 (i mod 4), register i / 4 (checked on
// the device with a one-wave probe).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fcw_device.h"

namespace fcw {
namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kPad = 2;
constexpr int kLoadsA = kTileM * kTileK / kThreads;  // 4
constexpr int kLoadsB = kTileK * kTileN / kThreads;  // 8
constexpr int kIdxCap = 1024;                         // rows whose indices are staged in LDS
static_assert(kTileM == 64 && kTileN == 64 && kTileK == 16, "MFMA sweep tiling assumes 64 x 64 x 32");

enum class Sweep { Forward, Backward };

template <Sweep MODE>
__global__ __launch_bounds__(kThreads) void sweep_kernel(
    const LevelTile* __restrict__ tiles, int ldk, const int* __restrict__ node_start, const int* __restrict__ node_s,
    const int* __restrict__ node_b, const int64_t* __restrict__ node_boff, const int* __restrict__ bidx,
    const int64_t* __restrict__ node_poff, const int64_t* __restrict__ node_uoff, const int* __restrict__ cown,
    const int* __restrict__ ctail, const double* __restrict__ panel, const double* __restrict__ Bsrc,
    const double* Zsrc, double* Out, double* U) {
    const LevelTile t = tiles[blockIdx.x];
    const int s = node_s[t.node], b = node_b[t.node], f = s + b, start = node_start[t.node];
    const double* __restrict__ P = panel + node_poff[t.node];
    const int* __restrict__ B = (MODE == Sweep::Backward) ? bidx + node_boff[t.node] : nullptr;
    const int n0 = blockIdx.y * kTileN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave & 1) * 32, wn0 = (wave >> 1) * 32;
    const int M = (MODE == Sweep::Forward) ? f : s;   // output rows of this node
    const int K = (MODE == Sweep::Forward) ? s : f;   // depth

    __shared__ double As[kTileK][kTileM + kPad];
    __shared__ double Bs[kTileK][kTileN + kPad];
    __shared__ int sidx[2 * kIdxCap];
    // Stage per-row indices: backward -> boundary positions (b rows), forward -> two child
    // contributor rows for each of the node's s own rows.
    const bool idx_lds = (MODE == Sweep::Backward) ? (b <= kIdxCap) : (s <= kIdxCap);
    if (idx_lds) {
        if (MODE == Sweep::Backward) {
            for (int i = tid; i < b; i += kThreads) sidx[i] = B[i];
        } else {
            for (int i = tid; i < 2 * s; i += kThreads) sidx[i] = cown[2 * start + i];
        }
    }
    __syncthreads();

    double ra[kLoadsA], rb[kLoadsB];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int i = 0; i < kLoadsA; ++i) {
            const int e = tid + i * kThreads;
            double v = 0.0;
            if (MODE == Sweep::Forward) {
                const int r = e % kTileM, kk = e / kTileM;  // P rows contiguous
                const int row = t.m0 + r, col = k0 + kk;
                if (row < f && col < s) v = P[row + (int64_t)col * f];
            } else {
                const int kk = e % kTileK, r = e / kTileK;  // P^T: walk a panel column
                const int orow = t.m0 + r, l = k0 + kk;
                if (orow < s && l < f) v = P[l + (int64_t)orow * f];
            }
            ra[i] = v;
        }
#pragma unroll
        for (int i = 0; i < kLoadsB; ++i) {
            const int e = tid + i * kThreads;
            const int c = e % kTileN, kk = e / kTileN;
            const int l = k0 + kk;
            double v = 0.0;
            if (MODE == Sweep::Forward) {
                if (l < s) {
                    const int pos = start + l;
                    v = Bsrc[(int64_t)pos * ldk + n0 + c];
                    const int ca = idx_lds ? sidx[2 * l] : cown[2 * pos];
                    const int cb = idx_lds ? sidx[2 * l + 1] : cown[2 * pos + 1];
                    if (ca >= 0) v -= U[(int64_t)ca * ldk + n0 + c];
                    if (cb >= 0) v -= U[(int64_t)cb * ldk + n0 + c];
                }
            } else {
                if (l < s)
                    v = Bsrc[(int64_t)(start + l) * ldk + n0 + c];
                else if (l < f)
                    v = -Zsrc[(int64_t)(idx_lds ? sidx[l - s] : B[l - s]) * ldk + n0 + c];
            }
            rb[i] = v;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int i = 0; i < kLoadsA; ++i) {
            const int e = tid + i * kThreads;
            if (MODE == Sweep::Forward)
                As[e / kTileM][e % kTileM] = ra[i];
            else
                As[e % kTileK][e / kTileK] = ra[i];
        }
#pragma unroll
        for (int i = 0; i < kLoadsB; ++i) {
            const int e = tid + i * kThreads;
            Bs[e / kTileN][e % kTileN] = rb[i];
        }
    };

    dbl4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};

    const int lrow = lane & 15, lk = lane >> 4;
    if (K > 0) fetch(0);
    for (int k0 = 0; k0 < K; k0 += kTileK) {
        __syncthreads();
        stash();
        __syncthreads();
        if (k0 + kTileK < K) fetch(k0 + kTileK);
#pragma unroll
        for (int k4 = 0; k4 < kTileK; k4 += 4) {
            double a0 = As[k4 + lk][wm0 + lrow];
            double a1 = As[k4 + lk][wm0 + 16 + lrow];
            double b0 = Bs[k4 + lk][wn0 + lrow];
            double b1 = Bs[k4 + lk][wn0 + 16 + lrow];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
    }

    const int64_t uoff = (MODE == Sweep::Forward) ? node_uoff[t.node] : 0;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = n0 + wn0 + 16 * j + lrow;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                // v_mfma_f64_16x16x4f64 result: lane l, register r holds D[(l >> 4) + 4 r][l & 15]
                const int row = t.m0 + wm0 + 16 * i + lk + 4 * r;
                if (row >= M) continue;
                const double v = acc[i][j][r];
                if (MODE == Sweep::Forward) {
                    if (row < s) {
                        Out[(int64_t)(start + row) * ldk + col] = v;
                    } else {
                        const int64_t ur = uoff + (row - s);
                        const int ta = ctail[2 * ur], tb = ctail[2 * ur + 1];
                        double w = v;
                        if (ta >= 0) w += U[(int64_t)ta * ldk + col];
                        if (tb >= 0) w += U[(int64_t)tb * ldk + col];
                        U[ur * ldk + col] = w;
                    }
                } else {
                    Out[(int64_t)(start + row) * ldk + col] = v;
                }
            }
        }
    }
}

}  // namespace

void launch_fwd_gemm(hipStream_t st, const LevelTile* tiles, int ntiles, int ldk, const int* node_start,
                     const int* node_s, const int* node_b, const int64_t* node_poff, const int64_t* node_uoff,
                     const int* contrib_own, const int* contrib_tail, const double* panel, const double* X, double* Y,
                     double* U) {
    if (ntiles <= 0) return;
    sweep_kernel<Sweep::Forward><<<dim3(ntiles, ldk / kTileN), kThreads, 0, st>>>(
        tiles, ldk, node_start, node_s, node_b, nullptr, nullptr, node_poff, node_uoff, contrib_own, contrib_tail,
        panel, X, nullptr, Y, U);
}

void launch_bwd_gemm(hipStream_t st, const LevelTile* tiles, int ntiles, int ldk, const int* node_start,
                     const int* node_s, const int* node_b, const int64_t* node_boff, const int* bidx,
                     const int64_t* node_poff, const double* panel, const double* Y, double* Z) {
    if (ntiles <= 0) return;
    sweep_kernel<Sweep::Backward><<<dim3(ntiles, ldk / kTileN), kThreads, 0, st>>>(
        tiles, ldk, node_start, node_s, node_b, node_boff, bidx, node_poff, nullptr, nullptr, nullptr, panel, Y, Z, Z,
        nullptr);
}

}  // namespace fcw
