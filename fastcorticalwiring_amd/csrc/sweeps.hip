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
// Workgroup tile 64 x 64, 32 x 128 or (backward) 16 x 256, chosen per tree level (4 waves, each
// holding four 16 x 16 MFMA blocks), depth 16 per step staged through LDS, next step prefetched
// into registers while the current one runs on the matrix cores. The node's boundary-row indices
// (backward) or child-contributor maps (forward) are staged in LDS once per workgroup (when they
// fit: kSweepIdxCap), so no tile load waits on an index load.
//
// Operand layout of v_mfma_f64_16x16x4f64 (one wave): A[i][k] in lane i + 16k, B[k][j] in lane
// j + 16k, and the 16 x 16 result D[i][j] in lane j + 16 (i mod 4), register i / 4 (checked on
// the device with a one-wave probe).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "fcw_device.h"

namespace fcw {
namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));
typedef short bf4 __attribute__((ext_vector_type(4)));  // four bf16 values

// Split-BF16 form of an FP32 value: hi = x rounded to bf16 (nearest even), lo = the remainder
// x - hi rounded to bf16; hi + lo carries 16 significant bits. Packed as (lo << 16) | hi.
__device__ __forceinline__ unsigned bf16_rne(float x) {
    const unsigned u = __float_as_uint(x);
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ unsigned split_bf16(float x) {
    const unsigned hi = bf16_rne(x);
    const float rest = x - __uint_as_float(hi << 16);
    return (bf16_rne(rest) << 16) | hi;
}

// One 16 x 16 x 4 matrix-core step per value type, and where it leaves row r of lane group q
// (checked on the device with one-wave probes): FP64 D rows q + 4 r, FP32 D rows 4 q + r.
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef dbl4 acc_t;
    static __device__ acc_t step(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static __device__ int row(int q, int r) { return q + 4 * r; }
};
template <>
struct Mfma<float> {
    typedef flt4 acc_t;
    static __device__ acc_t step(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __device__ int row(int q, int r) { return 4 * q + r; }
};

constexpr int kThreads = 256;
constexpr int kPad = 2;
constexpr int kIdxCap = kSweepIdxCap;
static_assert(kTileK == 16, "MFMA sweep depth steps are 16 deep");

// Workgroup tile shapes: TM rows x TN = 4096 / TM columns, four waves, each wave holding WBM x WBN
// result blocks of 16 x 16 (always four blocks, 16 accumulator values per lane). TM = 64: waves in
// 2 x 2, each 32 x 32; TM = 32: waves in 1 x 4, each 32 x 32; TM = 16: waves in 1 x 4, each 16 x 64.
// Every shape runs 16 matrix-core steps per wave per depth step; the shorter tiles waste fewer rows
// on nodes with few rows and stage more right-hand-side values per step.
template <int TM>
struct Shape {
    static_assert(TM == 64 || TM == 32 || TM == 16, "row tiles of 64, 32 or 16");
    static constexpr int TN = 4096 / TM;
    static constexpr int WGM = TM == 64 ? 2 : 1;      // waves along the rows
    static constexpr int WBM = TM / (16 * WGM);       // result blocks per wave along the rows
    static constexpr int WBN = 4 / WBM;               // ... along the columns
    static constexpr int LA = TM * kTileK / kThreads;  // panel values each thread stages per depth step
    static constexpr int LB = TN * kTileK / kThreads;  // right-hand-side values each thread stages
    // The source-major copy of the backward epilogue goes through LDS in two halves of HR rows x
    // HC columns (TM = 64: row halves, else column halves).
    static constexpr int HR = TM == 64 ? 32 : TM;
    static constexpr int HC = 2048 / HR;
};

enum class Sweep { Forward, Backward };

// T: value type of panels, right-hand sides and solution (double, or float with FP32
// accumulation for the refined solve). KIDS = false: every node of the launch is a leaf (no
// child update rows to gather). Backward, acc_mode 1 / 2 also stores / adds each solution value
// into the FP64 array acc64 (the iterate of the refined solve); with 4 added the solution itself is
// not stored (the refined solve's leaf level: no node reads leaf rows of Z, and the iterate is
// acc64); with Dt set the (FP64) solution value is also written source-major,
// Dt[col * ldd + position], for the metric kernel.
//
// Grid: one workgroup per (row tile, TN-column slice), laid out in 1-D so that each XCD
// (workgroups go to the 8 XCDs round robin) walks its own contiguous run of tiles, the column
// slices of a tile one after another: the slices read the same panel rows, and the row tiles of
// one node the same right-hand-side rows, which the first of them brings into that XCD's L2 for
// the others. Workgroups past the last tile exit.
//
// Split depth (SPLIT, FP64 and TM = 64 only) for the narrow upper levels, whose few workgroups each
// walk a long depth: SPLIT = 1 gives each (tile, slice) nsplit workgroups, part p multiplying depth
// steps [p kc, (p + 1) kc) and storing its 64 x 64 partial product to `part`; SPLIT = 2 (the next
// launch, one workgroup per (tile, slice)) adds the nsplit partials in part order, then runs the
// ordinary epilogue. The order is fixed, so the result does not depend on scheduling; it differs
// from the unsplit product only in rounding.
constexpr int kXcds = 8;

// PT = uint32_t (with T = float): split-BF16 panels, each entry packed hi | lo << 16; the right-hand
// side is split the same way while it is staged, and each 16 x 16 block of a depth step takes three
// v_mfma_f32_16x16x16bf16_1k products (hi*hi + hi*lo + lo*hi, FP32 accumulation; the lo*lo term,
// 2^-16 of the product, is dropped). Operand layout (checked on the device with a one-wave probe):
// A[i][4q + v] in lane i + 16q, element v; B[4q + v][j] in lane j + 16q; D rows 4q + r as for FP32.
constexpr int kBfPad = kTileK + 4;  // bf16 LDS row: the 16 depth values of one tile row / column

template <typename T, int TM, Sweep MODE, bool IDX_LDS, bool KIDS, bool DT = false, int SPLIT = 0, typename PT = T>
__global__ __launch_bounds__(kThreads) void sweep_kernel(
    int ntiles, const LevelTile* __restrict__ tiles, int ldk, const int* __restrict__ node_start, const int* __restrict__ node_s,
    const int* __restrict__ node_b, const int64_t* __restrict__ node_boff, const int* __restrict__ bidx,
    const int64_t* __restrict__ node_poff, const int64_t* __restrict__ node_uoff, const int* __restrict__ cown,
    const int* __restrict__ ctail, const PT* __restrict__ panel, const T* __restrict__ Bsrc, const T* Zsrc, T* Out,
    T* U, double* acc64, int acc_mode, double* __restrict__ Dt, int ldd, int zrow, int nsplit, double* __restrict__ part) {
    static_assert(SPLIT == 0 || (sizeof(T) == sizeof(double) && TM == 64), "split depth is FP64 with 64-row tiles only");
    constexpr bool BF = std::is_same<PT, uint32_t>::value;
    static_assert(!BF || std::is_same<T, float>::value, "split-BF16 panels go with FP32 right-hand sides");
    typedef Shape<TM> Sh;
    constexpr int TN = Sh::TN, WBM = Sh::WBM, WBN = Sh::WBN, LA = Sh::LA, LB = Sh::LB;
    const int ncol = ldk / TN;
    const int xcd = blockIdx.x % kXcds;
    const int ps = SPLIT == 1 ? (int)(blockIdx.x / kXcds) % nsplit : 0;  // this workgroup's depth part
    const int j = SPLIT == 1 ? (int)(blockIdx.x / kXcds) / nsplit : (int)(blockIdx.x / kXcds);
    const int run = (ntiles + kXcds - 1) / kXcds;  // tiles per XCD
    const int tile = xcd * run + j / ncol;
    if (tile >= ntiles) return;
    const LevelTile t = tiles[tile];
    const int s = node_s[t.node], b = node_b[t.node], f = s + b, start = node_start[t.node];
    const PT* __restrict__ P = panel + node_poff[t.node];
    const int* __restrict__ B = (MODE == Sweep::Backward) ? bidx + node_boff[t.node] : nullptr;
    const int n0 = (j - (j / ncol) * ncol) * TN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave % Sh::WGM) * 16 * WBM, wn0 = (wave / Sh::WGM) * 16 * WBN;
    const int M = (MODE == Sweep::Forward) ? f : s;   // output rows of this node
    const int K = (MODE == Sweep::Forward) ? s : f;   // depth
    // depth range of this workgroup (all of it unless split)
    int kbeg = 0, kend = K;
    if (SPLIT == 1) {
        const int kc = ((K + nsplit - 1) / nsplit + kTileK - 1) / kTileK * kTileK;
        kbeg = min(K, ps * kc);
        kend = min(K, kbeg + kc);
    }
    // The panel's top s x s block is L11^{-1}, lower triangular with its upper part stored as
    // zeros: backward output row i meets nonzeros only at depth k >= i, forward own row i only at
    // k <= i. Depth steps that would multiply nothing but those zeros are skipped (each skipped
    // product is a zero, so the sums are unchanged).
    if (MODE == Sweep::Backward)
        kbeg = max(kbeg, t.m0 / kTileK * kTileK);
    else if (t.m0 + TM <= s)
        kend = min(kend, t.m0 + TM);
    // 16 accumulator values per thread for each (tile, slice, part)
    double* const pslot =
        SPLIT ? part + (((int64_t)tile * ncol + (j - (j / ncol) * ncol)) * nsplit) * (kThreads * 16) + tid * 16 : nullptr;

    // A stage (depth x rows) and B stage (depth x columns); the epilogue's transpose after the loop
    constexpr int kLdA = TM + kPad, kLdB = TN + kPad;
    __shared__ T sbuf[kTileK * (kLdA + kLdB)];
    T* const As = sbuf;
    T* const Bs = sbuf + kTileK * kLdA;
    __shared__ int sidx[2 * kIdxCap];
    // split-BF16 stages, depth contiguous: A hi / lo (tile row, depth), B hi / lo (column, depth)
    __shared__ __align__(16) unsigned short bfa[BF ? 2 : 1][BF ? TM : 1][BF ? kBfPad : 1];
    __shared__ __align__(16) unsigned short bfb[BF ? 2 : 1][BF ? TN : 1][BF ? kBfPad : 1];
    // Stage per-row indices: backward -> the source row of every depth index (own rows as positions
    // in Y, boundary rows as positions in Z), forward -> two child contributor rows for each of the
    // node's s own rows.
    if (IDX_LDS && SPLIT != 2) {
        if (MODE == Sweep::Backward) {
            // The boundary-map loads are issued together from clamped indices (only the uniform
            // per-workgroup bounds are branched on), then written.
            constexpr int R = 2 * kIdxCap / kThreads;
            if (b > 0) {
                int q[R];
#pragma unroll
                for (int u = 0; u < R; ++u) q[u] = B[min(max(tid + u * kThreads - s, 0), b - 1)];
#pragma unroll
                for (int u = 0; u < R; ++u) {
                    const int i = tid + u * kThreads;
                    if (i < f) sidx[i] = i < s ? start + i : q[u];
                }
            } else {
                for (int i = tid; i < f; i += kThreads) sidx[i] = start + i;
            }
        } else if (KIDS) {
            for (int i = tid; i < 2 * s; i += kThreads) sidx[i] = cown[2 * start + i];
        }
    }
    __syncthreads();

    // Tile loads. Panel (A) element e of a thread: tile row / depth from e; right-hand-side (B)
    // element i: column c (the same for all of a thread's elements, since 256 is a multiple of TN)
    // and depth row l = k0 + wrow + (256 / TN) i, where wrow = tid / TN is the same in every lane of
    // a wave (TN >= 64): each B element is one row of the right-hand side, whose address is
    // wave-uniform (computed on the scalar unit), plus this lane's column.
    //
    // Every load is unconditional and in bounds (indices clamped). Depth steps past the node's
    // depth are made zero on the panel side only: a panel value at depth k >= K is staged as 0, so
    // the right-hand-side values it meets (a clamped, finite row) add nothing. The backward sweep's
    // boundary rows enter negated ([Y_S ; -Z_B]); the sign is put on the panel values at depths
    // s..f-1 instead of on the right-hand side (the products are the same). fetch() only issues the
    // loads of the next step; stash() applies masks and signs after the current step's MFMAs.
    const int wrow = __builtin_amdgcn_readfirstlane(tid / TN);
    const int c = n0 + (tid & (TN - 1));
    PT ra[LA];
    T rb[LB], ru[KIDS && MODE == Sweep::Forward ? LB : 1], rv[KIDS && MODE == Sweep::Forward ? LB : 1];
    int kA = 0;  // backward: this thread's panel depth (the same for its LA elements) of the staged step
    unsigned fwd_keep = 0;  // forward: panel elements inside the node's columns
    auto row_ptr = [&](const T* base, int row) {  // wave-uniform row base
        return base + (int64_t)__builtin_amdgcn_readfirstlane(row) * ldk;
    };
    auto fetch = [&](int k0) {
        fwd_keep = 0;
#pragma unroll
        for (int i = 0; i < LA; ++i) {
            const int e = tid + i * kThreads;
            int prow, pcol;  // element (prow, pcol) of the (s+b) x s panel
            if (MODE == Sweep::Forward) {
                prow = t.m0 + e % TM;  // P rows contiguous
                pcol = k0 + e / TM;
            } else {
                prow = k0 + e % kTileK;  // P^T: walk a panel column
                pcol = t.m0 + e / kTileK;
            }
            if (MODE == Sweep::Forward) fwd_keep |= pcol < s ? (1u << i) : 0u;
            ra[i] = P[min(prow, f - 1) + (int64_t)min(pcol, s - 1) * f];
        }
        if (MODE == Sweep::Backward) kA = k0 + (tid & (kTileK - 1));
#pragma unroll
        for (int i = 0; i < LB; ++i) {
            const int l = k0 + wrow + (kThreads / TN) * i;  // wave-uniform
            if (MODE == Sweep::Forward) {
                // F rows: own RHS row minus up to two children's update rows (a missing contributor
                // reads the zero row)
                const int lc = min(l, s - 1);
                rb[i] = row_ptr(Bsrc, start + lc)[c];
                if constexpr (KIDS) {
                    // contributor rows taken as uniform values first, so the choice of the zero row
                    // for a missing one is scalar work too
                    const int ca = __builtin_amdgcn_readfirstlane(IDX_LDS ? sidx[2 * lc] : cown[2 * (start + lc)]);
                    const int cb = __builtin_amdgcn_readfirstlane(IDX_LDS ? sidx[2 * lc + 1] : cown[2 * (start + lc) + 1]);
                    ru[i] = row_ptr(U, ca >= 0 ? ca : zrow)[c];
                    rv[i] = row_ptr(U, cb >= 0 ? cb : zrow)[c];
                }
            } else {
                // [Y_S ; Z_B] rows: own rows from Y, boundary rows from Z
                const int lc = min(l, f - 1);
                const int prow = IDX_LDS ? sidx[lc] : (lc < s ? start + lc : B[lc - s]);
                rb[i] = row_ptr(lc < s ? Bsrc : Zsrc, prow)[c];
            }
        }
    };
    // value of staged right-hand-side element i
    auto bval = [&](int i) -> T {
        if constexpr (MODE == Sweep::Forward && KIDS)
            return rb[i] - ru[i] - rv[i];
        else
            return rb[i];
    };
    // panel element i as staged: 0 outside the node's depth, negated at the backward's boundary depths
    auto aval = [&](int i) -> PT {
        if constexpr (MODE == Sweep::Forward) {
            return ((fwd_keep >> i) & 1u) ? ra[i] : PT(0);
        } else if constexpr (BF) {
            const unsigned v = kA < s ? ra[i] : ra[i] ^ 0x80008000u;  // both bf16 halves negated
            return kA < f ? v : 0u;
        } else {
            const T v = kA < s ? ra[i] : -ra[i];
            return kA < f ? v : T(0);
        }
    };
    auto stash = [&]() {
        if constexpr (BF) {
#pragma unroll
            for (int i = 0; i < LA; ++i) {
                const int e = tid + i * kThreads;
                const unsigned v = aval(i);
                const int m = (MODE == Sweep::Forward) ? e % TM : e / kTileK;
                const int k = (MODE == Sweep::Forward) ? e / TM : e % kTileK;
                bfa[0][m][k] = (unsigned short)(v & 0xffffu);
                bfa[1][m][k] = (unsigned short)(v >> 16);
            }
#pragma unroll
            for (int i = 0; i < LB; ++i) {
                const int e = tid + i * kThreads;
                const unsigned sv = split_bf16(bval(i));
                bfb[0][e % TN][e / TN] = (unsigned short)(sv & 0xffffu);
                bfb[1][e % TN][e / TN] = (unsigned short)(sv >> 16);
            }
            return;
        }
#pragma unroll
        for (int i = 0; i < LA; ++i) {
            const int e = tid + i * kThreads;
            const T v = aval(i);
            if (MODE == Sweep::Forward)
                As[(e / TM) * kLdA + e % TM] = v;
            else
                As[(e % kTileK) * kLdA + e / kTileK] = v;
        }
#pragma unroll
        for (int i = 0; i < LB; ++i) {
            const int e = tid + i * kThreads;
            Bs[(e / TN) * kLdB + e % TN] = bval(i);
        }
    };

    typename Mfma<T>::acc_t acc[WBM][WBN];
#pragma unroll
    for (int i = 0; i < WBM; ++i)
#pragma unroll
        for (int jj = 0; jj < WBN; ++jj) acc[i][jj] = typename Mfma<T>::acc_t{T(0), T(0), T(0), T(0)};
    // accumulator value q (0..15) of this lane: block (q / (4 WBN), (q / 4) mod WBN), register q mod 4
#define FCW_ACC(q) acc[(q) / (4 * WBN)][((q) / 4) % WBN][(q) % 4]

    const int lrow = lane & 15, lk = lane >> 4;
    if (SPLIT == 2) {
        // the partial products of the depth parts, added in part order
#pragma unroll
        for (int q = 0; q < 16; ++q) FCW_ACC(q) = T(pslot[q]);
        for (int p = 1; p < nsplit; ++p)
#pragma unroll
            for (int q = 0; q < 16; ++q) FCW_ACC(q) += T(pslot[(int64_t)p * (kThreads * 16) + q]);
        kend = 0;  // no depth steps of its own
    }
    if (kbeg < kend) fetch(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += kTileK) {
        __syncthreads();
        stash();
        __syncthreads();
        if (k0 + kTileK < kend) fetch(k0 + kTileK);
        if constexpr (BF) {
            bf4 ah[WBM], al[WBM], bh[WBN], bl[WBN];
#pragma unroll
            for (int q = 0; q < WBM; ++q) {
                ah[q] = *reinterpret_cast<const bf4*>(&bfa[0][wm0 + 16 * q + lrow][4 * lk]);
                al[q] = *reinterpret_cast<const bf4*>(&bfa[1][wm0 + 16 * q + lrow][4 * lk]);
            }
#pragma unroll
            for (int q = 0; q < WBN; ++q) {
                bh[q] = *reinterpret_cast<const bf4*>(&bfb[0][wn0 + 16 * q + lrow][4 * lk]);
                bl[q] = *reinterpret_cast<const bf4*>(&bfb[1][wn0 + 16 * q + lrow][4 * lk]);
            }
#pragma unroll
            for (int i = 0; i < WBM; ++i)
#pragma unroll
                for (int jj = 0; jj < WBN; ++jj) {
                    flt4 c = acc[i][jj];
                    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[i], bh[jj], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[i], bl[jj], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al[i], bh[jj], c, 0, 0, 0);
                    acc[i][jj] = c;
                }
            continue;
        }
#pragma unroll
        for (int k4 = 0; k4 < kTileK; k4 += 4) {
            // four depths past the end are zeros in both stages: nothing to add
            if (k0 + k4 >= kend) break;
            T a[WBM], bb[WBN];
#pragma unroll
            for (int q = 0; q < WBM; ++q) a[q] = As[(k4 + lk) * kLdA + wm0 + 16 * q + lrow];
#pragma unroll
            for (int q = 0; q < WBN; ++q) bb[q] = Bs[(k4 + lk) * kLdB + wn0 + 16 * q + lrow];
#pragma unroll
            for (int i = 0; i < WBM; ++i)
#pragma unroll
                for (int jj = 0; jj < WBN; ++jj) acc[i][jj] = Mfma<T>::step(a[i], bb[jj], acc[i][jj]);
        }
    }
    if (SPLIT == 1) {
        double* const slot = pslot + (int64_t)ps * (kThreads * 16);
#pragma unroll
        for (int q = 0; q < 16; ++q) slot[q] = (double)FCW_ACC(q);
        return;
    }
#undef FCW_ACC

    if (MODE == Sweep::Forward) {
        const int64_t uoff = node_uoff[t.node];
#pragma unroll
        for (int i = 0; i < WBM; ++i) {
#pragma unroll
            for (int jj = 0; jj < WBN; ++jj) {
                const int col = n0 + wn0 + 16 * jj + lrow;
                // lane l, register r holds D[row(l >> 4, r)][l & 15]. Own rows go to Y; boundary rows
                // to U, plus the children's update rows passing through (tails). The loads of the
                // four values' tails are issued together and unconditionally (rows that are not
                // boundary rows read in-bounds stand-ins, missing tails the zero row), so only the
                // stores depend on the row.
                T xa[4], xb[4];
                if (KIDS) {  // a leaf has no children, so no update rows pass through it
                    int ta[4], tb[4];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = t.m0 + wm0 + 16 * i + Mfma<T>::row(lk, r);
                        const bool bnd = row >= s && row < M;
                        const int64_t ur = bnd ? uoff + (row - s) : 0;
                        ta[r] = ctail[2 * ur];
                        tb[r] = ctail[2 * ur + 1];
                        ta[r] = (bnd && ta[r] >= 0) ? ta[r] : zrow;
                        tb[r] = (bnd && tb[r] >= 0) ? tb[r] : zrow;
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        xa[r] = U[(int64_t)ta[r] * ldk + col];
                        xb[r] = U[(int64_t)tb[r] * ldk + col];
                    }
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = t.m0 + wm0 + 16 * i + Mfma<T>::row(lk, r);
                    if (row >= M) continue;
                    const T v = acc[i][jj][r];
                    if (row < s) {
                        Out[(int64_t)(start + row) * ldk + col] = v;
                    } else {
                        const int64_t ur = uoff + (row - s);
                        U[ur * ldk + col] = KIDS ? v + xa[r] + xb[r] : v;
                    }
                }
            }
        }
        return;
    }

    // Backward: store one solution value (and the refined iterate); returns it in FP64, 0 past M.
    auto emit = [&](int i, int jj, int r) -> double {
        const int col = n0 + wn0 + 16 * jj + lrow;
        const int row = t.m0 + wm0 + 16 * i + Mfma<T>::row(lk, r);
        if (row >= M) return 0.0;
        const T v = acc[i][jj][r];
        const int64_t o = (int64_t)(start + row) * ldk + col;
        if (!(acc_mode & 4)) Out[o] = v;
        double x = (double)v;
        if ((acc_mode & 3) == 1) acc64[o] = x;
        if ((acc_mode & 3) == 2) {
            x += acc64[o];
            acc64[o] = x;
        }
        return x;
    };
    if (!DT) {
#pragma unroll
        for (int i = 0; i < WBM; ++i)
#pragma unroll
            for (int jj = 0; jj < WBN; ++jj)
#pragma unroll
                for (int r = 0; r < 4; ++r) (void)emit(i, jj, r);
    } else if (sizeof(T) == sizeof(double)) {
        // Source-major copy through LDS, half a tile (HR rows x HC columns) at a time, so each
        // column's rows go out as one run of HR values.
        constexpr int HR = Sh::HR, HC = Sh::HC;
        static_assert(kTileK * (kLdA + kLdB) >= HC * (HR + 1), "transpose stage fits the A/B stages");
        double* stg = reinterpret_cast<double*>(sbuf);
#pragma unroll 1
        for (int h = 0; h < 2; ++h) {
            const int hr0 = TM == 64 ? HR * h : 0, hc0 = TM == 64 ? 0 : HC * h;  // this half's origin
            __syncthreads();  // the stages (or the previous half) are no longer read
            if (wm0 >= hr0 && wm0 < hr0 + HR && wn0 >= hc0 && wn0 < hc0 + HC) {
#pragma unroll
                for (int i = 0; i < WBM; ++i)
#pragma unroll
                    for (int jj = 0; jj < WBN; ++jj)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            stg[(wn0 - hc0 + 16 * jj + lrow) * (HR + 1) + wm0 - hr0 + 16 * i + Mfma<T>::row(lk, r)] =
                                emit(i, jj, r);
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 2048 / kThreads; ++q) {
                const int idx = tid + q * kThreads;
                const int lc = idx / HR, lr = idx % HR;
                const int row = t.m0 + hr0 + lr;
                if (row < M) Dt[(int64_t)(n0 + hc0 + lc) * ldd + start + row] = stg[lc * (HR + 1) + lr];
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < WBM; ++i)
#pragma unroll
            for (int jj = 0; jj < WBN; ++jj)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const double x = emit(i, jj, r);
                    const int row = t.m0 + wm0 + 16 * i + Mfma<T>::row(lk, r);
                    if (row < M) Dt[(int64_t)(n0 + wn0 + 16 * jj + lrow) * ldd + start + row] = x;
                }
    }
}

// Workgroups of one sweep launch (one depth part) over ntiles row tiles of height tm.
inline int grid_of(int ntiles, int ldk, int tm) { return (ntiles + kXcds - 1) / kXcds * kXcds * (ldk / (4096 / tm)); }

template <typename T, int TM, typename PT>
void fwd_gemm_tm(hipStream_t st, const LevelTile* tiles, int ntiles, int ldk, const int* node_start, const int* node_s,
                 const int* node_b, const int64_t* node_poff, const int64_t* node_uoff, const int* contrib_own,
                 const int* contrib_tail, const PT* panel, const T* X, T* Y, T* U, bool idx_lds, bool leaves, int zrow,
                 int nsplit, double* part) {
    const dim3 grid(grid_of(ntiles, ldk, TM));  // 8 runs of ceil(ntiles / 8)
    const dim3 grid_split(grid.x * (nsplit > 1 ? nsplit : 1));
#define FCW_FWD(LDS, KIDS, SPL, G)                                                                                      \
    fcw::launch(sweep_kernel<T, TM, Sweep::Forward, LDS, KIDS, false, SPL, PT>, dim3(G), dim3(kThreads), 0, st, ntiles, \
                tiles, ldk, node_start, node_s, node_b, nullptr, nullptr, node_poff, node_uoff, contrib_own,           \
                contrib_tail, panel, X, nullptr, Y, U, nullptr, 0, nullptr, 0, zrow, nsplit, part)
    if constexpr (sizeof(T) == sizeof(double) && std::is_same<T, PT>::value && TM == 64) {
        if (nsplit > 1 && !leaves) {
            if (idx_lds) {
                FCW_FWD(true, true, 1, grid_split);
                FCW_FWD(true, true, 2, grid);
            } else {
                FCW_FWD(false, true, 1, grid_split);
                FCW_FWD(false, true, 2, grid);
            }
            return;
        }
    }
    if (leaves)
        FCW_FWD(true, false, 0, grid);
    else if (idx_lds)
        FCW_FWD(true, true, 0, grid);
    else
        FCW_FWD(false, true, 0, grid);
#undef FCW_FWD
}

template <typename T, typename PT = T>
void fwd_gemm(hipStream_t st, const LevelTile* tiles, int ntiles, int ldk, int tm, const int* node_start,
              const int* node_s, const int* node_b, const int64_t* node_poff, const int64_t* node_uoff,
              const int* contrib_own, const int* contrib_tail, const PT* panel, const T* X, T* Y, T* U, bool idx_lds,
              bool leaves, int zrow, int nsplit, double* part) {
    if (ntiles <= 0) return;
    if (tm == 32)
        fwd_gemm_tm<T, 32, PT>(st, tiles, ntiles, ldk, node_start, node_s, node_b, node_poff, node_uoff, contrib_own,
                               contrib_tail, panel, X, Y, U, idx_lds, leaves, zrow, 1, nullptr);
    else
        fwd_gemm_tm<T, 64, PT>(st, tiles, ntiles, ldk, node_start, node_s, node_b, node_poff, node_uoff, contrib_own,
                               contrib_tail, panel, X, Y, U, idx_lds, leaves, zrow, nsplit, part);
}

template <typename T, int TM, typename PT>
void bwd_gemm_tm(hipStream_t st, const LevelTile* tiles, int ntiles, int ldk, const int* node_start, const int* node_s,
                 const int* node_b, const int64_t* node_boff, const int* bidx, const int64_t* node_poff,
                 const PT* panel, const T* Y, T* Z, bool idx_lds, double* acc64, int acc_mode, double* Dt, int ldd,
                 int nsplit, double* part) {
    const dim3 grid(grid_of(ntiles, ldk, TM));
    const dim3 grid_split(grid.x * (nsplit > 1 ? nsplit : 1));
#define FCW_BWD(LDS, DTW, SPL, G)                                                                                      \
    fcw::launch(sweep_kernel<T, TM, Sweep::Backward, LDS, true, DTW, SPL, PT>, dim3(G), dim3(kThreads), 0, st, ntiles, \
                tiles, ldk, node_start, node_s, node_b, node_boff, bidx, node_poff, nullptr, nullptr, nullptr, panel,  \
                Y, Z, Z, nullptr, acc64, acc_mode, Dt, ldd, 0, nsplit, part)
#define FCW_BWD_ALL(SPL, G)           \
    if (idx_lds && Dt)                \
        FCW_BWD(true, true, SPL, G);  \
    else if (idx_lds)                 \
        FCW_BWD(true, false, SPL, G); \
    else if (Dt)                      \
        FCW_BWD(false, true, SPL, G); \
    else                              \
        FCW_BWD(false, false, SPL, G)
    if constexpr (sizeof(T) == sizeof(double) && std::is_same<T, PT>::value && TM == 64) {
        if (nsplit > 1) {
            FCW_BWD_ALL(1, grid_split);
            FCW_BWD_ALL(2, grid);
            return;
        }
    }
    FCW_BWD_ALL(0, grid);
#undef FCW_BWD_ALL
#undef FCW_BWD
}

template <typename T, typename PT = T>
void bwd_gemm(hipStream_t st, const LevelTile* tiles, int ntiles, int ldk, int tm, const int* node_start,
              const int* node_s, const int* node_b, const int64_t* node_boff, const int* bidx,
              const int64_t* node_poff, const PT* panel, const T* Y, T* Z, bool idx_lds, double* acc64, int acc_mode,
              double* Dt, int ldd, int nsplit, double* part) {
    if (ntiles <= 0) return;
    if (tm == 16)
        bwd_gemm_tm<T, 16, PT>(st, tiles, ntiles, ldk, node_start, node_s, node_b, node_boff, bidx, node_poff, panel, Y,
                               Z, idx_lds, acc64, acc_mode, Dt, ldd, 1, nullptr);
    else if (tm == 32)
        bwd_gemm_tm<T, 32, PT>(st, tiles, ntiles, ldk, node_start, node_s, node_b, node_boff, bidx, node_poff, panel, Y,
                               Z, idx_lds, acc64, acc_mode, Dt, ldd, 1, nullptr);
    else
        bwd_gemm_tm<T, 64, PT>(st, tiles, ntiles, ldk, node_start, node_s, node_b, node_boff, bidx, node_poff, panel, Y,
                               Z, idx_lds, acc64, acc_mode, Dt, ldd, nsplit, part);
}

__global__ void split_bf16_kernel(const double* __restrict__ src, uint32_t* __restrict__ dst, int64_t count) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        const double x = src[i];
        const unsigned hi = bf16_rne((float)x);
        const double rest = x - (double)__uint_as_float(hi << 16);
        dst[i] = (bf16_rne((float)rest) << 16) | hi;
    }
}

}  // namespace

void launch_to_bf16x2(hipStream_t st, const double* src, uint32_t* dst, int64_t count) {
    if (count <= 0) return;
    const int64_t blocks = std::min<int64_t>((count + 255) / 256, 65536);
    fcw::launch(split_bf16_kernel, dim3(dim3((unsigned)blocks)), dim3(256), 0, st, src, dst, count);
}

bool sweep_tile_ok(bool forward, int tm, int ldk) {
    if (tm == 64) return ldk % 64 == 0;
    if (tm == 32) return ldk % 128 == 0;
    return !forward && tm == 16 && ldk % 256 == 0;
}

void launch_fwd_gemm(hipStream_t st, const LevelTile* tiles, int ntiles, int ldk, int tm, const int* node_start,
                     const int* node_s, const int* node_b, const int64_t* node_poff, const int64_t* node_uoff,
                     const int* contrib_own, const int* contrib_tail, const uint32_t* panel, const float* X, float* Y,
                     float* U, bool idx_lds, bool leaves, int zrow) {
    fwd_gemm<float, uint32_t>(st, tiles, ntiles, ldk, tm, node_start, node_s, node_b, node_poff, node_uoff, contrib_own,
                              contrib_tail, panel, X, Y, U, idx_lds, leaves, zrow, 1, nullptr);
}

void launch_bwd_gemm(hipStream_t st, const LevelTile* tiles, int ntiles, int ldk, int tm, const int* node_start,
                     const int* node_s, const int* node_b, const int64_t* node_boff, const int* bidx,
                     const int64_t* node_poff, const uint32_t* panel, const float* Y, float* Z, bool idx_lds,
                     double* acc64, int acc_mode, double* Dt, int ldd) {
    bwd_gemm<float, uint32_t>(st, tiles, ntiles, ldk, tm, node_start, node_s, node_b, node_boff, bidx, node_poff, panel,
                              Y, Z, idx_lds, acc64, acc_mode, Dt, ldd, 1, nullptr);
}

void launch_fwd_gemm(hipStream_t st, const LevelTile* tiles, int ntiles, int ldk, int tm, const int* node_start,
                     const int* node_s, const int* node_b, const int64_t* node_poff, const int64_t* node_uoff,
                     const int* contrib_own, const int* contrib_tail, const double* panel, const double* X, double* Y,
                     double* U, bool idx_lds, bool leaves, int zrow, int nsplit, double* part) {
    fwd_gemm<double>(st, tiles, ntiles, ldk, tm, node_start, node_s, node_b, node_poff, node_uoff, contrib_own,
                     contrib_tail, panel, X, Y, U, idx_lds, leaves, zrow, nsplit, part);
}

void launch_fwd_gemm(hipStream_t st, const LevelTile* tiles, int ntiles, int ldk, int tm, const int* node_start,
                     const int* node_s, const int* node_b, const int64_t* node_poff, const int64_t* node_uoff,
                     const int* contrib_own, const int* contrib_tail, const float* panel, const float* X, float* Y,
                     float* U, bool idx_lds, bool leaves, int zrow) {
    fwd_gemm<float>(st, tiles, ntiles, ldk, tm, node_start, node_s, node_b, node_poff, node_uoff, contrib_own,
                    contrib_tail, panel, X, Y, U, idx_lds, leaves, zrow, 1, nullptr);
}

void launch_bwd_gemm(hipStream_t st, const LevelTile* tiles, int ntiles, int ldk, int tm, const int* node_start,
                     const int* node_s, const int* node_b, const int64_t* node_boff, const int* bidx,
                     const int64_t* node_poff, const double* panel, const double* Y, double* Z, bool idx_lds,
                     double* Dt, int ldd, int nsplit, double* part) {
    bwd_gemm<double>(st, tiles, ntiles, ldk, tm, node_start, node_s, node_b, node_boff, bidx, node_poff, panel, Y, Z,
                     idx_lds, nullptr, 0, Dt, ldd, nsplit, part);
}

void launch_bwd_gemm(hipStream_t st, const LevelTile* tiles, int ntiles, int ldk, int tm, const int* node_start,
                     const int* node_s, const int* node_b, const int64_t* node_boff, const int* bidx,
                     const int64_t* node_poff, const float* panel, const float* Y, float* Z, bool idx_lds,
                     double* acc64, int acc_mode, double* Dt, int ldd) {
    bwd_gemm<float>(st, tiles, ntiles, ldk, tm, node_start, node_s, node_b, node_boff, bidx, node_poff, panel, Y, Z,
                    idx_lds, acc64, acc_mode, Dt, ldd, 1, nullptr);
}

}  // namespace fcw
