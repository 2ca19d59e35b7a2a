// Device kernel launchers (kernels.hip) used by the C-ABI runtime (fcw_api.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>

namespace fcw {

// Every kernel of the library is started through launch(): issued on the stream, or, while a
// GraphRecorder is active on the calling thread, appended to the recorder's graph as the next node
// of one chain (each node depends on the one before). Graphs are built this way rather than by
// stream capture, so no stream ever enters capture mode and no other thread's work (in this
// library or not) is affected while a graph is being built.
struct GraphRecorder {
    hipGraph_t graph = nullptr;
    hipGraphNode_t last = nullptr;
};
extern thread_local GraphRecorder* t_recorder;

template <class Tuple, std::size_t... I>
inline void kernel_arg_ptrs(Tuple& t, void** ptrs, std::index_sequence<I...>) {
    ((ptrs[I] = static_cast<void*>(&std::get<I>(t))), ...);
}

template <class... KArgs, class... Args>
inline void launch(void (*kernel)(KArgs...), dim3 grid, dim3 block, size_t shmem, hipStream_t st, Args&&... args) {
    static_assert(sizeof...(KArgs) == sizeof...(Args), "kernel argument count");
    if (!t_recorder) {
        kernel<<<grid, block, shmem, st>>>(std::forward<Args>(args)...);
        return;
    }
    std::tuple<KArgs...> vals(std::forward<Args>(args)...);  // the kernel's own parameter types
    void* ptrs[sizeof...(KArgs) > 0 ? sizeof...(KArgs) : 1];
    kernel_arg_ptrs(vals, ptrs, std::index_sequence_for<KArgs...>{});
    hipKernelNodeParams p{};
    p.func = reinterpret_cast<void*>(kernel);
    p.gridDim = grid;
    p.blockDim = block;
    p.sharedMemBytes = (unsigned)shmem;
    p.kernelParams = ptrs;
    p.extra = nullptr;
    hipGraphNode_t node = nullptr;
    const hipError_t e = hipGraphAddKernelNode(&node, t_recorder->graph, t_recorder->last ? &t_recorder->last : nullptr,
                                               t_recorder->last ? 1 : 0, &p);
    if (e != hipSuccess) throw std::runtime_error(std::string("hipGraphAddKernelNode failed: ") + hipGetErrorString(e));
    t_recorder->last = node;
}

struct LevelTile {
    int node;
    int m0;
};

// Depth per step of the level panel products (the row tile height, 64, 32 or 16, is chosen per
// level and direction; a workgroup's column slice is 4096 / height wide).
constexpr int kTileK = 16;
constexpr int kColPad = 64;  // RHS block widths are padded to this many columns
// Per-node row indices staged in LDS by the sweeps: the forward sweep needs s <= kSweepIdxCap
// (two contributor rows per own row), the backward sweep s + b <= 2 * kSweepIdxCap (one source
// row per depth index).
constexpr int kSweepIdxCap = 1024;

// Doubles of band scratch per source for pool_cap band faces: a, b, c and the face index.
__host__ __device__ inline int64_t band_pool_doubles(int64_t pool_cap) { return 3 * pool_cap + (pool_cap + 1) / 2; }

// Metric kernel limits.
constexpr int kMaxScales = 32;
constexpr int kMaxExtras = 256;
constexpr int kMaxBins = 1536;
constexpr int kMaxChunks = 2048;  // 256-face chunks with recorded distance bounds (larger meshes: no skipping)
// 512 threads per source: a workgroup then holds half a CU's registers, so the phases' latency
// overlaps other streams' work on the same CU (148k-vertex surface, 256-source blocks: 1024 threads
// 90.5-91.5 k sources/s, 512 threads 93.3 k)
constexpr int kMetricThreads = 512;

// Per-source scalar outputs, in this order (double each).
enum ScalarSlot {
    kSlotFlag = 0,
    kSlotNeg,
    kSlotViol,
    kSlotDistB,
    kSlotMsdU,
    kSlotMsdW,
    kSlotNEval,
    kSlotNSamples,
    kSlotOverflow,
    kSlotCount
};

struct MeshDev {
    int n, m;
    const double* V;        // n x 3, positions
    // the surface's faces (metric kernels)
    const int* faces;       // m x 3, positions
    const double* larea;    // m (Gram area)
    const double* g11;
    const double* g12;
    const double* g22;
    // the operator triangulation (L, M, divergence): the surface's own faces, or its intrinsic
    // Delaunay triangulation (same vertices, other faces)
    int mo;
    const int* op_faces;    // mo x 3, positions
    const double* op_area;  // mo
    const double* op_gram;  // mo x 6: grad B_a . grad B_b, stored G00 G01 G02 G11 G12 G22
    const int* op_vf_ptr;   // n + 1
    const int* op_vf;       // face*3 + corner
};

// Row groups of the divergence (built once per mesh): consecutive vertex rows whose incidences
// touch at most kDivVerts distinct vertices, at most kDivInc incidences, at most kDivGroupRows rows.
constexpr int kDivVerts = 64;
constexpr int kDivInc = 128;
constexpr int kDivGroupRows = 64;
struct DivDev {
    int ngroups;
    const int4* grp;       // ngroups + 1: first row, first incidence, first distinct vertex (into verts), count
                           // (-1: one row beyond the limits, read from global memory); entry ngroups has .x = n
    const int* verts;      // distinct vertex positions of each group, ascending
    const unsigned* loc;   // per incidence: local indices of the face's vertices 0, 1, 2 (6 bits each)
    const double* geo;     // per incidence: A g1.ga, A g2.ga, |g1|^2, 2 g1.g2, |g2|^2
    const int* vf_ptr;     // n + 1, the operator triangulation's incidences per row
    const int* vf;         // face*3 + corner
    const int* faces;      // operator faces, positions
};

// Mesh preparation on the device (prep.hip): caller-owned host buffers, sized for the input mesh
// (n vertices, m faces, adjacency 6m).
struct MeshPrepOut {
    int n_sub, m_sub, n_boundary, n_adj;
    int* sub_to_orig;     // n_sub
    int* orig_to_sub;     // n (-1 outside the submesh)
    int* faces;           // 3 m_sub, submesh numbering
    double* vertices;     // 3 n_sub
    int* boundary;        // n_boundary, ascending
    int* order;           // n_sub, breadth-first order over every component
    double* face_areas;   // m_sub
    double* face_L;       // m_sub, longest edge
    double* vertex_areas; // n_sub
    int* adj_ptr;         // n_sub + 1
    int* adj;             // n_adj, ascending per vertex
};
int mesh_prep_device(int n, const double* V, int m, const int* F, const unsigned char* mask, int start, int device,
                     MeshPrepOut& out, std::string& err);

struct MetricsDev {
    int S, q, nb;
    const double* targets;  // S (ascending)
    const double* fracs;    // q
    const int* boundary;    // nb, positions
    const double* vw;       // n, vertex areas (positions)
    double eps, health_h, width, cap;  // cap < 0: no cap
    double rtol;
    int max_iter;
    int pool_cap;           // band faces per source held in the scratch (all bands together)
    double* band;           // per source band_pool_doubles(pool_cap): sorted a, b, c per band face (three
                            // arrays of pool_cap doubles), then the faces' indices (pool_cap ints)
    // Far-field summaries from chunk_stats (null: the metric kernel makes its own vertex pass and
    // chunk bounds): per source and 256-face chunk, the distance range of the chunk's vertices
    // and kChunkStats partial sums over the vertices the chunk owns.
    int nch;
    const float* ch_lo;        // rounded down, at chunk_range_index(j, c)
    const float* ch_hi;        // rounded up, at chunk_range_index(j, c)
    const double* part;        // at chunk_rec_index(j, c, q)
    const double* chunk_area;  // nch: area of the chunk's faces
    // Fields are stored unshifted when shift_src is set: source j's distances are
    // Dt[j][p] - shift_src[src_pos[j] * shift_ld + j] (its solution at its own position).
    const double* shift_src;
    int shift_ld;
};

// Partial record of chunk_stats: sum d (d > eps), sum d*w, sum w, max finite d, min finite d on the
// boundary, then as exact counts: d < -h, chord violations, d > eps, non-finite, finite boundary.
constexpr int kChunkStats = 10;

// Where chunk_stats puts source j's record q of chunk c (and its distance range). Sources are grouped
// by eight with the same j mod 8 within their 64-source column slice: a wave of 64 consecutive
// sources then writes whole 64-byte lines (eight groups of eight doubles), and the eight sources of a
// group are metric workgroups that run on one XCD (workgroup j goes to XCD j mod 8), so they share
// those lines in its L2. Arrays hold ldk x nch x kChunkStats records (ldk x nch ranges).
__host__ __device__ inline int64_t chunk_rec_index(int j, int c, int q, int nch) {
    const int64_t g = (int64_t)(j >> 6) * 8 + (j & 7);
    return ((g * nch + c) * kChunkStats + q) * 8 + ((j >> 3) & 7);
}
__host__ __device__ inline int64_t chunk_range_index(int j, int c, int nch) {
    const int64_t g = (int64_t)(j >> 6) * 8 + (j & 7);
    return (g * nch + c) * 8 + ((j >> 3) & 7);
}

struct MetricsOut {
    double* scal;      // k x kSlotCount
    double* radii;     // k x S
    double* perim;     // k x S
    double* extra_r;   // k x (S-1)q
    double* extra_a;   // k x (S-1)q
    float* samp_r;     // k x (S + (S-1)q)
    float* samp_a;     // k x (S + (S-1)q)
};

void launch_face_geometry(hipStream_t st, int m, const double* V, const int* faces, double* area, double* gram,
                          double* larea, double* g11, double* g12, double* g22);
void launch_assemble(hipStream_t st, int n, int nnz, const int* contrib_ptr, const int* contrib, const double* area,
                     const double* gram, const int* vf_ptr, const int* vf, double* lval, double* mass);
void launch_fill(hipStream_t st, double* p, int64_t count, double value);
void launch_calibration_copy(hipStream_t st, const void* src, void* dst, int64_t bytes);
void launch_delta_rhs(hipStream_t st, double* X, int n, int ldk, const int* src_pos, int k);
// X[src_pos[j]][j] = 1 only (the caller has zeroed every row the sweep reads).
void launch_set_delta(hipStream_t st, double* X, int ldk, const int* src_pos, int k);
// Zero the own rows (in Y) and, with U set, the update rows of the listed supernodes.
void launch_clear_rows(hipStream_t st, const int* nodes, int count, int ldk, const int* node_start,
                       const int* node_s, const int* node_b, const int64_t* node_uoff, double* Y, double* U);
void launch_fwd_gemm(hipStream_t st, const LevelTile* tiles, int ntiles, int ldk, int tm, const int* node_start,
                     const int* node_s, const int* node_b, const int64_t* node_poff, const int64_t* node_uoff,
                     const int* contrib_own, const int* contrib_tail, const double* panel, const double* X, double* Y,
                     double* U, bool idx_lds, bool leaves, int zrow, int nsplit = 1, double* part = nullptr);
void launch_bwd_gemm(hipStream_t st, const LevelTile* tiles, int ntiles, int ldk, int tm, const int* node_start,
                     const int* node_s, const int* node_b, const int64_t* node_boff, const int* bidx,
                     const int64_t* node_poff, const double* panel, const double* Y, double* Z, bool idx_lds,
                     double* Dt, int ldd, int nsplit = 1, double* part = nullptr);
// tm: row tile height of the launch's tiles (every tile's m0 a multiple of it): 64 (ldk a multiple
// of 64), 32 (ldk a multiple of 128) or, backward only, 16 (ldk a multiple of 256); see sweep_tile_ok.
bool sweep_tile_ok(bool forward, int tm, int ldk);
// Forward sweeps: U has a row zrow past the update rows that stays zero (it stands in for a
// missing child contribution, so every gather is an unconditional load). nsplit > 1 (FP64, not
// the leaf level): the depth is split over nsplit workgroups per (tile, slice) whose partial
// products go to `part` (ntiles x ldk/64 x nsplit x 4096 doubles), then added in order by a second
// launch that writes the results.
// Backward sweeps with Dt set also write the FP64 solution source-major, Dt[col * ldd + position]
// (unshifted; consumers subtract the source's own value).
// FP32 forms of the sweeps (refined solve): FP32 panels and right-hand sides, FP32 matrix-core
// accumulation; the backward sweep also stores (acc_mode 1) or adds (2) its solution into acc64, and
// with 4 added to acc_mode does not store the FP32 solution itself (rows no other node reads).
void launch_fwd_gemm(hipStream_t st, const LevelTile* tiles, int ntiles, int ldk, int tm, const int* node_start,
                     const int* node_s, const int* node_b, const int64_t* node_poff, const int64_t* node_uoff,
                     const int* contrib_own, const int* contrib_tail, const float* panel, const float* X, float* Y,
                     float* U, bool idx_lds, bool leaves, int zrow);
void launch_bwd_gemm(hipStream_t st, const LevelTile* tiles, int ntiles, int ldk, int tm, const int* node_start,
                     const int* node_s, const int* node_b, const int64_t* node_boff, const int* bidx,
                     const int64_t* node_poff, const float* panel, const float* Y, float* Z, bool idx_lds,
                     double* acc64, int acc_mode, double* Dt, int ldd);
// Split-BF16 forms (refined solve, FCW_BF16X3_IR): panels packed as bf16 hi | lo << 16 (from the FP64
// panels by launch_to_bf16x2), FP32 right-hand sides split while staged, three bf16 products per block.
void launch_to_bf16x2(hipStream_t st, const double* src, uint32_t* dst, int64_t count);
void launch_fwd_gemm(hipStream_t st, const LevelTile* tiles, int ntiles, int ldk, int tm, const int* node_start, const int* node_s,
                     const int* node_b, const int64_t* node_poff, const int64_t* node_uoff, const int* contrib_own,
                     const int* contrib_tail, const uint32_t* panel, const float* X, float* Y, float* U, bool idx_lds,
                     bool leaves, int zrow);
void launch_bwd_gemm(hipStream_t st, const LevelTile* tiles, int ntiles, int ldk, int tm, const int* node_start, const int* node_s,
                     const int* node_b, const int64_t* node_boff, const int* bidx, const int64_t* node_poff,
                     const uint32_t* panel, const float* Y, float* Z, bool idx_lds, double* acc64, int acc_mode,
                     double* Dt, int ldd);
// r = b - (L + shift M) x in FP64 (x null: r = b), rounded to FP32 (the refined solve's RHS).
void launch_residual32(hipStream_t st, int n, int ldk, const int* rowptr, const int* colind, const double* lval,
                       const double* mass, double shift, const double* b, const double* x, float* r);
// The same residual with x staged per divergence row group in LDS (x non-null): aval = L + shift M in
// CSR order (launch_shifted_values), cloc = each CSR entry's local vertex index in its row's group.
void launch_residual_grouped(hipStream_t st, const DivDev& d, int ldk, const int* rowptr, const int* colind,
                             const unsigned char* cloc, const double* aval, const double* b, const double* x,
                             float* r);
void launch_shifted_values(hipStream_t st, int n, const int* rowptr, const int* colind, const double* lval,
                           const double* mass, double shift, double* aval);
void launch_to_float(hipStream_t st, const double* src, float* dst, int64_t count);
// Per-incidence coefficients of the divergence, computed once per mesh (see DivDev::geo).
void launch_incidence_geometry(hipStream_t st, int ninc, const int* vf, const double* area, const double* gram,
                               double* geo);
// out32 (optional): the same values rounded to FP32 (the refined solve's first right-hand side).
void launch_divergence(hipStream_t st, const DivDev& d, int ldk, const double* u, double* out, float* out32 = nullptr);
// out[j][v] = Dt[j][iperm[v]] - phi[src_pos[j]][j]: fields in caller vertex order, 0 at the source.
void launch_gather_rows(hipStream_t st, int n, int ldk, int k, const double* Dt, const double* phi, const int* src_pos,
                        const int* iperm, double* out);
void launch_chunk_area(hipStream_t st, int m, const double* larea, double* chunk_area);
void launch_chunk_stats(hipStream_t st, int k, int ldk, int nch, const double* Z, const int* src_pos, const double* V,
                        const double* vw, const int* isb, const int* cvptr, const int* cverts, const int* cnown,
                        double eps, double h, float* ch_lo, float* ch_hi, double* part);
void launch_metrics(hipStream_t st, int k, const MeshDev& mesh, const MetricsDev& md, const double* Dt,
                    const int* src_pos, MetricsOut out, bool health_screen);

}  // namespace fcw
