/*
 * fcw.h - C ABI of libfcw.so, the MI355X (gfx950) heat-method geodesic and
 * wiring-metric engine. Plain C types only; the caller owns every host buffer,
 * the library owns device memory per context. Every call returns 0 on success
 * and a nonzero status on failure, with the message in fcw_last_error()
 * (thread-local). A context runs its RHS blocks on its own 1-16 HIP streams
 * (fcw_options.streams) and synchronises them before each call returns; calls on
 * one context must be serialised by the caller, contexts are independent.
 *
 * Vertex and face arrays describe the cortical submesh, as the reference hands
 * it to its distance engine (core_analysis.py:646-667).
 */
#ifndef FCW_H
#define FCW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fcw_ctx fcw_ctx;

/* Arithmetic of the two triangular sweeps per system. */
enum {
    FCW_FP64 = 0,    /* plain FP64 panels and RHS blocks */
    FCW_FP32_IR = 1, /* Poisson system: FP32 panel products with FP32 accumulation, refined against the FP64
                        residual; the heat system stays FP64 (its solution spans 1 to ~1e-141 at the
                        164k-vertex mesh, below FP32's range) */
    FCW_BF16X3_IR = 2 /* as FCW_FP32_IR with split-BF16 panel products: panel and right-hand side each split
                        into bf16 hi + lo, three bf16 matrix-core products (hi*hi + hi*lo + lo*hi) with FP32
                        accumulation, refined against the FP64 residual */
};

/* Scalar outputs per source, in this order (fcw_metrics_out.scalars is k x FCW_NSCALARS). */
enum {
    FCW_S_FLAG = 0,      /* 1 if the field failed the health screen (lean_geometry.py:343-357) */
    FCW_S_NEG_FRAC,      /* fraction of vertices with d < -h */
    FCW_S_VIOL_FRAC,     /* fraction with d < 0.9 |x - x_src| - h */
    FCW_S_DIST_BOUNDARY, /* min finite d over boundary vertices; +inf on a closed mesh */
    FCW_S_MSD_U,         /* mean d over vertices with d > eps */
    FCW_S_MSD_W,         /* vertex-area weighted mean over the same vertices */
    FCW_S_N_EVAL,        /* band evaluations used by the radius/perimeter/sample solves */
    FCW_S_N_SAMPLES,     /* cleaned (radius, area) pairs written for this source */
    FCW_S_OVERFLOW,      /* 1 if a band list overflowed and the full-face fallback ran */
    FCW_NSCALARS
};

typedef struct {
    double t;              /* diffusion time; <= 0 selects t_coef * mean_edge^2 */
    double t_coef;         /* default 1.0 (fastcw_lean.py:481, 485-489) */
    double poisson_shift;  /* default 1e-10: Poisson matrix L + shift*M (fastcw_lean.py:500) */
    int precision;         /* FCW_FP64, FCW_FP32_IR or FCW_BF16X3_IR */
    int block;             /* maximum sources per RHS block (padded to a multiple of 64); default 256 */
    int leaf_size;         /* nested-dissection leaf size; default 64 */
    int host_threads;      /* threads for the host factorisation; 0 = OpenMP default */
    int device;            /* HIP device ordinal */
    int ir_steps;          /* refined forms: residual + correction solves after the first solve; < 0 (default):
                              2 for FCW_FP32_IR (one step leaves ~2e-7 of the largest distance, 6e-5 relative
                              on the smallest distances to the boundary of the 164k-vertex surface), 6 for
                              FCW_BF16X3_IR (4 leave 4.5e-5 relative there) */
    int streams;           /* RHS blocks in flight on separate HIP streams (own workspaces), 1..16;
                              default 8 (run concurrently when GPU_MAX_HW_QUEUES >= streams) */
    int sparse_rhs;        /* 1: the heat forward sweep visits only supernodes on source-to-root paths */
    int robust;            /* 1: operators on the intrinsic Delaunay triangulation of the mollified surface,
                              potpourri3d's use_robust=True (distance_engines.py:78-79, 119-125); default 0 */
    int split_depth;       /* FP64 sweeps of the narrow upper tree levels split their depth over several
                              workgroups (lower latency per block, more work in all): 1 on, -1 off, 0 (default)
                              on when streams <= 2 */
    int band_cap;          /* band faces per scale held per source by the metric kernel (default 6144); a band
                              that does not fit is evaluated from every face instead (same result, slower) */
    int graphs;            /* 1 (default): each stream records the launches that are the same for every full
                              block (the heat backward sweep; divergence + Poisson solve) into HIP graphs once
                              (built node by node, no stream capture) and replays them; 0: every launch issued
                              one by one */
    int tile_rows;         /* row tile height of the panel sweeps: 0 (default) chooses per tree level and
                              direction among 64, 32 and 16 by a model of the work each executes (32 needs a
                              block width that is a multiple of 128; 16, backward only, a multiple of 256);
                              16, 32 or 64: the smallest allowed height >= this on every level */
    int index_staging;     /* 1 (default): each sweep workgroup stages its node's row maps (boundary rows,
                              child contributor rows) in LDS when every node's fit (fronts up to 2,048 rows);
                              0: every sweep reads them from global memory, the path larger fronts take */
} fcw_options;

typedef struct {
    int n, m, nnodes, nlevels, max_front, block;
    int64_t factor_nnz;         /* sum over supernodes of s(s+1)/2 + s*b */
    int64_t panel_elems;        /* stored panel entries per system */
    double factor_flops;        /* per numeric factorisation */
    double solve_flops_per_rhs; /* executed panel-product flops, one system, forward + backward */
    double t, mean_edge;
    double setup_seconds[4];    /* analyse, assemble, factor (both systems), upload */
    int64_t device_bytes;       /* device memory held by the context */
    int robust;                 /* operators on the intrinsic Delaunay triangulation */
    int intrinsic_flips;        /* edge flips made (robust) */
    int intrinsic_skipped;      /* non-Delaunay edges left because flipping them would fold a face */
    double mollify_delta;       /* length added to every edge by the mollification (robust) */
    int split_levels;           /* sweep launches per block whose depth is split (0: none) */
    int op_faces;               /* faces of the operator triangulation (robust: intrinsic, merged cover) */
    int tufted_cover;           /* robust: 1 if the operators were built on the tufted cover */
    int nonmanifold_edges;      /* robust: input edges with three or more faces */
} fcw_info;

typedef struct {
    double* scalars;     /* k x FCW_NSCALARS */
    double* radii;       /* k x S */
    double* perimeters;  /* k x S */
    double* extra_r;     /* k x (S-1)q */
    double* extra_a;     /* k x (S-1)q */
    float* sample_r;     /* k x (S + (S-1)q), first scalars[FCW_S_N_SAMPLES] entries valid */
    float* sample_a;     /* k x (S + (S-1)q) */
} fcw_metrics_out;

void fcw_default_options(fcw_options* opt);

/* Build operators, factor M + tL and L + shift*M, upload everything.
 * Replaces pp3d.MeshHeatMethodDistanceSolver(V, F, t_coef=..., use_robust=...)
 * (distance_engines.py:119-125) and ConsistentBatchHeatEngine.__init__
 * (fastcw_lean.py:481-502). Zero-area faces are rejected as the latter does
 * (fastcw_lean.py:383-384). */
int fcw_create(const double* V, int n, const int* F, int m, const fcw_options* opt, fcw_ctx** out);
int fcw_destroy(fcw_ctx* ctx);
int fcw_get_info(const fcw_ctx* ctx, fcw_info* out);

/* Distances from k sources, row per source: out is k x n float64 (input vertex order), shifted to
 * 0 at each source. Replaces compute_distance (distance_engines.py:131-133) and
 * ConsistentBatchHeatEngine.compute_distance_batch (fastcw_lean.py:504-531). k <= block. */
int fcw_distance_block(fcw_ctx* ctx, const int* src, int k, double* out);

/* The fused gradient -> normalise -> divergence of k fields on the context's operator faces:
 * out[v][j] = sum over faces f at v of A_f (X_f . grad B_v), X_f = -grad u_j / |grad u_j| on f (no term
 * where |grad u_j| vanishes). u and out are n x k, row per vertex (input vertex order), as the (n, k)
 * arrays of fused_normalized_divergence (fastcw_lean.py:416-442), which this replaces; the same kernel
 * runs between the heat and Poisson solves of fcw_distance_block and fcw_wiring_sources. k <= block. */
int fcw_divergence_block(fcw_ctx* ctx, const double* u, int k, double* out);

/* Per-source metric configuration (core_analysis.py:1513-1514, 1536, 1588, 705):
 * targets = sorted scale * total area (S <= 32), fracs = extra-sample fractions (q), boundary =
 * submesh boundary vertices, vertex_w = vertex areas (n), eps = MSD threshold (1e-6), health_h =
 * median face max-edge, bin_width = lean bucket width, cap_fraction < 0 disables the boundary cap,
 * rtol / max_iter = radius solve tolerance (1e-10, 60). */
int fcw_set_metrics(fcw_ctx* ctx, const double* targets, int S, const double* fracs, int q, const int* boundary,
                    int nb, const double* vertex_w, double eps, double health_h, double bin_width,
                    double cap_fraction, double rtol, int max_iter);

/* The per-source loop body of compute_all_wiring_costs (core_analysis.py:1550-1609) for nsrc sources,
 * processed in RHS blocks: distance -> health screen -> distance to boundary -> MSD -> radii,
 * perimeters, supplementary samples -> cleaned samples. Output rows follow `src`. */
int fcw_wiring_sources(fcw_ctx* ctx, const int* src, int nsrc, fcw_metrics_out* out);

/* As fcw_wiring_sources, with no device allocation in the call: for callers that cannot allocate
 * device memory while they run (tight serving loops, code that must not call hipMalloc). It has no
 * counterpart in the reference. The call's source positions, sweep plans and device outputs live
 * in a static device array of 16 MiB per GPU (held until the process ends); at most 4096 sources
 * per call, and fewer when the outputs per source are large (S scales, q samples: about
 * 1,100 bytes per source at S = 12, q = 3), else the call fails with a message saying so. The
 * context's own workspaces, created by fcw_create and fcw_set_metrics, are reused. Calls on one GPU
 * share its array and are serialised by a lock; profiling must be off. */
int fcw_wiring_sources_noalloc(fcw_ctx* ctx, const int* src, int nsrc, fcw_metrics_out* out);

/* Metrics of caller-supplied fields D (k x n, input vertex order): the batched counterpart of
 * lean_geometry.per_source_geometry (lean_geometry.py:302-340) when health_screen = 0, and of the
 * whole per-source loop body on given fields (core_analysis.py:1558-1609) when health_screen = 1.
 * src may be NULL (chord distances then use vertex 0). k <= block. */
int fcw_geometry_fields(fcw_ctx* ctx, const double* D, const int* src, int k, int health_screen,
                        fcw_metrics_out* out);

/* Accumulated seconds: [0] heat_solve, [1] fused_div, [2] poisson_solve, [3] metrics, [4] wall of the
 * last fcw_wiring_sources call, [5] host time that call spent issuing its blocks (the stage keys of
 * fastcw_lean.py:527-530 plus the metric stage). Stages 0-3 are filled only when profiling is on.
 * Returns the number of values written. */
int fcw_stage_times(fcw_ctx* ctx, double* out, int len);

/* Per-kernel timing with HIP events around every launch on the context's streams (0 = off).
 * fcw_kernel_times writes (total_ms, launches) per kernel kind, in the order of fcw_kernel_log's
 * kinds (FP64 forward sweep, FP64 backward sweep, metric kernel, FP32 forward, FP32 backward,
 * divergence, chunk statistics, residual): out[0..15]. */
int fcw_set_profiling(fcw_ctx* ctx, int on);
int fcw_kernel_times(fcw_ctx* ctx, double* out, int len);

/* Per-launch record of the timed kernels since profiling was last switched on, 6 doubles per launch:
 * kind (0 FP64 forward sweep, 1 FP64 backward sweep, 2 metric kernel, 3 FP32 forward sweep, 4 FP32
 * backward sweep, 5 divergence, 6 chunk statistics, 7 FP64 residual), system (0 heat, 1 Poisson,
 * -1 none), tree level (0 = leaves, -1 none), workgroups, Cholesky factor entries of the supernodes
 * the launch covers (sum of s(s+1)/2 + s*b; sweeps only), milliseconds. With out = NULL returns the
 * number of records held; else writes at most max_rows and returns how many. No reference
 * counterpart (the reference times whole stages, fastcw_lean.py:527-530). */
int fcw_kernel_log(fcw_ctx* ctx, double* out, int max_rows);

/* Library self-description for the backend self-check; replaces the binary inspection of
 * backend_check.verify_suitesparse (backend_check.py:32-73). */
int fcw_build_info(char* buf, int len);
const char* fcw_last_error(void);

/* One launch of a plain device copy of `bytes` bytes (reads `bytes`, writes `bytes`): the calibration
 * point for hardware byte counters when profiling the other kernels. No reference counterpart. */
int fcw_calibration_copy(int64_t bytes, int device);

/* Host-only check of the analysis and factorisation (no GPU): solves A X = B for
 * A = l_scale*L + m_scale*M assembled on the host from (V, F), with the CPU application of the
 * partitioned-inverse panels. B/X are n x k row-major in input vertex order. stats[0..3] =
 * nnodes, nlevels, factor_nnz, panel_elems. */
int fcw_host_factor_solve(const double* V, int n, const int* F, int m, double l_scale, double m_scale,
                          int leaf_size, const double* B, int k, double* X, double* stats);

/* Host-only: as fcw_host_intrinsic for any mesh, including non-manifold edges and neighbours of
 * inconsistent orientation, which potpourri3d's robust mode accepts (distance_engines.py:72-133;
 * core_analysis.py:609-636 lets such meshes through). Those meshes are flipped on their tufted cover
 * (each face once per side, glued around every edge in the cyclic order of its faces), whose faces
 * come back merged with their mirror copies. Buffers hold up to 2m faces: *mo = faces written,
 * faces_out mo x 3, len_out mo x 3, weight_out mo (1, or 1/2 for a cover face without its copy),
 * area_out mo (area x weight, may be NULL), gram_out mo x 6 (may be NULL). stats[0..4] = flips,
 * skipped flips, mollification delta, 1 if the cover was used, edges with three or more faces. */
int fcw_host_intrinsic_cover(const double* V, int n, const int* F, int m, double mollify, int* mo, int* faces_out,
                             double* len_out, double* weight_out, double* area_out, double* gram_out, double* stats);

/* Host-only view of the analysis (no GPU): the nested-dissection ordering and supernodes that
 * fcw_create builds, with the dissection's top `par_depth` levels split over threads (< 0: the
 * default rule, 2 on meshes of 20,000 vertices or more, else 0). The result does not depend on
 * par_depth. perm (n) = input vertex at each position; node_start / node_s / node_b (n each, the
 * first stats[0] entries written) = each supernode's first position, own and boundary row counts;
 * stats[0..3] = nnodes, nlevels, factor_nnz, panel_elems; panel_hash (may be NULL) = FNV-1a hash of
 * the FP64 panels of M + L. Replaces no reference interface: the ordering lives inside CHOLMOD in
 * the reference's solver (fastcw_lean.py:457-464). */
int fcw_host_analyse(const double* V, int n, const int* F, int m, int leaf_size, int par_depth, int* perm,
                     int* node_start, int* node_s, int* node_b, double* stats, uint64_t* panel_hash);

/* Mesh preparation on the device, the steps before the solves (core_analysis.py:651-683, 859-978,
 * 1336-1417, which build_cortex_submesh, boundary, face geometry, adjacency and
 * _compute_bfs_order_all do on the host): the cortical submesh (faces with all three vertices in
 * `mask` and area above 1e-12, vertices in ascending original index), face areas (zero below the
 * scale-aware threshold) and longest edges, vertex areas (each face's area / 3 per corner), the
 * boundary vertices (on an edge of one face), the sorted adjacency and the breadth-first order from
 * `start` (then from the smallest unvisited vertex), all bit-identical to the host computation.
 * The caller owns every buffer, sized for the input: sub_to_orig, orig_to_sub, boundary, order,
 * vertex_areas n; adj_ptr n + 1; vertices 3n; faces 3m; face_areas, face_L m; adj 6m. The counts
 * are returned in the struct. */
typedef struct {
    int n_sub, m_sub, n_boundary, n_adj;
    int* sub_to_orig;
    int* orig_to_sub;
    int* faces;
    double* vertices;
    int* boundary;
    int* order;
    double* face_areas;
    double* face_L;
    double* vertex_areas;
    int* adj_ptr;
    int* adj;
} fcw_mesh_prep_out;
int fcw_mesh_prep(const double* V, int n, const int* F, int m, const unsigned char* mask, int start, int device,
                  fcw_mesh_prep_out* out);

/* Host-only: the intrinsic Delaunay triangulation the robust operators are built on (the
 * triangulation inside potpourri3d's robust mode, distance_engines.py:119-125, which has no
 * callable counterpart). faces_out m x 3 (same vertices, flipped connectivity), len_out m x 3
 * (length of the edge opposite each corner), area_out m, gram_out m x 6 (grad B_a . grad B_b:
 * G00 G01 G02 G11 G12 G22). stats[0..2] = flips, skipped flips, mollification delta. Edge-manifold,
 * consistently oriented meshes only (else an error naming fcw_host_intrinsic_cover). */
int fcw_host_intrinsic(const double* V, int n, const int* F, int m, double mollify, int* faces_out, double* len_out,
                       double* area_out, double* gram_out, double* stats);

#ifdef __cplusplus
}
#endif

#endif /* FCW_H */
