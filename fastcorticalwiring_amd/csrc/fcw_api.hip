// C-ABI runtime of libfcw.so: context lifetime, device memory, and the
// per-block pipeline  delta RHS -> heat sweeps -> fused divergence ->
// Poisson sweeps -> shift/transpose -> metric kernel.
//
// A context owns the mesh, operators and both factor panel sets (shared), and
// `streams` workspaces, each with its own HIP stream and RHS-block buffers, so
// consecutive RHS blocks run concurrently: the narrow top-of-tree levels of one
// block overlap with the wide leaf levels and metric kernel of another.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fcw.h"
#include "fcw_device.h"
#include "fcw_internal.h"

namespace fcw {
thread_local GraphRecorder* t_recorder = nullptr;
}

namespace {

thread_local std::string g_err;

struct Fail : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIPCHK(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess)                                                                        \
            throw Fail(std::string(#x) + " failed: " + hipGetErrorString(e_) + " (" __FILE__ ":" + \
                       std::to_string(__LINE__) + ")");                                              \
    } while (0)

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t count, int64_t& tally) {
        if (p) {
            (void)hipFree(p);
            tally -= (int64_t)(n * sizeof(T));
            p = nullptr;
        }
        n = count;
        if (count) {
            HIPCHK(hipMalloc((void**)&p, count * sizeof(T)));
            tally += (int64_t)(count * sizeof(T));
        }
    }
    void upload(const T* h, size_t count, int64_t& tally) {
        alloc(count, tally);
        if (count) HIPCHK(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice));
    }
    // Per-call buffers: reallocated only to grow (a free can wait for every stream of the device,
    // other contexts' included), filled on `st`, which the caller synchronises or orders after.
    void refill(const T* h, size_t count, int64_t& tally, hipStream_t st) {
        if (count > n || !p) alloc(std::max<size_t>(count, 1), tally);
        if (count) HIPCHK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, st));
    }
};

struct EventPair {
    hipEvent_t a, b;
};

// Kernels timed per launch when profiling is on.
enum TimedKernel {
    kFwdGemm = 0,  // FP64 forward sweep launch
    kBwdGemm = 1,  // FP64 backward sweep launch
    kMetrics = 2,  // metric kernel
    kFwd32 = 3,    // FP32 forward sweep launch (refined solve)
    kBwd32 = 4,    // FP32 backward sweep launch
    kDivergence = 5,
    kStatsKernel = 6,  // chunk statistics
    kResidual = 7,  // FP64 residual of the refined solve
    kNumTimed = 8
};

// One timed launch: kind (TimedKernel), system (0 heat, 1 Poisson, -1 none), tree level (-1 none),
// workgroups, factor entries of the supernodes it covers (sweeps: sum of s(s+1)/2 + s*b), and, once
// collected, its duration.
struct LaunchRec {
    int kind, system, level, wgs;
    double entries;
    float ms;
};
constexpr size_t kMaxLaunchLog = 1 << 20;

// Active-node plan of one block's heat forward sweep: only supernodes on the paths from the
// block's sources to the root see a nonzero right-hand side.
struct BlockPlan {
    std::vector<int> tile_off;  // per level, into the call's tile array
    std::vector<int> node_off;  // per level, into the call's node array
    std::vector<double> level_entries;  // per level: factor entries of the active supernodes
    const fcw::LevelTile* tiles = nullptr;  // the call's tile and node arrays on the device
    const int* nodes = nullptr;
};

// Host-side plans of one call's blocks, with the tile and node arrays to upload.
struct CallPlans {
    std::vector<BlockPlan> blocks;
    std::vector<fcw::LevelTile> tiles;
    std::vector<int> nodes;
    void bind(const fcw::LevelTile* dtiles, const int* dnodes) {
        for (auto& b : blocks) {
            b.tiles = dtiles;
            b.nodes = dnodes;
        }
    }
};

// Device output arrays of one call (rows follow the call's sources).
struct OutBase {
    double *scal = nullptr, *radii = nullptr, *perim = nullptr, *er = nullptr, *ea = nullptr;
    float *sr = nullptr, *sa = nullptr;
};

}  // namespace

// Device scratch of the allocation-free wiring entry point (fcw_wiring_sources_noalloc): the
// call's source positions, sparse-sweep plans and device outputs. One copy per GPU, held until the
// process ends; host code reaches it only through the address hipGetSymbolAddress returns for the
// current device (fcw_create resolves it once per device), never through this name.
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
constexpr int kNoAllocMaxSources = 4096;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
struct ScratchSlot {
    unsigned char* ptr = nullptr;  // device address of `scratch` on this GPU
    std::mutex use;                // held by a call for as long as it uses the array
};
std::mutex g_scratch_init;
ScratchSlot g_scratch[kMaxDevices];

// The scratch array's address on `device`, which must be the current device.
ScratchSlot& scratch_slot(int device) {
    if (device < 0 || device >= kMaxDevices) throw Fail("device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_init);
    ScratchSlot& s = g_scratch[device];
    if (!s.ptr) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Fail("hipGetSymbolAddress returned no address for the scratch array");
        s.ptr = static_cast<unsigned char*>(p);
    }
    return s;
}

}  // namespace

// Per-stream RHS-block buffers. The heat pair (Yh, Uh) is all zero between calls: the sparse
// heat forward sweep writes only its active supernodes' rows and clears them afterwards, so
// inactive children always contribute zero. The Poisson pair (Yp, Up) is rewritten densely.
struct Workspace {
    hipStream_t stream = nullptr;
    DBuf<double> X, Yh, Yp, Z, Uh, Up, Dt, band;
    DBuf<double> part;  // partial products of the split-depth sweep launches
    DBuf<double> cpart;              // chunk_stats records (ldk x nch x kChunkStats, at chunk_rec_index)
    DBuf<float> X32, Y32, U32, Z32;  // refined solves: FP32 right-hand side and sweeps
    DBuf<float> ch_lo, ch_hi;        // chunk distance ranges (k x nch)
    DBuf<int> src_pos;
    std::vector<EventPair> ev_pool;
    std::vector<std::pair<LaunchRec, EventPair>> ev_pending;
    size_t ev_used = 0;
    // Captured launch sequences replayed for every full block on this stream (fcw_options.graphs):
    // the heat backward sweep, and divergence + Poisson solve. Their launches and arguments are the
    // same for every full block of this workspace.
    struct Captured {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    Captured g_heat_bwd, g_poisson;
};

struct fcw_ctx {
    fcw::Symbolic sym;
    fcw_options opt{};
    fcw_info info{};
    int n = 0, m = 0, mo = 0, ldk = 0;  // vertices, surface faces, operator faces, padded block width
    int64_t dev_bytes = 0;
    ScratchSlot* noalloc = nullptr;  // this device's scratch array (fcw_wiring_sources_noalloc)
    std::vector<std::unique_ptr<Workspace>> ws;
    std::vector<int> node_of_pos;

    // mesh and operators (positions)
    DBuf<double> V, area, gram, larea, g11, g12, g22, lval, mass;
    // robust mode: the operator triangulation (intrinsic Delaunay) beside the surface's faces
    bool robust = false;
    DBuf<int> op_faces;
    DBuf<double> op_area, op_gram;
    DBuf<int> faces, vf_ptr, vf, rowptr, colind, cptr, contrib, iperm;
    // divergence row groups
    int div_ngroups = 0;
    DBuf<int> div_grp, div_verts;
    DBuf<unsigned> div_loc;
    DBuf<double> div_geo;
    // refined forms: the residual's matrix values (L + shift M) and CSR entries' local indices in
    // the divergence groups, when every row's columns lie in its group (res_grouped)
    DBuf<unsigned char> res_cloc;
    DBuf<double> res_aval;
    bool res_grouped = false;
    // tree
    DBuf<int> node_start, node_s, node_b, bidx, cmap, child_ptr, child_list, level_nodes, cown, ctail;
    DBuf<int64_t> node_boff, node_poff, node_uoff;
    DBuf<fcw::LevelTile> fwd_tiles, bwd_tiles;
    std::vector<int> fwd_tile_off, bwd_tile_off, level_off;
    // depth parts per level of the dense FP64 sweeps (1: not split)
    std::vector<int> fwd_split, bwd_split;
    // row tile height per level of the forward and backward sweeps (64, 32 or 16)
    std::vector<int> fwd_tm, bwd_tm;
    bool fwd_idx_lds = true, bwd_idx_lds = true;  // every node's sweep indices fit in LDS
    DBuf<double> panel_heat, panel_pois;
    DBuf<float> panel_pois32;  // FCW_FP32_IR: the Poisson panels rounded to FP32
    DBuf<uint32_t> panel_pois_bf;  // FCW_BF16X3_IR: the Poisson panels split into bf16 hi | lo << 16
    DBuf<double> Dout;
    // per-call plans
    DBuf<fcw::LevelTile> plan_tiles;
    DBuf<int> plan_nodes;
    DBuf<int> call_pos;  // the current call's source positions
    // 256-face chunks: vertices (owned first, cnown of them), area
    int nch = 0;
    DBuf<int> cvptr, cverts, cnown, isb;
    DBuf<double> chunk_area;
    // metrics
    bool metrics_ready = false;
    int S = 0, q = 0, nb = 0, band_cap = 6144;  // scratch: S * band_cap band faces per source
    double eps = 1e-6, health_h = 0.0, width = 0.0, cap = -1.0, rtol = 1e-10;
    int max_iter = 60;
    DBuf<double> targets, fracs, vw;
    DBuf<int> boundary;
    DBuf<double> o_scal, o_radii, o_perim, o_er, o_ea;
    DBuf<float> o_sr, o_sa;
    size_t o_cap = 0;
    // timing
    double stage[6] = {0, 0, 0, 0, 0, 0};
    bool profiling = false;
    bool no_new_graphs = false;  // set during fcw_wiring_sources_noalloc: no capture may be created
    double ktime[kNumTimed] = {};
    int64_t kcount[kNumTimed] = {};
    std::vector<LaunchRec> launch_log;  // every timed launch while profiling (at most kMaxLaunchLog)
    std::vector<double> level_entries;   // per tree level: sum over its supernodes of s(s+1)/2 + s*b

    ~fcw_ctx() {
        for (auto& w : ws) {
            if (w->stream) (void)hipStreamSynchronize(w->stream);
            for (Workspace::Captured* g : {&w->g_heat_bwd, &w->g_poisson}) {
                if (g->exec) (void)hipGraphExecDestroy(g->exec);
                if (g->graph) (void)hipGraphDestroy(g->graph);
            }
            for (auto& e : w->ev_pool) {
                (void)hipEventDestroy(e.a);
                (void)hipEventDestroy(e.b);
            }
            if (w->stream) (void)hipStreamDestroy(w->stream);
        }
    }

    fcw::MeshDev mesh_dev() const {
        fcw::MeshDev md;
        md.n = n;
        md.m = m;
        md.V = V.p;
        md.faces = faces.p;
        md.larea = larea.p;
        md.g11 = g11.p;
        md.g12 = g12.p;
        md.g22 = g22.p;
        md.mo = mo;
        md.op_faces = robust ? op_faces.p : faces.p;
        md.op_area = robust ? op_area.p : area.p;
        md.op_gram = robust ? op_gram.p : gram.p;
        md.op_vf_ptr = vf_ptr.p;
        md.op_vf = vf.p;
        return md;
    }

    fcw::DivDev div_dev() const {
        fcw::DivDev d;
        d.ngroups = div_ngroups;
        d.grp = reinterpret_cast<const int4*>(div_grp.p);
        d.verts = div_verts.p;
        d.loc = div_loc.p;
        d.geo = div_geo.p;
        d.vf_ptr = vf_ptr.p;
        d.vf = vf.p;
        d.faces = robust ? op_faces.p : faces.p;
        return d;
    }

    EventPair* ev_begin(Workspace& w, int which, int system = -1, int level = -1, int wgs = 0, double entries = 0.0) {
        if (!profiling) return nullptr;
        if (w.ev_used == w.ev_pool.size()) {
            EventPair e;
            HIPCHK(hipEventCreate(&e.a));
            HIPCHK(hipEventCreate(&e.b));
            w.ev_pool.push_back(e);
        }
        EventPair* e = &w.ev_pool[w.ev_used++];
        HIPCHK(hipEventRecord(e->a, w.stream));
        w.ev_pending.push_back({LaunchRec{which, system, level, wgs, entries, 0.f}, *e});
        return e;
    }
    void ev_end(Workspace& w, EventPair* e) {
        if (e) HIPCHK(hipEventRecord(e->b, w.stream));
    }
    void ev_collect(Workspace& w) {
        if (w.ev_pending.empty()) return;
        HIPCHK(hipStreamSynchronize(w.stream));
        for (auto& pe : w.ev_pending) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, pe.second.a, pe.second.b));
            ktime[pe.first.kind] += ms;
            kcount[pe.first.kind] += 1;
            if (launch_log.size() < kMaxLaunchLog) {
                pe.first.ms = ms;
                launch_log.push_back(pe.first);
            }
        }
        w.ev_pending.clear();
        w.ev_used = 0;
    }

    void alloc_workspace(Workspace& w) {
        const size_t ld = (size_t)ldk;
        HIPCHK(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking));
        // update rows plus one row that stays zero: a missing child contribution reads it
        const size_t urows = (size_t)sym.urows + 1;
        w.X.alloc((size_t)n * ld, dev_bytes);
        w.Yh.alloc((size_t)n * ld, dev_bytes);
        w.Yp.alloc((size_t)n * ld, dev_bytes);
        w.Z.alloc((size_t)n * ld, dev_bytes);
        w.Uh.alloc(urows * ld, dev_bytes);
        w.Up.alloc(urows * ld, dev_bytes);
        w.Dt.alloc((size_t)n * ld, dev_bytes);
        w.src_pos.alloc(ld, dev_bytes);
        if (opt.precision != FCW_FP64) {
            w.X32.alloc((size_t)n * ld, dev_bytes);
            w.Y32.alloc((size_t)n * ld, dev_bytes);
            w.Z32.alloc((size_t)n * ld, dev_bytes);
            w.U32.alloc(urows * ld, dev_bytes);
            HIPCHK(hipMemsetAsync(w.U32.p + (size_t)sym.urows * ld, 0, ld * sizeof(float), w.stream));
        }
        fcw::launch_fill(w.stream, w.Yh.p, (int64_t)n * ldk, 0.0);
        fcw::launch_fill(w.stream, w.Uh.p, (int64_t)urows * ldk, 0.0);
        fcw::launch_fill(w.stream, w.Up.p + (size_t)sym.urows * ld, ldk, 0.0);
        HIPCHK(hipStreamSynchronize(w.stream));
    }

    // One system: forward then backward sweep, X (RHS) -> Z (solution) through Y/U. With a plan the
    // forward sweep visits only the plan's active supernodes (Y/U must then be the heat pair, zero on
    // every other row) and their rows are cleared again after the backward sweep.
    // Cholesky factor entries of one supernode: its triangle and its boundary rows.
    double node_entries(int id) const {
        const double s_ = sym.node_s[id], b_ = sym.node_b[id];
        return s_ * (s_ + 1.0) / 2.0 + s_ * b_;
    }

    // Workgroups of a sweep launch over ntiles row tiles of height tm (the kernel's grid, one depth part).
    int sweep_wgs(int ntiles, int tm) const { return (ntiles + 7) / 8 * 8 * (ldk / (4096 / tm)); }

    void solve(Workspace& w, const double* panel, const BlockPlan* plan, double* Y, double* U, double* Dt = nullptr,
               int system = 0) {
        solve_fwd(w, panel, plan, Y, U, system);
        solve_bwd(w, panel, Y, Dt, system);
        if (plan) clear_plan_rows(w, plan, Y, U);
    }

    void clear_plan_rows(Workspace& w, const BlockPlan* plan, double* Y, double* U) {
        const int L = sym.nlevels;
        fcw::launch_clear_rows(w.stream, plan->nodes + plan->node_off[0], plan->node_off[L] - plan->node_off[0], ldk,
                               node_start.p, node_s.p, node_b.p, node_uoff.p, Y, U);
    }

    void solve_fwd(Workspace& w, const double* panel, const BlockPlan* plan, double* Y, double* U, int system) {
        const int L = sym.nlevels;
        for (int l = 0; l < L; ++l) {
            const fcw::LevelTile* tiles;
            int ntiles;
            double ent;
            if (plan) {
                tiles = plan->tiles + plan->tile_off[l];
                ntiles = plan->tile_off[l + 1] - plan->tile_off[l];
                ent = plan->level_entries[l];
            } else {
                tiles = fwd_tiles.p + fwd_tile_off[l];
                ntiles = fwd_tile_off[l + 1] - fwd_tile_off[l];
                ent = level_entries[l];
            }
            EventPair* e = ntiles > 0 ? ev_begin(w, kFwdGemm, system, l, sweep_wgs(ntiles, fwd_tm[l]), ent) : nullptr;
            fcw::launch_fwd_gemm(w.stream, tiles, ntiles, ldk, fwd_tm[l], node_start.p, node_s.p, node_b.p, node_poff.p,
                                 node_uoff.p, cown.p, ctail.p, panel, w.X.p, Y, U, fwd_idx_lds, l == 0,
                                 (int)sym.urows, plan ? 1 : fwd_split[l], w.part.p);
            ev_end(w, e);
        }
    }

    void solve_bwd(Workspace& w, const double* panel, double* Y, double* Dt, int system) {
        const int L = sym.nlevels;
        for (int l = L - 1; l >= 0; --l) {
            EventPair* e = ev_begin(w, kBwdGemm, system, l, sweep_wgs(bwd_tile_off[l + 1] - bwd_tile_off[l], bwd_tm[l]),
                                    level_entries[l]);
            fcw::launch_bwd_gemm(w.stream, bwd_tiles.p + bwd_tile_off[l], bwd_tile_off[l + 1] - bwd_tile_off[l], ldk,
                                 bwd_tm[l],
                                 node_start.p, node_s.p, node_b.p, node_boff.p, bidx.p, node_poff.p, panel, Y, w.Z.p,
                                 bwd_idx_lds, Dt, n, bwd_split[l], w.part.p);
            ev_end(w, e);
        }
    }

    // Run `issue` on w's stream, or, when graphs are on for this call, launch its graph (built the
    // first time: the launches `issue` makes are recorded as a chain of kernel nodes instead of
    // being started, so no stream is captured).
    template <class Fn>
    void run_captured(Workspace& w, Workspace::Captured& g, bool graphable, Fn&& issue) {
        if (!graphable) {
            issue();
            return;
        }
        if (!g.exec) {
            hipGraph_t graph = nullptr;
            HIPCHK(hipGraphCreate(&graph, 0));
            fcw::GraphRecorder rec;
            rec.graph = graph;
            fcw::t_recorder = &rec;
            try {
                issue();
            } catch (...) {
                fcw::t_recorder = nullptr;
                (void)hipGraphDestroy(graph);
                throw;
            }
            fcw::t_recorder = nullptr;
            g.graph = graph;
            HIPCHK(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
        }
        HIPCHK(hipGraphLaunch(g.exec, w.stream));
    }

    // FP32 sweeps of one system: w.X32 -> w.Z32, the solution also stored / added into acc64 (and
    // written source-major into Dt when set).
    template <typename PT>
    void solve32(Workspace& w, const PT* panel, double* acc64, int acc_mode, double* Dt = nullptr) {
        const int L = sym.nlevels;
        for (int l = 0; l < L; ++l) {
            EventPair* e = ev_begin(w, kFwd32, 1, l, sweep_wgs(fwd_tile_off[l + 1] - fwd_tile_off[l], fwd_tm[l]),
                                    level_entries[l]);
            fcw::launch_fwd_gemm(w.stream, fwd_tiles.p + fwd_tile_off[l], fwd_tile_off[l + 1] - fwd_tile_off[l], ldk,
                                 fwd_tm[l], node_start.p, node_s.p, node_b.p, node_poff.p, node_uoff.p, cown.p, ctail.p, panel,
                                 w.X32.p, w.Y32.p, w.U32.p, fwd_idx_lds, l == 0, (int)sym.urows);
            ev_end(w, e);
        }
        for (int l = L - 1; l >= 0; --l) {
            EventPair* e = ev_begin(w, kBwd32, 1, l, sweep_wgs(bwd_tile_off[l + 1] - bwd_tile_off[l], bwd_tm[l]),
                                    level_entries[l]);
            fcw::launch_bwd_gemm(w.stream, bwd_tiles.p + bwd_tile_off[l], bwd_tile_off[l + 1] - bwd_tile_off[l], ldk,
                                 bwd_tm[l], node_start.p, node_s.p, node_b.p, node_boff.p, bidx.p, node_poff.p, panel, w.Y32.p,
                                 w.Z32.p, bwd_idx_lds, acc64, acc_mode | (l == 0 ? 4 : 0), Dt, n);
            ev_end(w, e);
        }
    }

    // Poisson system by refinement: an FP32 solve of the right-hand side, then ir_steps times the
    // FP64 residual r = b - A x rounded to FP32 and an FP32 solve of A dx = r added to x (FP64).
    void poisson_refined(Workspace& w) {
        if (opt.precision == FCW_BF16X3_IR)
            poisson_refined_with(w, panel_pois_bf.p);
        else
            poisson_refined_with(w, panel_pois32.p);
    }

    // The first FP32 right-hand side, w.X32 = w.X rounded to FP32, is written by the divergence.
    template <typename PT>
    void poisson_refined_with(Workspace& w, const PT* panel) {
        solve32(w, panel, w.Z.p, 1, opt.ir_steps == 0 ? w.Dt.p : nullptr);
        for (int it = 0; it < opt.ir_steps; ++it) {
            EventPair* e = ev_begin(w, kResidual, 1);
            if (res_grouped)
                fcw::launch_residual_grouped(w.stream, div_dev(), ldk, rowptr.p, colind.p, res_cloc.p, res_aval.p,
                                             w.X.p, w.Z.p, w.X32.p);
            else
                fcw::launch_residual32(w.stream, n, ldk, rowptr.p, colind.p, lval.p, mass.p, opt.poisson_shift,
                                       w.X.p, w.Z.p, w.X32.p);
            ev_end(w, e);
            solve32(w, panel, w.Z.p, 2, it + 1 == opt.ir_steps ? w.Dt.p : nullptr);
        }
    }

    // Distances for the k sources whose positions are in w.src_pos[0:k]: the Poisson solution in w.Z
    // (n x ldk) and source-major in w.Dt (ldk x n), both unshifted; source j's distances are its
    // column minus its value at its own position.
    void distances(Workspace& w, int k, const BlockPlan* plan) {
        // full blocks of the default configuration replay the captured sequences
        const bool graphable = opt.graphs > 0 && !profiling && !no_new_graphs && plan && k == ldk;
        hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
        if (profiling) {
            for (auto& e : ev) HIPCHK(hipEventCreate(&e));
            HIPCHK(hipEventRecord(ev[0], w.stream));
        }
        if (plan) {
            // the sparse forward sweep reads only the own rows of the plan's supernodes: those rows
            // of X are cleared (X holds the previous block's divergence), not the whole block
            const int L = sym.nlevels;
            fcw::launch_clear_rows(w.stream, plan->nodes + plan->node_off[0], plan->node_off[L] - plan->node_off[0],
                                   ldk, node_start.p, node_s.p, node_b.p, node_uoff.p, w.X.p, nullptr);
            fcw::launch_set_delta(w.stream, w.X.p, ldk, w.src_pos.p, k);
            solve_fwd(w, panel_heat.p, plan, w.Yh.p, w.Uh.p, 0);
            run_captured(w, w.g_heat_bwd, graphable, [&] { solve_bwd(w, panel_heat.p, w.Yh.p, nullptr, 0); });
            clear_plan_rows(w, plan, w.Yh.p, w.Uh.p);
        } else {
            fcw::launch_delta_rhs(w.stream, w.X.p, n, ldk, w.src_pos.p, k);
            solve(w, panel_heat.p, nullptr, w.Yp.p, w.Up.p, nullptr, 0);
        }
        if (profiling) HIPCHK(hipEventRecord(ev[1], w.stream));
        auto divergence = [&] {
            EventPair* ed = ev_begin(w, kDivergence);
            // the refined forms' first FP32 right-hand side is the divergence rounded to FP32
            fcw::launch_divergence(w.stream, div_dev(), ldk, w.Z.p, w.X.p, opt.precision != FCW_FP64 ? w.X32.p : nullptr);
            ev_end(w, ed);
        };
        auto poisson = [&] {
            if (opt.precision != FCW_FP64)
                poisson_refined(w);
            else
                solve(w, panel_pois.p, nullptr, w.Yp.p, w.Up.p, w.Dt.p, 1);
        };
        if (graphable) {
            run_captured(w, w.g_poisson, true, [&] {
                divergence();
                poisson();
            });
        } else {
            divergence();
            if (profiling) HIPCHK(hipEventRecord(ev[2], w.stream));
            poisson();
        }
        if (profiling) {
            HIPCHK(hipEventRecord(ev[3], w.stream));
            HIPCHK(hipEventSynchronize(ev[3]));
            for (int i = 0; i < 3; ++i) {
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
                stage[i] += ms * 1e-3;
            }
            for (auto& e : ev) (void)hipEventDestroy(e);
        }
    }

    // Host plans for consecutive blocks of `pos` (source positions); the caller uploads the tile and
    // node arrays and binds their device copies.
    CallPlans plan_blocks(const std::vector<int>& pos) {
        const int L = sym.nlevels, nn = sym.nnodes;
        const int nblocks = ((int)pos.size() + ldk - 1) / ldk;
        CallPlans cp;
        std::vector<BlockPlan>& plans = cp.blocks;
        plans.resize(nblocks);
        std::vector<fcw::LevelTile>& tiles = cp.tiles;
        std::vector<int>& nodes = cp.nodes;
        std::vector<int> mark(nn, -1);
        std::vector<std::vector<int>> per_level(L);
        for (int b = 0; b < nblocks; ++b) {
            for (auto& v : per_level) v.clear();
            const int j1 = std::min((int)pos.size(), (b + 1) * ldk);
            for (int j = b * ldk; j < j1; ++j) {
                for (int id = node_of_pos[pos[j]]; id >= 0 && mark[id] != b; id = sym.node_parent[id]) {
                    mark[id] = b;
                    per_level[sym.node_level[id]].push_back(id);
                }
            }
            BlockPlan& P = plans[b];
            P.tile_off.push_back((int)tiles.size());
            P.node_off.push_back((int)nodes.size());
            for (int l = 0; l < L; ++l) {
                std::sort(per_level[l].begin(), per_level[l].end());
                double ent = 0.0;
                for (int id : per_level[l]) {
                    nodes.push_back(id);
                    const int f = sym.node_s[id] + sym.node_b[id];
                    for (int m0 = 0; m0 < f; m0 += fwd_tm[l]) tiles.push_back({id, m0});
                    ent += node_entries(id);
                }
                P.level_entries.push_back(ent);
                P.tile_off.push_back((int)tiles.size());
                P.node_off.push_back((int)nodes.size());
            }
        }
        return cp;
    }

    // Plans (and, with `dpos`, the source positions) uploaded to the context's own per-call arrays;
    // the copies are complete when this returns.
    CallPlans make_plans(const std::vector<int>& pos, const int** dpos = nullptr) {
        CallPlans cp = plan_blocks(pos);
        hipStream_t st = ws[0]->stream;
        plan_tiles.refill(cp.tiles.data(), cp.tiles.size(), dev_bytes, st);
        plan_nodes.refill(cp.nodes.data(), cp.nodes.size(), dev_bytes, st);
        if (dpos) {
            call_pos.refill(pos.data(), pos.size(), dev_bytes, st);
            *dpos = call_pos.p;
        }
        HIPCHK(hipStreamSynchronize(st));
        cp.bind(plan_tiles.p, plan_nodes.p);
        return cp;
    }

    void ensure_outputs(size_t k) {
        if (k <= o_cap) return;
        const size_t NE = S > 0 ? (size_t)(S - 1) * q : 0;
        o_scal.alloc(k * fcw::kSlotCount, dev_bytes);
        o_radii.alloc(k * S, dev_bytes);
        o_perim.alloc(k * S, dev_bytes);
        o_er.alloc(std::max<size_t>(1, k * NE), dev_bytes);
        o_ea.alloc(std::max<size_t>(1, k * NE), dev_bytes);
        o_sr.alloc(k * (S + NE), dev_bytes);
        o_sa.alloc(k * (S + NE), dev_bytes);
        o_cap = k;
    }

    fcw::MetricsDev metrics_dev(Workspace& w) {
        fcw::MetricsDev md;
        md.S = S;
        md.q = q;
        md.nb = nb;
        md.targets = targets.p;
        md.fracs = fracs.p;
        md.boundary = boundary.p;
        md.vw = vw.p;
        md.eps = eps;
        md.health_h = health_h;
        md.width = width;
        md.cap = cap;
        md.rtol = rtol;
        md.max_iter = max_iter;
        md.pool_cap = S * band_cap;
        md.band = w.band.p;
        md.nch = nch;
        md.ch_lo = nullptr;  // set by the wiring path after chunk_stats
        md.ch_hi = nullptr;
        md.part = nullptr;
        md.chunk_area = chunk_area.p;
        md.shift_src = nullptr;  // set by the wiring path (fields stored unshifted)
        md.shift_ld = ldk;
        return md;
    }

    OutBase own_outputs() const {
        OutBase b;
        b.scal = o_scal.p;
        b.radii = o_radii.p;
        b.perim = o_perim.p;
        b.er = o_er.p;
        b.ea = o_ea.p;
        b.sr = o_sr.p;
        b.sa = o_sa.p;
        return b;
    }

    // Device bytes of the outputs of k sources, laid out as carve_outputs places them.
    size_t output_bytes(size_t k) const {
        const size_t NE = S > 0 ? (size_t)(S - 1) * q : 0;
        auto up = [](size_t b) { return (b + 255) / 256 * 256; };
        return up(k * fcw::kSlotCount * 8) + 2 * up(k * S * 8) + 2 * up(std::max<size_t>(1, k * NE) * 8) +
               2 * up(k * (S + NE) * 4);
    }
    OutBase carve_outputs(unsigned char* base, size_t k) const {
        const size_t NE = S > 0 ? (size_t)(S - 1) * q : 0;
        auto up = [](size_t b) { return (b + 255) / 256 * 256; };
        OutBase b;
        size_t off = 0;
        auto take = [&](size_t bytes) {
            unsigned char* p = base + off;
            off += up(bytes);
            return p;
        };
        b.scal = reinterpret_cast<double*>(take(k * fcw::kSlotCount * 8));
        b.radii = reinterpret_cast<double*>(take(k * S * 8));
        b.perim = reinterpret_cast<double*>(take(k * S * 8));
        b.er = reinterpret_cast<double*>(take(std::max<size_t>(1, k * NE) * 8));
        b.ea = reinterpret_cast<double*>(take(std::max<size_t>(1, k * NE) * 8));
        b.sr = reinterpret_cast<float*>(take(k * (S + NE) * 4));
        b.sa = reinterpret_cast<float*>(take(k * (S + NE) * 4));
        return b;
    }

    fcw::MetricsOut outputs_at(const OutBase& b, size_t row) const {
        const size_t NE = S > 0 ? (size_t)(S - 1) * q : 0;
        fcw::MetricsOut o;
        o.scal = b.scal + row * fcw::kSlotCount;
        o.radii = b.radii + row * S;
        o.perim = b.perim + row * S;
        o.extra_r = b.er + row * NE;
        o.extra_a = b.ea + row * NE;
        o.samp_r = b.sr + row * (S + NE);
        o.samp_a = b.sa + row * (S + NE);
        return o;
    }

    void sync_all() {
        for (auto& w : ws) HIPCHK(hipStreamSynchronize(w->stream));
    }

    void copy_outputs(const OutBase& b, size_t k, fcw_metrics_out* out) {
        sync_all();
        const size_t NE = S > 0 ? (size_t)(S - 1) * q : 0;
        // copies ordered on the context's own stream (not the legacy stream)
        hipStream_t st = ws[0]->stream;
        auto cp = [&](void* dst, const void* src, size_t bytes) {
            if (dst && bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
        };
        cp(out->scalars, b.scal, k * fcw::kSlotCount * sizeof(double));
        cp(out->radii, b.radii, k * S * sizeof(double));
        cp(out->perimeters, b.perim, k * S * sizeof(double));
        cp(out->extra_r, b.er, k * NE * sizeof(double));
        cp(out->extra_a, b.ea, k * NE * sizeof(double));
        cp(out->sample_r, b.sr, k * (S + NE) * sizeof(float));
        cp(out->sample_a, b.sa, k * (S + NE) * sizeof(float));
        HIPCHK(hipStreamSynchronize(st));
    }

    // The blocks of one wiring call: source positions `dpos` (device, nsrc), plans bound to device
    // arrays, outputs at `ob`. Issues every block on the workspaces' streams; no allocation here
    // (profiling events aside, which only the timing mode creates).
    void run_wiring(const int* dpos, int nsrc, const CallPlans& plans, const OutBase& ob) {
        const fcw::MeshDev mdv = mesh_dev();
        const int NS = (int)ws.size();
        for (int b0 = 0, bi = 0; b0 < nsrc; b0 += ldk, ++bi) {
            Workspace& w = *ws[bi % NS];
            const int k = std::min(ldk, nsrc - b0);
            HIPCHK(hipMemcpyAsync(w.src_pos.p, dpos + b0, sizeof(int) * k, hipMemcpyDeviceToDevice, w.stream));
            distances(w, k, opt.sparse_rhs ? &plans.blocks[bi] : nullptr);
            hipEvent_t ma = nullptr, mb = nullptr;
            if (profiling) {
                HIPCHK(hipEventCreate(&ma));
                HIPCHK(hipEventCreate(&mb));
                HIPCHK(hipEventRecord(ma, w.stream));
            }
            EventPair* ec = ev_begin(w, kStatsKernel);
            fcw::launch_chunk_stats(w.stream, k, ldk, nch, w.Z.p, w.src_pos.p, V.p, vw.p, isb.p, cvptr.p, cverts.p,
                                    cnown.p, eps, health_h, w.ch_lo.p, w.ch_hi.p, w.cpart.p);
            ev_end(w, ec);
            EventPair* e = ev_begin(w, kMetrics);
            fcw::MetricsDev md = metrics_dev(w);
            md.ch_lo = w.ch_lo.p;
            md.ch_hi = w.ch_hi.p;
            md.part = w.cpart.p;
            md.shift_src = w.Z.p;
            fcw::launch_metrics(w.stream, k, mdv, md, w.Dt.p, w.src_pos.p, outputs_at(ob, (size_t)b0), true);
            ev_end(w, e);
            HIPCHK(hipGetLastError());
            if (profiling) {
                HIPCHK(hipEventRecord(mb, w.stream));
                HIPCHK(hipEventSynchronize(mb));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, ma, mb));
                stage[3] += ms * 1e-3;
                (void)hipEventDestroy(ma);
                (void)hipEventDestroy(mb);
                ev_collect(w);
            }
        }
    }
};

namespace {

// Row tile height of one level's sweep launch: among 64, 32 and 16 (16: backward only) the height
// the block width allows with the least modelled cost: the 16-deep depth steps every tile walks
// (its depth range after the triangular skip), counted as the matrix-core work they execute (the
// range rounded up to 4) plus half the per-thread staging loads of each step relative to the 8 of
// a 64-row backward tile. Short tiles waste fewer rows on nodes with few own rows; tall tiles stage
// fewer right-hand-side values per result. forced > 0 picks the smallest allowed height >= forced.
int choose_tile_rows(const fcw::Symbolic& S, int l, bool forward, int ldk, int forced) {
    static const int cand[3] = {16, 32, 64};
    if (forced > 0) {
        for (int tm : cand)
            if (tm >= forced && fcw::sweep_tile_ok(forward, tm, ldk)) return tm;
        return 64;
    }
    int pick = 64;
    double best = 0.0;
    bool have = false;
    for (int tm : cand) {
        if (!fcw::sweep_tile_ok(forward, tm, ldk)) continue;
        const int tn = 4096 / tm;
        const bool kids = forward && l > 0;  // forward non-leaf steps also stage two update rows
        const double loads = (tm * fcw::kTileK + (kids ? 3 : 1) * tn * fcw::kTileK) / 256.0;
        double cost = 0.0;
        for (int i = S.level_ptr[l]; i < S.level_ptr[l + 1]; ++i) {
            const int id = S.level_nodes[i];
            const int sn = S.node_s[id], f = sn + S.node_b[id];
            const int rows = forward ? f : sn, K = forward ? sn : f;
            for (int m0 = 0; m0 < rows; m0 += tm) {
                int kb = 0, ke = K;
                if (!forward)
                    kb = m0 / fcw::kTileK * fcw::kTileK;
                else if (m0 + tm <= sn)
                    ke = std::min(ke, m0 + tm);
                if (ke <= kb) continue;
                const double depth4 = (ke - kb + 3) / 4 * 4, steps = (ke - kb + fcw::kTileK - 1) / fcw::kTileK;
                cost += (double)(ldk / tn) * (depth4 / fcw::kTileK + 0.5 * steps * loads / 8.0);
            }
        }
        if (!have || cost < best) {
            best = cost;
            pick = tm;
            have = true;
        }
    }
    return pick;
}

void host_operators(int n, const double* V, int m, const int* F, std::vector<std::vector<std::pair<int, double>>>& rows,
                    std::vector<double>& mass) {
    rows.assign(n, {});
    mass.assign(n, 0.0);
    for (int f = 0; f < m; ++f) {
        const int id[3] = {F[3 * f], F[3 * f + 1], F[3 * f + 2]};
        double p0[3], e1[3], e2[3];
        for (int c = 0; c < 3; ++c) {
            p0[c] = V[3 * id[0] + c];
            e1[c] = V[3 * id[1] + c] - p0[c];
            e2[c] = V[3 * id[2] + c] - p0[c];
        }
        double nr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const double nl = std::sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
        if (!(nl > 0.0)) throw Fail("Zero-area faces must be removed before building the heat engine.");
        const double A = 0.5 * nl, inv2a = 1.0 / (2.0 * A);
        double nh[3] = {nr[0] / nl, nr[1] / nl, nr[2] / nl};
        double g[3][3];
        g[1][0] = (e2[1] * nh[2] - e2[2] * nh[1]) * inv2a;
        g[1][1] = (e2[2] * nh[0] - e2[0] * nh[2]) * inv2a;
        g[1][2] = (e2[0] * nh[1] - e2[1] * nh[0]) * inv2a;
        g[2][0] = (nh[1] * e1[2] - nh[2] * e1[1]) * inv2a;
        g[2][1] = (nh[2] * e1[0] - nh[0] * e1[2]) * inv2a;
        g[2][2] = (nh[0] * e1[1] - nh[1] * e1[0]) * inv2a;
        for (int c = 0; c < 3; ++c) g[0][c] = -g[1][c] - g[2][c];
        for (int a = 0; a < 3; ++a) {
            mass[id[a]] += A / 3.0;
            for (int b = 0; b < 3; ++b)
                rows[id[a]].push_back({id[b], A * (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2])});
        }
    }
}

double mean_edge_length(const fcw::Symbolic& S, const double* Vpos) {
    double sum = 0.0;
    int64_t cnt = 0;
    for (int p = 0; p < S.n; ++p)
        for (int e = S.rowptr[p]; e < S.rowptr[p + 1]; ++e) {
            const int q = S.colind[e];
            if (q <= p) continue;
            const double dx = Vpos[3 * q] - Vpos[3 * p], dy = Vpos[3 * q + 1] - Vpos[3 * p + 1],
                         dz = Vpos[3 * q + 2] - Vpos[3 * p + 2];
            sum += std::sqrt(dx * dx + dy * dy + dz * dz);
            ++cnt;
        }
    return cnt ? sum / (double)cnt : 0.0;
}

// Row groups of the divergence: consecutive rows added while the group's incidences touch at most
// kDivVerts distinct vertices and number at most kDivInc, and it has at most kDivGroupRows rows.
// grp holds four ints per group (first row, first incidence, first entry in verts, distinct count;
// -1 for a lone row beyond the limits) and a closing entry (n, total incidences, |verts|, 0);
// loc[e] packs the local indices of incidence e's three face vertices, 6 bits each. With a CSR
// pattern (rowptr, colind), cloc[e] is the local index of CSR entry e's column in its row's group;
// returns false if some column is not among its group's vertices (then cloc is not usable).
bool build_divergence_groups(int n, const std::vector<int>& vf_ptr, const std::vector<int>& vf,
                             const std::vector<int>& faces, std::vector<int>& grp, std::vector<int>& verts,
                             std::vector<unsigned>& loc, const std::vector<int>* rowptr = nullptr,
                             const std::vector<int>* colind = nullptr, std::vector<unsigned char>* cloc = nullptr) {
    bool cloc_ok = true;
    std::vector<int> mark(n, -1), seen(n, -1), lidx(n, 0), cur;
    int r = 0, gid = 0, attempt = 0;  // a row that does not fit is tried again as the next group's first
    while (r < n) {
        const int r0 = r, e0 = vf_ptr[r0];
        cur.clear();
        while (r < n && r - r0 < fcw::kDivGroupRows) {
            int add = 0;
            ++attempt;
            for (int e = vf_ptr[r]; e < vf_ptr[r + 1]; ++e)
                for (int c = 0; c < 3; ++c) {
                    const int p = faces[3 * (size_t)(vf[e] / 3) + c];
                    if (mark[p] != gid && seen[p] != attempt) {
                        seen[p] = attempt;
                        ++add;
                    }
                }
            const bool fits = (int)cur.size() + add <= fcw::kDivVerts && vf_ptr[r + 1] - e0 <= fcw::kDivInc;
            if (!fits) break;
            for (int e = vf_ptr[r]; e < vf_ptr[r + 1]; ++e)
                for (int c = 0; c < 3; ++c) {
                    const int p = faces[3 * (size_t)(vf[e] / 3) + c];
                    if (mark[p] != gid) {
                        mark[p] = gid;
                        cur.push_back(p);
                    }
                }
            ++r;
        }
        if (r == r0) {  // this row alone exceeds the limits
            grp.insert(grp.end(), {r0, e0, (int)verts.size(), -1});
            ++r;
            ++gid;
            continue;
        }
        std::sort(cur.begin(), cur.end());
        for (size_t i = 0; i < cur.size(); ++i) lidx[cur[i]] = (int)i;
        for (int e = e0; e < vf_ptr[r]; ++e) {
            const int* fv = faces.data() + 3 * (size_t)(vf[e] / 3);
            loc[e] = (unsigned)lidx[fv[0]] | ((unsigned)lidx[fv[1]] << 6) | ((unsigned)lidx[fv[2]] << 12);
        }
        if (cloc)
            for (int q = r0; q < r; ++q)
                for (int e = (*rowptr)[q]; e < (*rowptr)[q + 1]; ++e) {
                    const int c = (*colind)[e];
                    if (mark[c] != gid) cloc_ok = false;
                    (*cloc)[e] = (unsigned char)(mark[c] == gid ? lidx[c] : 0);
                }
        grp.insert(grp.end(), {r0, e0, (int)verts.size(), (int)cur.size()});
        verts.insert(verts.end(), cur.begin(), cur.end());
        ++gid;
    }
    grp.insert(grp.end(), {n, vf_ptr[n], (int)verts.size(), 0});
    return cloc_ok;
}

// Mean length of the surface's distinct edges (distance_engines.py:17-21).
double mean_edge_faces(const double* V, int m, const int* F) {
    std::vector<uint64_t> e;
    e.reserve(3 * (size_t)m);
    for (int f = 0; f < m; ++f)
        for (int c = 0; c < 3; ++c) {
            uint32_t a = (uint32_t)F[3 * (size_t)f + c], b = (uint32_t)F[3 * (size_t)f + (c + 1) % 3];
            if (a > b) std::swap(a, b);
            e.push_back(((uint64_t)a << 32) | b);
        }
    std::sort(e.begin(), e.end());
    e.erase(std::unique(e.begin(), e.end()), e.end());
    double sum = 0.0;
    for (uint64_t k : e) {
        const double* p = V + 3 * (size_t)(k >> 32);
        const double* q = V + 3 * (size_t)(k & 0xffffffffu);
        sum += std::sqrt((p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) + (p[2] - q[2]) * (p[2] - q[2]));
    }
    return e.empty() ? 0.0 : sum / (double)e.size();
}

void check_faces(int n, const double* V, int m, const int* F) {
    if (n <= 0 || m <= 0) throw Fail("empty mesh");
    for (int64_t i = 0; i < 3LL * m; ++i)
        if (F[i] < 0 || F[i] >= n) throw Fail("face index out of range");
    for (int f = 0; f < m; ++f) {
        const double* a = V + 3 * F[3 * f];
        const double* b = V + 3 * F[3 * f + 1];
        const double* c = V + 3 * F[3 * f + 2];
        const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        const double x = e1[1] * e2[2] - e1[2] * e2[1], y = e1[2] * e2[0] - e1[0] * e2[2],
                     z = e1[0] * e2[1] - e1[1] * e2[0];
        if (!(std::sqrt(x * x + y * y + z * z) > 0.0))
            throw Fail("Zero-area faces must be removed before building the heat engine.");
    }
}

template <class Fn>
int guard(Fn&& fn) {
    try {
        fn();
        return 0;
    } catch (const Fail& e) {
        g_err = e.what();
        return 1;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return 2;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 3;
    }
}

}  // namespace

extern "C" {

void fcw_default_options(fcw_options* opt) {
    std::memset(opt, 0, sizeof(*opt));
    opt->t = 0.0;
    opt->t_coef = 1.0;
    opt->poisson_shift = 1e-10;
    opt->precision = FCW_FP64;
    opt->block = 256;
    opt->leaf_size = 64;
    opt->host_threads = 0;
    opt->device = 0;
    opt->ir_steps = -1;
    opt->streams = 8;
    opt->robust = 0;
    opt->sparse_rhs = 1;
    opt->split_depth = 0;
    opt->band_cap = 6144;
    opt->graphs = 1;
    opt->index_staging = 1;
}

const char* fcw_last_error(void) { return g_err.c_str(); }

int fcw_build_info(char* buf, int len) {
    char tmp[512];
    std::snprintf(tmp, sizeof(tmp),
                  "libfcw gfx950 heat-method engine; solver=separator-tree multifrontal FP64 Cholesky, "
                  "partitioned-inverse panel sweeps; hip=%d.%d; abi=1",
                  HIP_VERSION_MAJOR, HIP_VERSION_MINOR);
    if (buf && len > 0) {
        std::strncpy(buf, tmp, (size_t)len - 1);
        buf[len - 1] = '\0';
    }
    return (int)std::strlen(tmp);
}

int fcw_create(const double* Vin, int n, const int* F, int m, const fcw_options* opt_in, fcw_ctx** out) {
    return guard([&] {
        if (!out) throw Fail("null output pointer");
        *out = nullptr;
        fcw_options opt;
        fcw_default_options(&opt);
        if (opt_in) opt = *opt_in;
        if (opt.precision != FCW_FP64 && opt.precision != FCW_FP32_IR && opt.precision != FCW_BF16X3_IR)
            throw Fail("unknown precision");
        if (opt.ir_steps < 0) opt.ir_steps = opt.precision == FCW_BF16X3_IR ? 6 : 2;
        if (opt.ir_steps > 8) throw Fail("ir_steps must be in 0..8 (or negative for the form's default)");
        if (opt.block <= 0) opt.block = 256;
        if (opt.streams <= 0) opt.streams = 1;
        if (opt.streams > 16) throw Fail("streams must be in 1..16");
        if (!(opt.poisson_shift > 0.0) || !std::isfinite(opt.poisson_shift))
            throw Fail("poisson_shift must be positive and finite");
        check_faces(n, Vin, m, F);
        std::unique_ptr<fcw_ctx> C(new fcw_ctx());
        C->opt = opt;
        C->n = n;
        C->m = m;
        C->ldk = ((opt.block + fcw::kColPad - 1) / fcw::kColPad) * fcw::kColPad;
        if (opt.band_cap < 0) throw Fail("band_cap must be >= 0 (0 = default)");
        if (opt.tile_rows != 0 && opt.tile_rows != 16 && opt.tile_rows != 32 && opt.tile_rows != 64)
            throw Fail("tile_rows must be 0 (chosen per level), 16, 32 or 64");
        if (opt.band_cap > 0) C->band_cap = opt.band_cap;
        // the divergence addresses one RHS block with 32-bit byte offsets
        if ((uint64_t)n * (uint64_t)C->ldk * 8u >= (1ull << 32))
            throw Fail("an RHS block of this mesh would exceed 4 GiB; use a smaller block");
        HIPCHK(hipSetDevice(opt.device));
        C->noalloc = &scratch_slot(opt.device);
        double t0 = now_s();
        std::string err;
        fcw::Symbolic& S = C->sym;
        // robust: the operators live on the intrinsic Delaunay triangulation of the mollified surface
        // (same vertices, flipped faces); the metric kernels keep the surface's own faces
        fcw::Intrinsic intr;
        const int* Fop = F;
        int mop = m;  // faces of the operator triangulation
        C->robust = opt.robust != 0;
        if (C->robust) {
            if (fcw::intrinsic_delaunay(n, Vin, m, F, fcw::kMollify, intr, err)) throw Fail(err);
            Fop = intr.faces.data();
            mop = (int)(intr.faces.size() / 3);
        }
        C->mo = mop;
        if (fcw::analyse(n, Vin, mop, Fop, opt.leaf_size > 0 ? opt.leaf_size : 64, S, err)) throw Fail(err);
        C->node_of_pos.assign(n, -1);
        for (int id = 0; id < S.nnodes; ++id)
            for (int p = S.node_start[id]; p < S.node_start[id] + S.node_s[id]; ++p) C->node_of_pos[p] = id;
        double t1 = now_s();
        std::vector<double> Vp((size_t)n * 3);
        for (int p = 0; p < n; ++p)
            for (int c = 0; c < 3; ++c) Vp[3 * (size_t)p + c] = Vin[3 * (size_t)S.perm[p] + c];
        const double mean_edge = C->robust ? mean_edge_faces(Vin, m, F) : mean_edge_length(S, Vp.data());
        // the surface's faces in positions, ordered by their smallest position (S.faces when the
        // operators use the same triangulation)
        std::vector<int> gfaces;
        if (C->robust) {
            std::vector<int> idx(m);
            std::iota(idx.begin(), idx.end(), 0);
            auto lo = [&](int f) {
                return std::min({S.iperm[F[3 * (size_t)f]], S.iperm[F[3 * (size_t)f + 1]], S.iperm[F[3 * (size_t)f + 2]]});
            };
            std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return lo(a) < lo(b); });
            gfaces.resize(3 * (size_t)m);
            for (int f = 0; f < m; ++f)
                for (int a = 0; a < 3; ++a) gfaces[3 * (size_t)f + a] = S.iperm[F[3 * (size_t)idx[f] + a]];
        }
        const std::vector<int>& G = C->robust ? gfaces : S.faces;
        const double t = (opt.t > 0.0) ? opt.t : opt.t_coef * mean_edge * mean_edge;
        if (!(t > 0.0) || !std::isfinite(t)) throw Fail("diffusion time must be positive and finite");
        int64_t& tb = C->dev_bytes;
        for (int i = 0; i < opt.streams; ++i) {
            C->ws.emplace_back(new Workspace());
            C->alloc_workspace(*C->ws.back());
        }
        hipStream_t st = C->ws[0]->stream;
        C->V.upload(Vp.data(), Vp.size(), tb);
        C->faces.upload(G.data(), G.size(), tb);
        C->vf_ptr.upload(S.vf_ptr.data(), S.vf_ptr.size(), tb);
        C->vf.upload(S.vf.data(), S.vf.size(), tb);
        C->rowptr.upload(S.rowptr.data(), S.rowptr.size(), tb);
        C->colind.upload(S.colind.data(), S.colind.size(), tb);
        C->cptr.upload(S.contrib_ptr.data(), S.contrib_ptr.size(), tb);
        C->contrib.upload(S.contrib.data(), S.contrib.size(), tb);
        C->iperm.upload(S.iperm.data(), S.iperm.size(), tb);
        C->area.alloc(m, tb);
        C->gram.alloc((size_t)m * 6, tb);
        C->larea.alloc(m, tb);
        C->g11.alloc(m, tb);
        C->g12.alloc(m, tb);
        C->g22.alloc(m, tb);
        const int nnz = S.rowptr[n];
        C->lval.alloc(nnz, tb);
        C->mass.alloc(n, tb);
        fcw::launch_face_geometry(st, m, C->V.p, C->faces.p, C->area.p, C->gram.p, C->larea.p, C->g11.p, C->g12.p,
                                  C->g22.p);
        if (C->robust) {
            // operator faces in the analysis's order; corner c keeps its input corner's opposite edge.
            // A face of the tufted cover left without its mirror copy counts half its area.
            std::vector<double> len(3 * (size_t)mop), oa(mop), og(6 * (size_t)mop);
            for (int f = 0; f < mop; ++f)
                for (int c = 0; c < 3; ++c) len[3 * (size_t)f + c] = intr.len[3 * (size_t)S.face_perm[f] + c];
            fcw::intrinsic_operators(mop, len.data(), oa.data(), og.data());
            for (int f = 0; f < mop; ++f) oa[f] *= intr.weight[S.face_perm[f]];
            C->op_faces.upload(S.faces.data(), S.faces.size(), tb);
            C->op_area.upload(oa.data(), oa.size(), tb);
            C->op_gram.upload(og.data(), og.size(), tb);
        }
        fcw::launch_assemble(st, n, nnz, C->cptr.p, C->contrib.p, C->robust ? C->op_area.p : C->area.p,
                             C->robust ? C->op_gram.p : C->gram.p, C->vf_ptr.p, C->vf.p, C->lval.p, C->mass.p);
        {
            // divergence row groups over the operator triangulation's incidences
            std::vector<int> grp, verts;
            std::vector<unsigned> loc(std::max<size_t>(1, S.vf.size()), 0u);
            // the refined forms' residual reads x through the same groups
            const bool refined = opt.precision != FCW_FP64;
            std::vector<unsigned char> cloc(refined ? std::max<size_t>(1, S.colind.size()) : 0);
            const bool cloc_ok = build_divergence_groups(n, S.vf_ptr, S.vf, S.faces, grp, verts, loc, &S.rowptr,
                                                         &S.colind, refined ? &cloc : nullptr);
            if (refined && cloc_ok) {
                C->res_cloc.upload(cloc.data(), cloc.size(), tb);
                C->res_aval.alloc(std::max<size_t>(1, S.colind.size()), tb);
                fcw::launch_shifted_values(st, n, C->rowptr.p, C->colind.p, C->lval.p, C->mass.p, opt.poisson_shift,
                                           C->res_aval.p);
                C->res_grouped = true;
            }
            C->div_ngroups = (int)(grp.size() / 4) - 1;
            C->div_grp.upload(grp.data(), grp.size(), tb);
            C->div_verts.upload(verts.data(), std::max<size_t>(1, verts.size()), tb);
            C->div_loc.upload(loc.data(), loc.size(), tb);
            C->div_geo.alloc(std::max<size_t>(1, S.vf.size()) * 5, tb);
            fcw::launch_incidence_geometry(st, (int)S.vf.size(), C->vf.p, C->robust ? C->op_area.p : C->area.p,
                                           C->robust ? C->op_gram.p : C->gram.p, C->div_geo.p);
        }
        {
            // 256-face chunks of the face order: their vertices, each vertex owned by the first
            // chunk that holds it (a vertex in no face by the last chunk), owned vertices first
            constexpr int kChunk = 256;
            const int nch = (m + kChunk - 1) / kChunk;
            std::vector<int> owner(n, -1), mark(n, -1);
            for (int f = 0; f < m; ++f)
                for (int a = 0; a < 3; ++a) {
                    const int p = G[3 * (size_t)f + a];
                    if (owner[p] < 0) owner[p] = f / kChunk;
                }
            std::vector<std::vector<int>> owned(std::max(1, nch));
            for (int p = 0; p < n; ++p) owned[owner[p] >= 0 ? owner[p] : std::max(0, nch - 1)].push_back(p);
            std::vector<int> vptr(nch + 1, 0), verts, nown(std::max(1, nch), 0), other;
            for (int c = 0; c < nch; ++c) {
                other.clear();
                for (int p : owned[c]) mark[p] = c;
                for (int f = c * kChunk; f < std::min(m, (c + 1) * kChunk); ++f)
                    for (int a = 0; a < 3; ++a) {
                        const int p = G[3 * (size_t)f + a];
                        if (mark[p] == c) continue;
                        mark[p] = c;
                        other.push_back(p);
                    }
                std::sort(other.begin(), other.end());
                nown[c] = (int)owned[c].size();
                verts.insert(verts.end(), owned[c].begin(), owned[c].end());
                verts.insert(verts.end(), other.begin(), other.end());
                vptr[c + 1] = (int)verts.size();
            }
            C->nch = nch;
            C->cvptr.upload(vptr.data(), vptr.size(), tb);
            C->cverts.upload(verts.data(), std::max<size_t>(1, verts.size()), tb);
            C->cnown.upload(nown.data(), std::max<size_t>(1, nown.size()), tb);
            C->chunk_area.alloc(std::max(1, nch), tb);
            fcw::launch_chunk_area(st, m, C->larea.p, C->chunk_area.p);
        }
        HIPCHK(hipGetLastError());
        std::vector<double> lval(nnz), mass(n);
        HIPCHK(hipStreamSynchronize(st));
        HIPCHK(hipMemcpy(lval.data(), C->lval.p, sizeof(double) * nnz, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(mass.data(), C->mass.p, sizeof(double) * n, hipMemcpyDeviceToHost));
        double t2 = now_s();
        std::vector<double> ph((size_t)S.panel_elems), pp((size_t)S.panel_elems);
        if (fcw::factor(S, lval.data(), mass.data(), t, 1.0, ph.data(), err, opt.host_threads))
            throw Fail("heat: " + err);
        if (fcw::factor(S, lval.data(), mass.data(), 1.0, opt.poisson_shift, pp.data(), err, opt.host_threads))
            throw Fail("poisson: " + err);
        double t3 = now_s();
        C->panel_heat.upload(ph.data(), ph.size(), tb);
        C->panel_pois.upload(pp.data(), pp.size(), tb);
        if (opt.precision == FCW_FP32_IR) {
            C->panel_pois32.alloc(pp.size(), tb);
            fcw::launch_to_float(st, C->panel_pois.p, C->panel_pois32.p, (int64_t)pp.size());
        }
        if (opt.precision == FCW_BF16X3_IR) {
            C->panel_pois_bf.alloc(pp.size(), tb);
            fcw::launch_to_bf16x2(st, C->panel_pois.p, C->panel_pois_bf.p, (int64_t)pp.size());
        }
        C->node_start.upload(S.node_start.data(), S.node_start.size(), tb);
        C->node_s.upload(S.node_s.data(), S.node_s.size(), tb);
        C->node_b.upload(S.node_b.data(), S.node_b.size(), tb);
        C->bidx.upload(S.bidx.data(), std::max<size_t>(1, S.bidx.size()), tb);
        C->cmap.upload(S.cmap.data(), std::max<size_t>(1, S.cmap.size()), tb);
        C->child_ptr.upload(S.child_ptr.data(), S.child_ptr.size(), tb);
        C->child_list.upload(S.child_list.data(), std::max<size_t>(1, S.child_list.size()), tb);
        C->level_nodes.upload(S.level_nodes.data(), S.level_nodes.size(), tb);
        C->cown.upload(S.contrib_own.data(), S.contrib_own.size(), tb);
        C->ctail.upload(S.contrib_tail.data(), S.contrib_tail.size(), tb);
        C->node_boff.upload(S.node_boff.data(), S.node_boff.size(), tb);
        C->node_poff.upload(S.node_poff.data(), S.node_poff.size(), tb);
        C->node_uoff.upload(S.node_uoff.data(), S.node_uoff.size(), tb);
        C->level_off.assign(S.level_ptr.begin(), S.level_ptr.end());
        C->level_entries.assign(S.nlevels, 0.0);
        for (int l = 0; l < S.nlevels; ++l)
            for (int i = S.level_ptr[l]; i < S.level_ptr[l + 1]; ++i) C->level_entries[l] += C->node_entries(S.level_nodes[i]);
        // Row tiles of height tm per level: forward over each node's s + b rows, backward over its s.
        auto count_tiles = [&](int l, bool forward, int tm) {
            int nt = 0;
            for (int i = S.level_ptr[l]; i < S.level_ptr[l + 1]; ++i) {
                const int id = S.level_nodes[i];
                nt += ((forward ? S.node_s[id] + S.node_b[id] : S.node_s[id]) + tm - 1) / tm;
            }
            return nt;
        };
        C->fwd_split.assign(S.nlevels, 1);
        C->bwd_split.assign(S.nlevels, 1);
        const bool split = opt.split_depth > 0 || (opt.split_depth == 0 && opt.streams <= 2);
        if (split) {
            // Split the depth of a level's launch when it has few workgroups and a long depth: aim
            // at about kSplitTarget workgroups, at least two depth steps per part, at most 8 parts
            // (64-row tiles). With several blocks in flight the narrow levels of one block already run
            // beside the wide launches of others, and the split's extra workgroups cost throughput, so
            // it is on by default only for one or two streams.
            constexpr int kSplitTarget = 512;
            const int ncol = C->ldk / 64;
            auto choose = [&](int ntiles, int kmax) {
                const int wgs = (ntiles + 7) / 8 * 8 * ncol;
                if (ntiles <= 0 || wgs >= kSplitTarget) return 1;
                return std::max(1, std::min({8, kSplitTarget / wgs, kmax / (2 * fcw::kTileK)}));
            };
            size_t part_elems = 0;
            for (int l = 0; l < S.nlevels; ++l) {
                int kf = 0, kb = 0;
                for (int i = S.level_ptr[l]; i < S.level_ptr[l + 1]; ++i) {
                    const int id = S.level_nodes[i];
                    kf = std::max(kf, S.node_s[id]);
                    kb = std::max(kb, S.node_s[id] + S.node_b[id]);
                }
                const int nf = count_tiles(l, true, 64), nb = count_tiles(l, false, 64);
                if (l > 0) C->fwd_split[l] = choose(nf, kf);  // the leaf level never splits
                C->bwd_split[l] = choose(nb, kb);
                if (C->fwd_split[l] > 1) part_elems = std::max(part_elems, (size_t)nf * ncol * C->fwd_split[l] * 4096);
                if (C->bwd_split[l] > 1) part_elems = std::max(part_elems, (size_t)nb * ncol * C->bwd_split[l] * 4096);
            }
            for (auto& w : C->ws) w->part.alloc(std::max<size_t>(1, part_elems), tb);
        }
        // tile heights (a split level keeps 64-row tiles), then the tile lists
        C->fwd_tm.assign(S.nlevels, 64);
        C->bwd_tm.assign(S.nlevels, 64);
        for (int l = 0; l < S.nlevels; ++l) {
            if (C->fwd_split[l] == 1) C->fwd_tm[l] = choose_tile_rows(S, l, true, C->ldk, opt.tile_rows);
            if (C->bwd_split[l] == 1) C->bwd_tm[l] = choose_tile_rows(S, l, false, C->ldk, opt.tile_rows);
        }
        std::vector<fcw::LevelTile> ft, bt;
        C->fwd_tile_off.assign(1, 0);
        C->bwd_tile_off.assign(1, 0);
        for (int l = 0; l < S.nlevels; ++l) {
            for (int i = S.level_ptr[l]; i < S.level_ptr[l + 1]; ++i) {
                const int id = S.level_nodes[i];
                const int f = S.node_s[id] + S.node_b[id];
                for (int m0 = 0; m0 < f; m0 += C->fwd_tm[l]) ft.push_back({id, m0});
                for (int m0 = 0; m0 < S.node_s[id]; m0 += C->bwd_tm[l]) bt.push_back({id, m0});
            }
            C->fwd_tile_off.push_back((int)ft.size());
            C->bwd_tile_off.push_back((int)bt.size());
        }
        C->fwd_tiles.upload(ft.data(), std::max<size_t>(1, ft.size()), tb);
        C->bwd_tiles.upload(bt.data(), std::max<size_t>(1, bt.size()), tb);
        for (int l = 0; l < S.nlevels; ++l) C->info.split_levels += (C->fwd_split[l] > 1) + (C->bwd_split[l] > 1);
        const int max_s = S.nnodes ? *std::max_element(S.node_s.begin(), S.node_s.end()) : 0;
        C->fwd_idx_lds = opt.index_staging != 0 && max_s <= fcw::kSweepIdxCap;
        C->bwd_idx_lds = opt.index_staging != 0 && S.max_front <= 2 * fcw::kSweepIdxCap;
        double t4 = now_s();
        fcw_info& I = C->info;
        I.n = n;
        I.m = m;
        I.nnodes = S.nnodes;
        I.nlevels = S.nlevels;
        I.max_front = S.max_front;
        I.block = C->ldk;
        I.factor_nnz = S.factor_nnz;
        I.panel_elems = S.panel_elems;
        I.factor_flops = S.factor_flops;
        I.solve_flops_per_rhs = S.solve_flops_per_rhs;
        I.t = t;
        I.robust = C->robust ? 1 : 0;
        I.intrinsic_flips = intr.flips;
        I.intrinsic_skipped = intr.skipped;
        I.mollify_delta = intr.delta;
        I.op_faces = mop;
        I.tufted_cover = intr.cover;
        I.nonmanifold_edges = intr.nonmanifold_edges;
        I.mean_edge = mean_edge;
        I.setup_seconds[0] = t1 - t0;
        I.setup_seconds[1] = t2 - t1;
        I.setup_seconds[2] = t3 - t2;
        I.setup_seconds[3] = t4 - t3;
        *out = C.release();
    });
}

int fcw_destroy(fcw_ctx* ctx) {
    return guard([&] { delete ctx; });
}

int fcw_get_info(const fcw_ctx* ctx, fcw_info* out) {
    return guard([&] {
        if (!ctx || !out) throw Fail("null argument");
        *out = ctx->info;
        out->device_bytes = ctx->dev_bytes;
    });
}

static std::vector<int> source_positions(fcw_ctx* C, const int* src, int k) {
    std::vector<int> pos(k);
    for (int j = 0; j < k; ++j) {
        if (src[j] < 0 || src[j] >= C->n)
            throw Fail("source_idx out of range: " + std::to_string(src[j]) + " for mesh with " +
                       std::to_string(C->n) + " vertices");
        pos[j] = C->sym.iperm[src[j]];
    }
    return pos;
}

int fcw_distance_block(fcw_ctx* C, const int* src, int k, double* out) {
    return guard([&] {
        if (!C || !src || !out) throw Fail("null argument");
        if (k <= 0 || k > C->ldk) throw Fail("source count must be in 1..block (" + std::to_string(C->ldk) + ")");
        std::vector<int> pos = source_positions(C, src, k);
        Workspace& w = *C->ws[0];
        HIPCHK(hipMemcpy(w.src_pos.p, pos.data(), sizeof(int) * k, hipMemcpyHostToDevice));
        CallPlans plans = C->make_plans(pos);
        C->distances(w, k, C->opt.sparse_rhs ? &plans.blocks[0] : nullptr);
        if (!C->Dout.p) C->Dout.alloc((size_t)C->n * C->ldk, C->dev_bytes);
        fcw::launch_gather_rows(w.stream, C->n, C->ldk, k, w.Dt.p, w.Z.p, w.src_pos.p, C->iperm.p, C->Dout.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(w.stream));
        HIPCHK(hipMemcpy(out, C->Dout.p, sizeof(double) * (size_t)k * C->n, hipMemcpyDeviceToHost));
        C->ev_collect(w);
    });
}

int fcw_divergence_block(fcw_ctx* C, const double* u, int k, double* out) {
    return guard([&] {
        if (!C || !u || !out) throw Fail("null argument");
        if (k <= 0 || k > C->ldk) throw Fail("column count must be in 1..block (" + std::to_string(C->ldk) + ")");
        const int n = C->n, ldk = C->ldk;
        const std::vector<int>& iperm = C->sym.iperm;
        // vertex v's row at its position, columns past k zero
        std::vector<double> h((size_t)n * ldk, 0.0);
        for (int v = 0; v < n; ++v)
            std::memcpy(h.data() + (size_t)iperm[v] * ldk, u + (size_t)v * k, sizeof(double) * k);
        Workspace& w = *C->ws[0];
        HIPCHK(hipMemcpyAsync(w.Z.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, w.stream));
        fcw::launch_divergence(w.stream, C->div_dev(), ldk, w.Z.p, w.X.p, nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(h.data(), w.X.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, w.stream));
        HIPCHK(hipStreamSynchronize(w.stream));
        for (int v = 0; v < n; ++v)
            std::memcpy(out + (size_t)v * k, h.data() + (size_t)iperm[v] * ldk, sizeof(double) * k);
    });
}

int fcw_set_metrics(fcw_ctx* C, const double* targets, int S, const double* fracs, int q, const int* boundary, int nb,
                    const double* vertex_w, double eps, double health_h, double bin_width, double cap_fraction,
                    double rtol, int max_iter) {
    return guard([&] {
        if (!C || !targets || !vertex_w) throw Fail("null argument");
        if (S < 1 || S > fcw::kMaxScales) throw Fail("number of scales must be in 1..32");
        if (q < 0 || (S - 1) * q > fcw::kMaxExtras) throw Fail("too many supplementary samples ((S-1)*q <= 256)");
        if (q > 0 && !fracs) throw Fail("null fracs");
        if (nb > 0 && !boundary) throw Fail("null boundary");
        for (int s = 1; s < S; ++s)
            if (targets[s] < targets[s - 1]) throw Fail("targets must be sorted ascending");
        if (!(bin_width > 0.0) || !std::isfinite(bin_width)) throw Fail("bin_width must be positive");
        if (max_iter < 1) throw Fail("max_iter must be >= 1");
        int64_t& tb = C->dev_bytes;
        C->sync_all();
        C->S = S;
        C->q = q;
        C->nb = nb;
        C->targets.upload(targets, S, tb);
        std::vector<double> fr(std::max(1, q), 0.0);
        for (int j = 0; j < q; ++j) fr[j] = fracs[j];
        C->fracs.upload(fr.data(), fr.size(), tb);
        std::vector<int> bpos(std::max(1, nb), 0);
        for (int i = 0; i < nb; ++i) {
            if (boundary[i] < 0 || boundary[i] >= C->n) throw Fail("boundary index out of range");
            bpos[i] = C->sym.iperm[boundary[i]];
        }
        C->boundary.upload(bpos.data(), bpos.size(), tb);
        std::vector<double> wv(C->n);
        for (int v = 0; v < C->n; ++v) wv[C->sym.iperm[v]] = vertex_w[v];
        C->vw.upload(wv.data(), wv.size(), tb);
        C->eps = eps;
        C->health_h = health_h;
        C->width = bin_width;
        C->cap = cap_fraction;
        C->rtol = rtol;
        C->max_iter = max_iter;
        for (auto& w : C->ws) {
            w->band.alloc((size_t)C->ldk * fcw::band_pool_doubles((int64_t)S * C->band_cap), tb);
            w->cpart.alloc((size_t)C->ldk * std::max(1, C->nch) * fcw::kChunkStats, tb);
            w->ch_lo.alloc((size_t)C->ldk * std::max(1, C->nch), tb);
            w->ch_hi.alloc((size_t)C->ldk * std::max(1, C->nch), tb);
        }
        std::vector<int> flag(C->n, 0);
        for (int i = 0; i < nb; ++i) flag[bpos[i]] = 1;
        C->isb.upload(flag.data(), flag.size(), tb);
        C->o_cap = 0;
        C->metrics_ready = true;
    });
}

int fcw_wiring_sources(fcw_ctx* C, const int* src, int nsrc, fcw_metrics_out* out) {
    return guard([&] {
        if (!C || !src || !out) throw Fail("null argument");
        if (!C->metrics_ready) throw Fail("fcw_set_metrics must be called first");
        if (nsrc <= 0) return;
        const double w0 = now_s();
        std::vector<int> pos = source_positions(C, src, nsrc);
        const int* dpos = nullptr;
        CallPlans plans = C->make_plans(pos, &dpos);
        C->ensure_outputs((size_t)nsrc);
        C->run_wiring(dpos, nsrc, plans, C->own_outputs());
        C->stage[5] = now_s() - w0;
        C->copy_outputs(C->own_outputs(), (size_t)nsrc, out);
        for (auto& w : C->ws) C->ev_collect(*w);
        C->stage[4] = now_s() - w0;
    });
}

int fcw_wiring_sources_noalloc(fcw_ctx* C, const int* src, int nsrc, fcw_metrics_out* out) {
    return guard([&] {
        if (!C || !src || !out) throw Fail("null argument");
        if (!C->metrics_ready) throw Fail("fcw_set_metrics must be called first");
        if (nsrc <= 0) return;
        if (nsrc > kNoAllocMaxSources)
            throw Fail("fcw_wiring_sources_noalloc takes at most " + std::to_string(kNoAllocMaxSources) +
                       " sources per call (got " + std::to_string(nsrc) + ")");
        if (C->profiling) throw Fail("fcw_wiring_sources_noalloc does not run with profiling on (it creates events)");
        const double w0 = now_s();
        HIPCHK(hipSetDevice(C->opt.device));
        std::vector<int> pos = source_positions(C, src, nsrc);
        CallPlans plans = C->plan_blocks(pos);
        auto up = [](size_t b) { return (b + 255) / 256 * 256; };
        const size_t b_pos = up(pos.size() * sizeof(int));
        const size_t b_tiles = up(std::max<size_t>(1, plans.tiles.size()) * sizeof(fcw::LevelTile));
        const size_t b_nodes = up(std::max<size_t>(1, plans.nodes.size()) * sizeof(int));
        const size_t need = b_pos + b_tiles + b_nodes + C->output_bytes((size_t)nsrc);
        if (need > kNoAllocScratchBytes)
            throw Fail("these " + std::to_string(nsrc) + " sources need " + std::to_string(need) +
                       " bytes of device scratch; fcw_wiring_sources_noalloc holds " +
                       std::to_string(kNoAllocScratchBytes) + ": pass fewer sources per call");
        // one call at a time uses this GPU's array; the previous call has synchronised its streams
        std::lock_guard<std::mutex> lk(C->noalloc->use);
        unsigned char* base = C->noalloc->ptr;
        int* dpos = reinterpret_cast<int*>(base);
        fcw::LevelTile* dtiles = reinterpret_cast<fcw::LevelTile*>(base + b_pos);
        int* dnodes = reinterpret_cast<int*>(base + b_pos + b_tiles);
        const OutBase ob = C->carve_outputs(base + b_pos + b_tiles + b_nodes, (size_t)nsrc);
        hipStream_t st = C->ws[0]->stream;
        HIPCHK(hipMemcpyAsync(dpos, pos.data(), pos.size() * sizeof(int), hipMemcpyHostToDevice, st));
        if (!plans.tiles.empty())
            HIPCHK(hipMemcpyAsync(dtiles, plans.tiles.data(), plans.tiles.size() * sizeof(fcw::LevelTile),
                                  hipMemcpyHostToDevice, st));
        if (!plans.nodes.empty())
            HIPCHK(hipMemcpyAsync(dnodes, plans.nodes.data(), plans.nodes.size() * sizeof(int),
                                  hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));  // every stream's blocks read these
        plans.bind(dtiles, dnodes);
        C->no_new_graphs = true;  // instantiating a capture allocates; this path issues every launch itself
        try {
            C->run_wiring(dpos, nsrc, plans, ob);
        } catch (...) {
            C->no_new_graphs = false;
            throw;
        }
        C->no_new_graphs = false;
        C->stage[5] = now_s() - w0;
        C->copy_outputs(ob, (size_t)nsrc, out);
        C->stage[4] = now_s() - w0;
    });
}

int fcw_geometry_fields(fcw_ctx* C, const double* D, const int* src, int k, int health_screen,
                        fcw_metrics_out* out) {
    return guard([&] {
        if (!C || !D || !out) throw Fail("null argument");
        if (!C->metrics_ready) throw Fail("fcw_set_metrics must be called first");
        if (k <= 0 || k > C->ldk) throw Fail("field count must be in 1..block");
        std::vector<int> zeros(k, 0);
        std::vector<int> pos = source_positions(C, src ? src : zeros.data(), k);
        Workspace& w = *C->ws[0];
        C->sync_all();
        HIPCHK(hipMemcpy(w.src_pos.p, pos.data(), sizeof(int) * k, hipMemcpyHostToDevice));
        std::vector<double> Dp((size_t)k * C->n);
        for (int j = 0; j < k; ++j)
            for (int p = 0; p < C->n; ++p) Dp[(size_t)j * C->n + p] = D[(size_t)j * C->n + C->sym.perm[p]];
        HIPCHK(hipMemcpy(w.Dt.p, Dp.data(), sizeof(double) * Dp.size(), hipMemcpyHostToDevice));
        C->ensure_outputs((size_t)k);
        fcw::launch_metrics(w.stream, k, C->mesh_dev(), C->metrics_dev(w), w.Dt.p, w.src_pos.p,
                            C->outputs_at(C->own_outputs(), 0), health_screen != 0);
        HIPCHK(hipGetLastError());
        C->copy_outputs(C->own_outputs(), (size_t)k, out);
    });
}

int fcw_stage_times(fcw_ctx* C, double* out, int len) {
    if (!C || !out) return 0;
    int c = std::min(len, 6);
    for (int i = 0; i < c; ++i) out[i] = C->stage[i];
    return c;
}

int fcw_set_profiling(fcw_ctx* C, int on) {
    return guard([&] {
        if (!C) throw Fail("null context");
        C->sync_all();
        C->profiling = on != 0;
        for (int i = 0; i < kNumTimed; ++i) {
            C->ktime[i] = 0.0;
            C->kcount[i] = 0;
        }
        C->launch_log.clear();
        for (int i = 0; i < 6; ++i) C->stage[i] = 0.0;
    });
}

int fcw_kernel_times(fcw_ctx* C, double* out, int len) {
    if (!C || !out) return 0;
    int c = 0;
    for (int i = 0; i < kNumTimed; ++i) {
        if (c < len) out[c++] = C->ktime[i];
        if (c < len) out[c++] = (double)C->kcount[i];
    }
    return c;
}

int fcw_kernel_log(fcw_ctx* C, double* out, int max_rows) {
    if (!C) return 0;
    const int n = (int)C->launch_log.size();
    if (!out) return n;
    const int c = std::min(n, max_rows);
    for (int i = 0; i < c; ++i) {
        const LaunchRec& r = C->launch_log[i];
        double* o = out + 6 * (size_t)i;
        o[0] = r.kind;
        o[1] = r.system;
        o[2] = r.level;
        o[3] = r.wgs;
        o[4] = r.entries;
        o[5] = r.ms;
    }
    return c;
}

int fcw_calibration_copy(int64_t bytes, int device) {
    return guard([&] {
        if (bytes <= 0 || bytes % 32) throw Fail("bytes must be a positive multiple of 32");
        HIPCHK(hipSetDevice(device));
        void *a = nullptr, *b = nullptr;
        HIPCHK(hipMalloc(&a, (size_t)bytes));
        HIPCHK(hipMalloc(&b, (size_t)bytes));
        fcw::launch_calibration_copy(nullptr, a, b, 0);  // nothing copied: warm the code object
        fcw::launch_calibration_copy(nullptr, a, b, bytes);
        hipError_t e = hipDeviceSynchronize();
        (void)hipFree(a);
        (void)hipFree(b);
        HIPCHK(e);
    });
}

int fcw_mesh_prep(const double* V, int n, const int* F, int m, const unsigned char* mask, int start, int device,
                  fcw_mesh_prep_out* out) {
    return guard([&] {
        if (!V || !F || !mask || !out) throw Fail("null argument");
        if (n <= 0 || m <= 0) throw Fail("empty mesh");
        if ((int64_t)m * 6 >= INT32_MAX) throw Fail("too many faces");
        for (int64_t i = 0; i < 3LL * m; ++i)
            if (F[i] < 0 || F[i] >= n) throw Fail("face index out of range");
        if (!out->sub_to_orig || !out->orig_to_sub || !out->faces || !out->vertices || !out->boundary ||
            !out->order || !out->face_areas || !out->face_L || !out->vertex_areas || !out->adj_ptr || !out->adj)
            throw Fail("null output buffer");
        fcw::MeshPrepOut o{0, 0, 0, 0, out->sub_to_orig, out->orig_to_sub, out->faces, out->vertices, out->boundary,
                           out->order, out->face_areas, out->face_L, out->vertex_areas, out->adj_ptr, out->adj};
        std::string err;
        if (fcw::mesh_prep_device(n, V, m, F, mask, start, device, o, err)) throw Fail(err);
        out->n_sub = o.n_sub;
        out->m_sub = o.m_sub;
        out->n_boundary = o.n_boundary;
        out->n_adj = o.n_adj;
    });
}

int fcw_host_intrinsic(const double* V, int n, const int* F, int m, double mollify, int* faces_out, double* len_out,
                       double* area_out, double* gram_out, double* stats) {
    return guard([&] {
        if (!V || !F || !faces_out || !len_out) throw Fail("null argument");
        check_faces(n, V, m, F);
        fcw::Intrinsic I;
        std::string err;
        if (fcw::intrinsic_delaunay(n, V, m, F, mollify, I, err)) throw Fail(err);
        if (I.cover)
            throw Fail("non-manifold or inconsistently oriented edge: the triangulation is built on the tufted cover, "
                       "use fcw_host_intrinsic_cover");
        std::copy(I.faces.begin(), I.faces.end(), faces_out);
        std::copy(I.len.begin(), I.len.end(), len_out);
        if (area_out && gram_out) fcw::intrinsic_operators(m, I.len.data(), area_out, gram_out);
        if (stats) {
            stats[0] = I.flips;
            stats[1] = I.skipped;
            stats[2] = I.delta;
        }
    });
}

}  // extern "C"

namespace {

// Stiffness values in the analysis's CSR order and the lumped mass in positions, assembled on the
// host from the input surface.
void host_assemble(const fcw::Symbolic& S, const double* V, const int* F, std::vector<double>& lval,
                   std::vector<double>& mpos) {
    const int n = S.n;
    std::vector<std::vector<std::pair<int, double>>> rows;
    std::vector<double> mass;
    host_operators(n, V, S.m, F, rows, mass);
    lval.assign(S.rowptr[n], 0.0);
    mpos.assign(n, 0.0);
    for (int v = 0; v < n; ++v) {
        const int p = S.iperm[v];
        mpos[p] = mass[v];
        const int* row = S.colind.data() + S.rowptr[p];
        const int len = S.rowptr[p + 1] - S.rowptr[p];
        for (auto& e : rows[v]) {
            const int qv = S.iperm[e.first];
            const int pos = (int)(std::lower_bound(row, row + len, qv) - row);
            lval[S.rowptr[p] + pos] += e.second;
        }
    }
}

}  // namespace

extern "C" {

int fcw_host_intrinsic_cover(const double* V, int n, const int* F, int m, double mollify, int* mo, int* faces_out,
                             double* len_out, double* weight_out, double* area_out, double* gram_out, double* stats) {
    return guard([&] {
        if (!V || !F || !mo || !faces_out || !len_out || !weight_out) throw Fail("null argument");
        check_faces(n, V, m, F);
        fcw::Intrinsic I;
        std::string err;
        if (fcw::intrinsic_delaunay(n, V, m, F, mollify, I, err)) throw Fail(err);
        const int k = (int)(I.faces.size() / 3);
        *mo = k;
        std::copy(I.faces.begin(), I.faces.end(), faces_out);
        std::copy(I.len.begin(), I.len.end(), len_out);
        std::copy(I.weight.begin(), I.weight.end(), weight_out);
        if (area_out && gram_out) {
            fcw::intrinsic_operators(k, I.len.data(), area_out, gram_out);
            for (int f = 0; f < k; ++f) area_out[f] *= I.weight[f];
        }
        if (stats) {
            stats[0] = I.flips;
            stats[1] = I.skipped;
            stats[2] = I.delta;
            stats[3] = I.cover;
            stats[4] = I.nonmanifold_edges;
        }
    });
}

int fcw_host_analyse(const double* V, int n, const int* F, int m, int leaf_size, int par_depth, int* perm,
                     int* node_start, int* node_s, int* node_b, double* stats, uint64_t* panel_hash) {
    return guard([&] {
        if (!V || !F || !perm || !node_start || !node_s || !node_b) throw Fail("null argument");
        check_faces(n, V, m, F);
        fcw::Symbolic S;
        std::string err;
        if (fcw::analyse(n, V, m, F, leaf_size > 0 ? leaf_size : 64, S, err, par_depth)) throw Fail(err);
        std::copy(S.perm.begin(), S.perm.end(), perm);
        std::copy(S.node_start.begin(), S.node_start.end(), node_start);
        std::copy(S.node_s.begin(), S.node_s.end(), node_s);
        std::copy(S.node_b.begin(), S.node_b.end(), node_b);
        if (stats) {
            stats[0] = S.nnodes;
            stats[1] = S.nlevels;
            stats[2] = (double)S.factor_nnz;
            stats[3] = (double)S.panel_elems;
        }
        if (panel_hash) {
            std::vector<double> lval, mpos;
            host_assemble(S, V, F, lval, mpos);
            std::vector<double> panel((size_t)S.panel_elems);
            if (fcw::factor(S, lval.data(), mpos.data(), 1.0, 1.0, panel.data(), err, 0)) throw Fail(err);
            uint64_t h = 1469598103934665603ull;  // FNV-1a over the panel bytes
            const unsigned char* b = reinterpret_cast<const unsigned char*>(panel.data());
            for (size_t i = 0; i < panel.size() * sizeof(double); ++i) h = (h ^ b[i]) * 1099511628211ull;
            *panel_hash = h;
        }
    });
}

int fcw_host_factor_solve(const double* V, int n, const int* F, int m, double l_scale, double m_scale, int leaf_size,
                          const double* B, int k, double* X, double* stats) {
    return guard([&] {
        check_faces(n, V, m, F);
        fcw::Symbolic S;
        std::string err;
        if (fcw::analyse(n, V, m, F, leaf_size > 0 ? leaf_size : 64, S, err)) throw Fail(err);
        std::vector<double> lval, mpos;
        host_assemble(S, V, F, lval, mpos);
        std::vector<double> panel((size_t)S.panel_elems);
        if (fcw::factor(S, lval.data(), mpos.data(), l_scale, m_scale, panel.data(), err, 0)) throw Fail(err);
        std::vector<double> Xp((size_t)n * k);
        for (int v = 0; v < n; ++v)
            std::memcpy(Xp.data() + (size_t)S.iperm[v] * k, B + (size_t)v * k, sizeof(double) * k);
        fcw::panel_solve_host(S, panel.data(), Xp.data(), k);
        for (int v = 0; v < n; ++v)
            std::memcpy(X + (size_t)v * k, Xp.data() + (size_t)S.iperm[v] * k, sizeof(double) * k);
        if (stats) {
            stats[0] = S.nnodes;
            stats[1] = S.nlevels;
            stats[2] = (double)S.factor_nnz;
            stats[3] = (double)S.panel_elems;
        }
    });
}

}  // extern "C"
