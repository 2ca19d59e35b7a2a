// C-ABI runtime of libfcw.so: context lifetime, device memory, and the
// per-block pipeline  delta RHS -> heat sweeps -> fused divergence ->
// Poisson sweeps -> shift/transpose -> metric kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fcw.h"
#include "fcw_device.h"
#include "fcw_internal.h"

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
};

struct EventPair {
    hipEvent_t a, b;
};

}  // namespace

struct fcw_ctx {
    fcw::Symbolic sym;
    fcw_options opt{};
    fcw_info info{};
    int n = 0, m = 0, ldk = 0;
    hipStream_t stream = nullptr;
    int64_t dev_bytes = 0;

    // mesh and operators (positions)
    DBuf<double> V, area, gb, larea, g11, g12, g22, lval, mass;
    DBuf<int> faces, vf_ptr, vf, rowptr, colind, cptr, contrib, iperm;
    // tree
    DBuf<int> node_start, node_s, node_b, bidx, cmap, child_ptr, child_list, level_nodes;
    DBuf<int64_t> node_boff, node_poff, node_uoff;
    DBuf<fcw::LevelTile> fwd_tiles, bwd_tiles;
    std::vector<int> fwd_tile_off, bwd_tile_off, level_off;
    DBuf<double> panel_heat, panel_pois;
    // block work arrays
    DBuf<double> X, Y, Z, U, Dt, Dout;
    DBuf<int> src_pos;
    // metrics
    bool metrics_ready = false;
    int S = 0, q = 0, nb = 0, band_cap = 16384;
    double eps = 1e-6, health_h = 0.0, width = 0.0, cap = -1.0, rtol = 1e-10;
    int max_iter = 60;
    DBuf<double> targets, fracs, vw;
    DBuf<int> boundary, band;
    DBuf<double> o_scal, o_radii, o_perim, o_er, o_ea;
    DBuf<float> o_sr, o_sa;
    size_t o_cap = 0;
    // timing
    double stage[5] = {0, 0, 0, 0, 0};
    bool profiling = false;
    double ktime[3] = {0, 0, 0};
    int64_t kcount[3] = {0, 0, 0};
    std::vector<EventPair> ev_pool;
    std::vector<std::pair<int, EventPair>> ev_pending;
    size_t ev_used = 0;

    ~fcw_ctx() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& e : ev_pool) {
            (void)hipEventDestroy(e.a);
            (void)hipEventDestroy(e.b);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }

    fcw::MeshDev mesh_dev() const {
        fcw::MeshDev md;
        md.n = n;
        md.m = m;
        md.V = V.p;
        md.faces = faces.p;
        md.area = area.p;
        md.gb = gb.p;
        md.larea = larea.p;
        md.g11 = g11.p;
        md.g12 = g12.p;
        md.g22 = g22.p;
        md.vf_ptr = vf_ptr.p;
        md.vf = vf.p;
        return md;
    }

    EventPair* ev_begin(int which) {
        if (!profiling) return nullptr;
        if (ev_used == ev_pool.size()) {
            EventPair e;
            HIPCHK(hipEventCreate(&e.a));
            HIPCHK(hipEventCreate(&e.b));
            ev_pool.push_back(e);
        }
        EventPair* e = &ev_pool[ev_used++];
        HIPCHK(hipEventRecord(e->a, stream));
        ev_pending.push_back({which, *e});
        return e;
    }
    void ev_end(EventPair* e) {
        if (e) HIPCHK(hipEventRecord(e->b, stream));
    }
    void ev_collect() {
        if (ev_pending.empty()) return;
        HIPCHK(hipStreamSynchronize(stream));
        for (auto& pe : ev_pending) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, pe.second.a, pe.second.b));
            ktime[pe.first] += ms;
            kcount[pe.first] += 1;
        }
        ev_pending.clear();
        ev_used = 0;
    }

    // One system: forward then backward sweep. X (RHS, overwritten) -> Z (solution), Y scratch.
    void solve(const double* panel, double* Xb, double* Yb, double* Zb) {
        const int L = sym.nlevels;
        for (int l = 0; l < L; ++l) {
            if (l > 0)
                fcw::launch_fwd_assemble(stream, level_nodes.p + level_off[l], level_off[l + 1] - level_off[l], ldk,
                                         node_start.p, node_s.p, node_b.p, node_boff.p, node_uoff.p, child_ptr.p,
                                         child_list.p, cmap.p, Xb, U.p);
            EventPair* e = ev_begin(0);
            fcw::launch_fwd_gemm(stream, fwd_tiles.p + fwd_tile_off[l], fwd_tile_off[l + 1] - fwd_tile_off[l], ldk,
                                 node_start.p, node_s.p, node_b.p, node_poff.p, node_uoff.p, panel, Xb, Yb, U.p,
                                 l > 0);
            ev_end(e);
        }
        for (int l = L - 1; l >= 0; --l) {
            EventPair* e = ev_begin(1);
            fcw::launch_bwd_gemm(stream, bwd_tiles.p + bwd_tile_off[l], bwd_tile_off[l + 1] - bwd_tile_off[l], ldk,
                                 node_start.p, node_s.p, node_b.p, node_boff.p, bidx.p, node_poff.p, panel, Yb, Zb);
            ev_end(e);
        }
    }

    // Distances for the k sources whose positions are already in src_pos[0:k]; result in Dt (k x n).
    void distances(int k, double* st_times) {
        hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
        if (st_times) {
            HIPCHK(hipEventCreate(&e0));
            HIPCHK(hipEventCreate(&e1));
            HIPCHK(hipEventCreate(&e2));
            HIPCHK(hipEventCreate(&e3));
            HIPCHK(hipEventRecord(e0, stream));
        }
        fcw::launch_delta_rhs(stream, X.p, n, ldk, src_pos.p, k);
        solve(panel_heat.p, X.p, Y.p, Z.p);
        if (st_times) HIPCHK(hipEventRecord(e1, stream));
        fcw::launch_divergence(stream, mesh_dev(), ldk, Z.p, X.p);
        if (st_times) HIPCHK(hipEventRecord(e2, stream));
        solve(panel_pois.p, X.p, Y.p, Z.p);
        fcw::launch_shift_transpose(stream, n, ldk, k, Z.p, src_pos.p, Dt.p);
        if (st_times) {
            HIPCHK(hipEventRecord(e3, stream));
            HIPCHK(hipEventSynchronize(e3));
            float a = 0, b = 0, c = 0;
            HIPCHK(hipEventElapsedTime(&a, e0, e1));
            HIPCHK(hipEventElapsedTime(&b, e1, e2));
            HIPCHK(hipEventElapsedTime(&c, e2, e3));
            st_times[0] += a * 1e-3;
            st_times[1] += b * 1e-3;
            st_times[2] += c * 1e-3;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            (void)hipEventDestroy(e2);
            (void)hipEventDestroy(e3);
        }
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

    fcw::MetricsDev metrics_dev() {
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
        md.band_cap = band_cap;
        md.band = band.p;
        return md;
    }

    fcw::MetricsOut outputs_at(size_t row) {
        const size_t NE = S > 0 ? (size_t)(S - 1) * q : 0;
        fcw::MetricsOut o;
        o.scal = o_scal.p + row * fcw::kSlotCount;
        o.radii = o_radii.p + row * S;
        o.perim = o_perim.p + row * S;
        o.extra_r = o_er.p + row * NE;
        o.extra_a = o_ea.p + row * NE;
        o.samp_r = o_sr.p + row * (S + NE);
        o.samp_a = o_sa.p + row * (S + NE);
        return o;
    }

    void copy_outputs(size_t k, fcw_metrics_out* out) {
        const size_t NE = S > 0 ? (size_t)(S - 1) * q : 0;
        auto cp = [&](void* dst, const void* src, size_t bytes) {
            if (dst && bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
        };
        cp(out->scalars, o_scal.p, k * fcw::kSlotCount * sizeof(double));
        cp(out->radii, o_radii.p, k * S * sizeof(double));
        cp(out->perimeters, o_perim.p, k * S * sizeof(double));
        cp(out->extra_r, o_er.p, k * NE * sizeof(double));
        cp(out->extra_a, o_ea.p, k * NE * sizeof(double));
        cp(out->sample_r, o_sr.p, k * (S + NE) * sizeof(float));
        cp(out->sample_a, o_sa.p, k * (S + NE) * sizeof(float));
        HIPCHK(hipStreamSynchronize(stream));
    }
};

namespace {

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
    opt->ir_steps = 2;
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
        if (opt.precision != FCW_FP64 && opt.precision != FCW_FP32_IR) throw Fail("unknown precision");
        if (opt.precision == FCW_FP32_IR) throw Fail("FCW_FP32_IR is not available in this build");
        if (opt.block <= 0) opt.block = 256;
        if (!(opt.poisson_shift > 0.0) || !std::isfinite(opt.poisson_shift))
            throw Fail("poisson_shift must be positive and finite");
        check_faces(n, Vin, m, F);
        std::unique_ptr<fcw_ctx> C(new fcw_ctx());
        C->opt = opt;
        C->n = n;
        C->m = m;
        C->ldk = ((opt.block + fcw::kColPad - 1) / fcw::kColPad) * fcw::kColPad;
        HIPCHK(hipSetDevice(opt.device));
        HIPCHK(hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking));
        double t0 = now_s();
        std::string err;
        fcw::Symbolic& S = C->sym;
        if (fcw::analyse(n, Vin, m, F, opt.leaf_size > 0 ? opt.leaf_size : 64, S, err)) throw Fail(err);
        double t1 = now_s();
        // positions
        std::vector<double> Vp((size_t)n * 3);
        for (int p = 0; p < n; ++p)
            for (int c = 0; c < 3; ++c) Vp[3 * (size_t)p + c] = Vin[3 * (size_t)S.perm[p] + c];
        const double mean_edge = mean_edge_length(S, Vp.data());
        const double t = (opt.t > 0.0) ? opt.t : opt.t_coef * mean_edge * mean_edge;
        if (!(t > 0.0) || !std::isfinite(t)) throw Fail("diffusion time must be positive and finite");
        int64_t& tb = C->dev_bytes;
        C->V.upload(Vp.data(), Vp.size(), tb);
        C->faces.upload(S.faces.data(), S.faces.size(), tb);
        C->vf_ptr.upload(S.vf_ptr.data(), S.vf_ptr.size(), tb);
        C->vf.upload(S.vf.data(), S.vf.size(), tb);
        C->rowptr.upload(S.rowptr.data(), S.rowptr.size(), tb);
        C->colind.upload(S.colind.data(), S.colind.size(), tb);
        C->cptr.upload(S.contrib_ptr.data(), S.contrib_ptr.size(), tb);
        C->contrib.upload(S.contrib.data(), S.contrib.size(), tb);
        C->iperm.upload(S.iperm.data(), S.iperm.size(), tb);
        C->area.alloc(m, tb);
        C->gb.alloc((size_t)m * 9, tb);
        C->larea.alloc(m, tb);
        C->g11.alloc(m, tb);
        C->g12.alloc(m, tb);
        C->g22.alloc(m, tb);
        const int nnz = S.rowptr[n];
        C->lval.alloc(nnz, tb);
        C->mass.alloc(n, tb);
        fcw::launch_face_geometry(C->stream, m, C->V.p, C->faces.p, C->area.p, C->gb.p, C->larea.p, C->g11.p,
                                  C->g12.p, C->g22.p);
        fcw::launch_assemble(C->stream, n, nnz, C->cptr.p, C->contrib.p, C->area.p, C->gb.p, C->vf_ptr.p, C->vf.p,
                             C->lval.p, C->mass.p);
        HIPCHK(hipGetLastError());
        std::vector<double> lval(nnz), mass(n);
        HIPCHK(hipMemcpyAsync(lval.data(), C->lval.p, sizeof(double) * nnz, hipMemcpyDeviceToHost, C->stream));
        HIPCHK(hipMemcpyAsync(mass.data(), C->mass.p, sizeof(double) * n, hipMemcpyDeviceToHost, C->stream));
        HIPCHK(hipStreamSynchronize(C->stream));
        double t2 = now_s();
        std::vector<double> ph((size_t)S.panel_elems), pp((size_t)S.panel_elems);
        if (fcw::factor(S, lval.data(), mass.data(), t, 1.0, ph.data(), err, opt.host_threads)) throw Fail("heat: " + err);
        if (fcw::factor(S, lval.data(), mass.data(), 1.0, opt.poisson_shift, pp.data(), err, opt.host_threads))
            throw Fail("poisson: " + err);
        double t3 = now_s();
        C->panel_heat.upload(ph.data(), ph.size(), tb);
        C->panel_pois.upload(pp.data(), pp.size(), tb);
        // tree arrays and level tile lists
        C->node_start.upload(S.node_start.data(), S.node_start.size(), tb);
        C->node_s.upload(S.node_s.data(), S.node_s.size(), tb);
        C->node_b.upload(S.node_b.data(), S.node_b.size(), tb);
        C->bidx.upload(S.bidx.data(), std::max<size_t>(1, S.bidx.size()), tb);
        C->cmap.upload(S.cmap.data(), std::max<size_t>(1, S.cmap.size()), tb);
        C->child_ptr.upload(S.child_ptr.data(), S.child_ptr.size(), tb);
        C->child_list.upload(S.child_list.data(), std::max<size_t>(1, S.child_list.size()), tb);
        C->level_nodes.upload(S.level_nodes.data(), S.level_nodes.size(), tb);
        C->node_boff.upload(S.node_boff.data(), S.node_boff.size(), tb);
        C->node_poff.upload(S.node_poff.data(), S.node_poff.size(), tb);
        C->node_uoff.upload(S.node_uoff.data(), S.node_uoff.size(), tb);
        std::vector<fcw::LevelTile> ft, bt;
        C->fwd_tile_off.assign(1, 0);
        C->bwd_tile_off.assign(1, 0);
        C->level_off.assign(S.level_ptr.begin(), S.level_ptr.end());
        for (int l = 0; l < S.nlevels; ++l) {
            for (int i = S.level_ptr[l]; i < S.level_ptr[l + 1]; ++i) {
                const int id = S.level_nodes[i];
                const int f = S.node_s[id] + S.node_b[id];
                for (int m0 = 0; m0 < f; m0 += fcw::kTileM) ft.push_back({id, m0});
                for (int m0 = 0; m0 < S.node_s[id]; m0 += fcw::kTileM) bt.push_back({id, m0});
            }
            C->fwd_tile_off.push_back((int)ft.size());
            C->bwd_tile_off.push_back((int)bt.size());
        }
        C->fwd_tiles.upload(ft.data(), std::max<size_t>(1, ft.size()), tb);
        C->bwd_tiles.upload(bt.data(), std::max<size_t>(1, bt.size()), tb);
        const size_t ld = (size_t)C->ldk;
        C->X.alloc((size_t)n * ld, tb);
        C->Y.alloc((size_t)n * ld, tb);
        C->Z.alloc((size_t)n * ld, tb);
        C->U.alloc(std::max<size_t>(1, (size_t)S.urows * ld), tb);
        C->Dt.alloc((size_t)n * ld, tb);
        C->src_pos.alloc(ld, tb);
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

static void upload_sources(fcw_ctx* C, const int* src, int k) {
    if (k <= 0 || k > C->ldk) throw Fail("source count must be in 1..block (" + std::to_string(C->ldk) + ")");
    std::vector<int> pos(k);
    for (int j = 0; j < k; ++j) {
        if (src[j] < 0 || src[j] >= C->n)
            throw Fail("source_idx out of range: " + std::to_string(src[j]) + " for mesh with " +
                       std::to_string(C->n) + " vertices");
        pos[j] = C->sym.iperm[src[j]];
    }
    HIPCHK(hipMemcpyAsync(C->src_pos.p, pos.data(), sizeof(int) * k, hipMemcpyHostToDevice, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
}

int fcw_distance_block(fcw_ctx* C, const int* src, int k, double* out) {
    return guard([&] {
        if (!C || !src || !out) throw Fail("null argument");
        upload_sources(C, src, k);
        C->distances(k, C->profiling ? C->stage : nullptr);
        if (!C->Dout.p) C->Dout.alloc((size_t)C->n * C->ldk, C->dev_bytes);
        fcw::launch_gather_rows(C->stream, C->n, k, C->Dt.p, C->iperm.p, C->Dout.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, C->Dout.p, sizeof(double) * (size_t)k * C->n, hipMemcpyDeviceToHost, C->stream));
        HIPCHK(hipStreamSynchronize(C->stream));
        C->ev_collect();
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
        std::vector<double> w(C->n);
        for (int v = 0; v < C->n; ++v) w[C->sym.iperm[v]] = vertex_w[v];
        C->vw.upload(w.data(), w.size(), tb);
        C->eps = eps;
        C->health_h = health_h;
        C->width = bin_width;
        C->cap = cap_fraction;
        C->rtol = rtol;
        C->max_iter = max_iter;
        C->band.alloc((size_t)C->ldk * S * C->band_cap, tb);
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
        std::vector<int> pos(nsrc);
        for (int j = 0; j < nsrc; ++j) {
            if (src[j] < 0 || src[j] >= C->n)
                throw Fail("source_idx out of range: " + std::to_string(src[j]) + " for mesh with " +
                           std::to_string(C->n) + " vertices");
            pos[j] = C->sym.iperm[src[j]];
        }
        DBuf<int> all_pos;
        int64_t dummy = 0;
        all_pos.upload(pos.data(), pos.size(), dummy);
        C->ensure_outputs((size_t)nsrc);
        const fcw::MeshDev mdv = C->mesh_dev();
        const fcw::MetricsDev met = C->metrics_dev();
        for (int b0 = 0; b0 < nsrc; b0 += C->ldk) {
            const int k = std::min(C->ldk, nsrc - b0);
            HIPCHK(hipMemcpyAsync(C->src_pos.p, all_pos.p + b0, sizeof(int) * k, hipMemcpyDeviceToDevice, C->stream));
            C->distances(k, C->profiling ? C->stage : nullptr);
            hipEvent_t ea = nullptr, eb = nullptr;
            EventPair* e = C->ev_begin(2);
            if (C->profiling) {
                HIPCHK(hipEventCreate(&ea));
                HIPCHK(hipEventCreate(&eb));
                HIPCHK(hipEventRecord(ea, C->stream));
            }
            fcw::launch_metrics(C->stream, k, mdv, met, C->Dt.p, C->src_pos.p, C->outputs_at((size_t)b0), true);
            C->ev_end(e);
            if (C->profiling) {
                HIPCHK(hipEventRecord(eb, C->stream));
                HIPCHK(hipEventSynchronize(eb));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, ea, eb));
                C->stage[3] += ms * 1e-3;
                (void)hipEventDestroy(ea);
                (void)hipEventDestroy(eb);
                C->ev_collect();
            }
            HIPCHK(hipGetLastError());
        }
        C->copy_outputs((size_t)nsrc, out);
        C->ev_collect();
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
        upload_sources(C, src ? src : zeros.data(), k);
        // fields to positions: Dt[j][p] = D[j][perm[p]]
        std::vector<double> Dp((size_t)k * C->n);
        for (int j = 0; j < k; ++j)
            for (int p = 0; p < C->n; ++p) Dp[(size_t)j * C->n + p] = D[(size_t)j * C->n + C->sym.perm[p]];
        HIPCHK(hipMemcpyAsync(C->Dt.p, Dp.data(), sizeof(double) * Dp.size(), hipMemcpyHostToDevice, C->stream));
        C->ensure_outputs((size_t)k);
        fcw::launch_metrics(C->stream, k, C->mesh_dev(), C->metrics_dev(), C->Dt.p, C->src_pos.p, C->outputs_at(0),
                            health_screen != 0);
        HIPCHK(hipGetLastError());
        C->copy_outputs((size_t)k, out);
    });
}

int fcw_stage_times(fcw_ctx* C, double* out, int len) {
    if (!C || !out) return 0;
    int c = std::min(len, 5);
    for (int i = 0; i < c; ++i) out[i] = C->stage[i];
    return c;
}

int fcw_set_profiling(fcw_ctx* C, int on) {
    return guard([&] {
        if (!C) throw Fail("null context");
        C->profiling = on != 0;
        for (int i = 0; i < 3; ++i) {
            C->ktime[i] = 0.0;
            C->kcount[i] = 0;
        }
        for (int i = 0; i < 5; ++i) C->stage[i] = 0.0;
    });
}

int fcw_kernel_times(fcw_ctx* C, double* out, int len) {
    if (!C || !out) return 0;
    int c = 0;
    for (int i = 0; i < 3 && c + 1 < len + 1; ++i) {
        if (c < len) out[c++] = C->ktime[i];
        if (c < len) out[c++] = (double)C->kcount[i];
    }
    return c;
}

int fcw_host_factor_solve(const double* V, int n, const int* F, int m, double l_scale, double m_scale, int leaf_size,
                          const double* B, int k, double* X, double* stats) {
    return guard([&] {
        check_faces(n, V, m, F);
        fcw::Symbolic S;
        std::string err;
        if (fcw::analyse(n, V, m, F, leaf_size > 0 ? leaf_size : 64, S, err)) throw Fail(err);
        std::vector<std::vector<std::pair<int, double>>> rows;
        std::vector<double> mass;
        host_operators(n, V, m, F, rows, mass);
        // values in S's CSR (positions)
        std::vector<double> lval(S.rowptr[n], 0.0), mpos(n);
        for (int v = 0; v < n; ++v) {
            const int p = S.iperm[v];
            mpos[p] = mass[v];
            const int* row = S.colind.data() + S.rowptr[p];
            const int len = S.rowptr[p + 1] - S.rowptr[p];
            for (auto& e : rows[v]) {
                const int q = S.iperm[e.first];
                const int pos = (int)(std::lower_bound(row, row + len, q) - row);
                lval[S.rowptr[p] + pos] += e.second;
            }
        }
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
