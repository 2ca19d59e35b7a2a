// Mesh preparation on the device: cortical submesh, face geometry, vertex areas, boundary vertices,
// adjacency and the breadth-first source order (core_analysis.py:651-683, 859-978, 1336-1417).
//
// Every output equals the host restatement bit for bit: the floating-point expressions are
// evaluated in the same order with contraction off, sums are taken in the host's order, and
// every list is built in a deterministic order (per-vertex lists are filled through atomics and
// then sorted, so the fill order does not matter). The breadth-first order is the FIFO order of a
// queue over ascending neighbour lists, restarted at the smallest unvisited vertex: level by level,
// each vertex of the next level is claimed by the earliest (frontier position, neighbour position)
// that reaches it, and each frontier vertex then writes its claimed neighbours in order after the
// previous frontier vertices' claims.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "fcw_device.h"

#pragma clang fp contract(off)

namespace fcw {
namespace {

constexpr int kScanBlock = 1024;

inline dim3 blocks_for(int64_t n, int bs) { return dim3((unsigned)std::max<int64_t>(1, (n + bs - 1) / bs)); }

// Exclusive scan, three passes: block sums, a scan of the block sums, block-local scans plus offsets.
__global__ void scan_block_sums(const int* __restrict__ in, int n, int* __restrict__ sums) {
    __shared__ int s[kScanBlock / 64];
    const int i = blockIdx.x * kScanBlock + threadIdx.x;
    int v = i < n ? in[i] : 0;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kScanBlock / 64; ++w) t += s[w];
        sums[blockIdx.x] = t;
    }
}

__global__ void scan_sums_serial(int* __restrict__ sums, int nb, int* __restrict__ total) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int run = 0;
    for (int b = 0; b < nb; ++b) {
        const int v = sums[b];
        sums[b] = run;
        run += v;
    }
    *total = run;
}

__global__ void scan_apply(const int* __restrict__ in, int n, const int* __restrict__ offs, int* __restrict__ out) {
    __shared__ int s[kScanBlock];
    const int i = blockIdx.x * kScanBlock + threadIdx.x;
    s[threadIdx.x] = i < n ? in[i] : 0;
    __syncthreads();
    for (int o = 1; o < kScanBlock; o <<= 1) {  // inclusive Hillis-Steele within the block
        const int v = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    if (i < n) out[i] = offs[blockIdx.x] + s[threadIdx.x] - (i < n ? in[i] : 0);
}

// Faces kept: all three vertices in the mask and 0.5 |(p1 - p0) x (p2 - p0)| > 1e-12
// (core_analysis.py:859-911), the norm as sqrt((x^2 + y^2) + z^2).
__global__ void keep_faces_kernel(int m, const int* __restrict__ F, const double* __restrict__ V,
                                  const unsigned char* __restrict__ mask, int* __restrict__ keep) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m) return;
    const int a = F[3 * f], b = F[3 * f + 1], c = F[3 * f + 2];
    int k = 0;
    if (mask[a] && mask[b] && mask[c]) {
        const double* p0 = V + 3 * (int64_t)a;
        const double* p1 = V + 3 * (int64_t)b;
        const double* p2 = V + 3 * (int64_t)c;
        const double u0 = p1[0] - p0[0], u1 = p1[1] - p0[1], u2 = p1[2] - p0[2];
        const double w0 = p2[0] - p0[0], w1 = p2[1] - p0[1], w2 = p2[2] - p0[2];
        const double x = u1 * w2 - u2 * w1, y = u2 * w0 - u0 * w2, z = u0 * w1 - u1 * w0;
        const double area = 0.5 * sqrt((x * x + y * y) + z * z);
        k = area > 1e-12 ? 1 : 0;
    }
    keep[f] = k;
}

__global__ void mark_used_kernel(int m, const int* __restrict__ F, const int* __restrict__ keep, int* __restrict__ used) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m || !keep[f]) return;
    used[F[3 * f]] = 1;
    used[F[3 * f + 1]] = 1;
    used[F[3 * f + 2]] = 1;
}

__global__ void vertex_maps_kernel(int n, const int* __restrict__ used, const int* __restrict__ vpos,
                                   const double* __restrict__ V, int* __restrict__ sub_to_orig,
                                   int* __restrict__ orig_to_sub, double* __restrict__ Vs) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    if (used[v]) {
        const int p = vpos[v];
        sub_to_orig[p] = v;
        orig_to_sub[v] = p;
        for (int c = 0; c < 3; ++c) Vs[3 * (int64_t)p + c] = V[3 * (int64_t)v + c];
    } else {
        orig_to_sub[v] = -1;
    }
}

__global__ void sub_faces_kernel(int m, const int* __restrict__ F, const int* __restrict__ keep,
                                 const int* __restrict__ fpos, const int* __restrict__ orig_to_sub, int* __restrict__ Fs) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m || !keep[f]) return;
    const int p = fpos[f];
    for (int c = 0; c < 3; ++c) Fs[3 * (int64_t)p + c] = orig_to_sub[F[3 * f + c]];
}

__device__ inline double norm3(double x, double y, double z) { return sqrt((x * x + y * y) + z * z); }

// Face areas (0 below the scale-aware threshold eps * 10 * L^2) and longest edges L
// (core_analysis.py:913-978).
__global__ void face_geometry_prep_kernel(int m, const int* __restrict__ F, const double* __restrict__ V,
                                          double* __restrict__ areas, double* __restrict__ L) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m) return;
    const double* p0 = V + 3 * (int64_t)F[3 * f];
    const double* p1 = V + 3 * (int64_t)F[3 * f + 1];
    const double* p2 = V + 3 * (int64_t)F[3 * f + 2];
    const double a0 = p1[0] - p0[0], a1 = p1[1] - p0[1], a2 = p1[2] - p0[2];  // e01
    const double b0 = p2[0] - p1[0], b1 = p2[1] - p1[1], b2 = p2[2] - p1[2];  // e12
    const double c0 = p0[0] - p2[0], c1 = p0[1] - p2[1], c2 = p0[2] - p2[2];  // e20
    const double w0 = p2[0] - p0[0], w1 = p2[1] - p0[1], w2 = p2[2] - p0[2];
    const double x = a1 * w2 - a2 * w1, y = a2 * w0 - a0 * w2, z = a0 * w1 - a1 * w0;
    const double nn = norm3(x, y, z);
    const double Lf = fmax(fmax(norm3(a0, a1, a2), norm3(b0, b1, b2)), norm3(c0, c1, c2));
    const double tiny = 2.220446049250313e-16 * (Lf * Lf) * 10.0;
    areas[f] = nn < tiny ? 0.0 : 0.5 * nn;
    L[f] = Lf;
}

// Per-vertex lists: count, then fill through atomic cursors (the order is fixed by sorting later).
__global__ void count_incidences_kernel(int m, const int* __restrict__ F, int* __restrict__ ic, int* __restrict__ hc) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m) return;
    for (int c = 0; c < 3; ++c) {
        atomicAdd(&ic[F[3 * f + c]], 1);
        atomicAdd(&hc[F[3 * f + c]], 2);  // two directed half-edges leave each corner
    }
}

__global__ void fill_lists_kernel(int m, const int* __restrict__ F, const int* __restrict__ ioff, int* __restrict__ icur,
                                  int* __restrict__ ikey, const int* __restrict__ hoff, int* __restrict__ hcur,
                                  int* __restrict__ hnb) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m) return;
    const int v[3] = {F[3 * f], F[3 * f + 1], F[3 * f + 2]};
    for (int c = 0; c < 3; ++c) {
        const int a = v[c];
        ikey[ioff[a] + atomicAdd(&icur[a], 1)] = c * m + f;  // host order: corner pass, then face
        hnb[hoff[a] + atomicAdd(&hcur[a], 1)] = v[(c + 1) % 3];
        hnb[hoff[a] + atomicAdd(&hcur[a], 1)] = v[(c + 2) % 3];
    }
}

__device__ inline void insertion_sort(int* x, int k) {
    for (int i = 1; i < k; ++i) {
        const int v = x[i];
        int j = i - 1;
        while (j >= 0 && x[j] > v) {
            x[j + 1] = x[j];
            --j;
        }
        x[j + 1] = v;
    }
}

// Vertex areas as np.add.at(va, F[:, c], areas / 3) for c = 0, 1, 2: each vertex sums its
// incidences in (corner, face) order. Neighbours: sorted, each undirected edge counted once per
// face holding it; a vertex on an edge held by one face is a boundary vertex (core_analysis.py:651-665).
__global__ void vertex_lists_kernel(int n, int m, const int* __restrict__ ioff, int* __restrict__ ikey,
                                    const double* __restrict__ areas, double* __restrict__ va,
                                    const int* __restrict__ hoff, int* __restrict__ hnb, int* __restrict__ ucnt,
                                    int* __restrict__ bflag) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int* K = ikey + ioff[v];
    const int nk = ioff[v + 1] - ioff[v];
    insertion_sort(K, nk);
    double s = 0.0;
    for (int i = 0; i < nk; ++i) s += areas[K[i] % m] / 3.0;
    va[v] = s;
    int* N = hnb + hoff[v];
    const int nh = hoff[v + 1] - hoff[v];
    insertion_sort(N, nh);
    int u = 0, bnd = 0;
    for (int i = 0; i < nh;) {
        int j = i + 1;
        while (j < nh && N[j] == N[i]) ++j;
        bnd |= (j - i) == 1;  // the edge (v, N[i]) lies in one face
        ++u;
        i = j;
    }
    ucnt[v] = u;
    bflag[v] = bnd;
}

__global__ void unique_neighbours_kernel(int n, const int* __restrict__ hoff, const int* __restrict__ hnb,
                                         const int* __restrict__ uoff, int* __restrict__ cols) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int* N = hnb + hoff[v];
    const int nh = hoff[v + 1] - hoff[v];
    int o = uoff[v];
    for (int i = 0; i < nh; ++i)
        if (i == 0 || N[i] != N[i - 1]) cols[o++] = N[i];
}

__global__ void compact_flags_kernel(int n, const int* __restrict__ flag, const int* __restrict__ pos, int* __restrict__ out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n && flag[v]) out[pos[v]] = v;
}

// Breadth-first levels. The frontier is order[f0 .. f0 + fn).
__global__ void bfs_claim_kernel(const int* __restrict__ order, int f0, int fn, const int* __restrict__ indptr,
                                 const int* __restrict__ cols, const int* __restrict__ visited,
                                 unsigned long long* __restrict__ claim) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fn) return;
    const int u = order[f0 + i];
    for (int p = indptr[u]; p < indptr[u + 1]; ++p) {
        const int w = cols[p];
        if (!visited[w]) atomicMin(&claim[w], ((unsigned long long)i << 32) | (unsigned)(p - indptr[u]));
    }
}

__global__ void bfs_count_kernel(const int* __restrict__ order, int f0, int fn, const int* __restrict__ indptr,
                                 const int* __restrict__ cols, const int* __restrict__ visited,
                                 const unsigned long long* __restrict__ claim, int* __restrict__ cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fn) return;
    const int u = order[f0 + i];
    int c = 0;
    for (int p = indptr[u]; p < indptr[u + 1]; ++p) {
        const int w = cols[p];
        c += !visited[w] && claim[w] == (((unsigned long long)i << 32) | (unsigned)(p - indptr[u]));
    }
    cnt[i] = c;
}

__global__ void bfs_emit_kernel(int* __restrict__ order, int f0, int fn, const int* __restrict__ indptr,
                                const int* __restrict__ cols, int* __restrict__ visited,
                                unsigned long long* __restrict__ claim, const int* __restrict__ off) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fn) return;
    const int u = order[f0 + i];
    int o = f0 + fn + off[i];
    for (int p = indptr[u]; p < indptr[u + 1]; ++p) {
        const int w = cols[p];
        if (!visited[w] && claim[w] == (((unsigned long long)i << 32) | (unsigned)(p - indptr[u]))) order[o++] = w;
    }
}

__global__ void bfs_mark_kernel(const int* __restrict__ order, int b0, int bn, int* __restrict__ visited,
                                unsigned long long* __restrict__ claim) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= bn) return;
    const int w = order[b0 + i];
    visited[w] = 1;
    claim[w] = ~0ull;
}

__global__ void first_unvisited_kernel(int n, const int* __restrict__ visited, int* __restrict__ best) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n && !visited[v]) atomicMin(best, v);
}

__global__ void fill_int_kernel(int* __restrict__ p, int64_t n, int v) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void fill_u64_kernel(unsigned long long* __restrict__ p, int64_t n, unsigned long long v) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void start_component_kernel(int* __restrict__ order, int pos, int s, int* __restrict__ visited,
                                       unsigned long long* __restrict__ claim) {
    order[pos] = s;
    visited[s] = 1;
    claim[s] = ~0ull;
}

template <typename T>
struct Tmp {
    T* p = nullptr;
    Tmp() = default;
    explicit Tmp(size_t count) { alloc(count); }
    Tmp(const Tmp&) = delete;
    Tmp& operator=(const Tmp&) = delete;
    ~Tmp() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t count) {
        if (p) (void)hipFree(p);
        p = nullptr;
        if (hipMalloc((void**)&p, std::max<size_t>(1, count) * sizeof(T)) != hipSuccess) throw std::string("hipMalloc failed");
    }
};

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::string(what) + ": " + hipGetErrorString(e);
}

// out = exclusive scan of in[0:n]; returns the total. sums holds one int per 1024 elements.
int scan(hipStream_t st, const int* in, int n, int* out, int* total_dev, int* sums) {
    const int nb = std::max(1, (n + kScanBlock - 1) / kScanBlock);
    fcw::launch(scan_block_sums, dim3(nb), dim3(kScanBlock), 0, st, in, n, sums);
    fcw::launch(scan_sums_serial, dim3(1), dim3(1), 0, st, sums, nb, total_dev);
    fcw::launch(scan_apply, dim3(nb), dim3(kScanBlock), 0, st, in, n, sums, out);
    int total = 0;
    check(hipMemcpyAsync(&total, total_dev, sizeof(int), hipMemcpyDeviceToHost, st), "scan total");
    check(hipStreamSynchronize(st), "scan");
    return total;
}

}  // namespace

int mesh_prep_device(int n, const double* Vh, int m, const int* Fh, const unsigned char* maskh, int start, int device,
                     MeshPrepOut& out, std::string& err) {
    try {
        check(hipSetDevice(device), "hipSetDevice");
        hipStream_t st;
        check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } sg{st};
        const int bs = 256;
        Tmp<double> V(3 * (size_t)n);
        Tmp<int> F(3 * (size_t)m);
        Tmp<unsigned char> mask(n);
        check(hipMemcpyAsync(V.p, Vh, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st), "upload V");
        check(hipMemcpyAsync(F.p, Fh, sizeof(int) * 3 * (size_t)m, hipMemcpyHostToDevice, st), "upload F");
        check(hipMemcpyAsync(mask.p, maskh, n, hipMemcpyHostToDevice, st), "upload mask");
        Tmp<int> total(1), keep(m), fpos(m), used(n), vpos(n);
        Tmp<int> sums((std::max(n, m) + 1) / kScanBlock + 2);
        fcw::launch(keep_faces_kernel, dim3(blocks_for(m, bs)), dim3(bs), 0, st, m, F.p, V.p, mask.p, keep.p);
        const int ms = scan(st, keep.p, m, fpos.p, total.p, sums.p);
        fcw::launch(fill_int_kernel, dim3(blocks_for(n, bs)), dim3(bs), 0, st, used.p, n, 0);
        fcw::launch(mark_used_kernel, dim3(blocks_for(m, bs)), dim3(bs), 0, st, m, F.p, keep.p, used.p);
        const int ns = scan(st, used.p, n, vpos.p, total.p, sums.p);
        out.n_sub = ns;
        out.m_sub = ms;
        Tmp<int> s2o(ns), o2s(n), Fs(3 * (size_t)ms);
        Tmp<double> Vs(3 * (size_t)ns);
        fcw::launch(vertex_maps_kernel, dim3(blocks_for(n, bs)), dim3(bs), 0, st, n, used.p, vpos.p, V.p, s2o.p, o2s.p, Vs.p);
        fcw::launch(sub_faces_kernel, dim3(blocks_for(m, bs)), dim3(bs), 0, st, m, F.p, keep.p, fpos.p, o2s.p, Fs.p);
        // submesh geometry and per-vertex lists
        Tmp<double> areas(ms), L(ms), va(ns);
        fcw::launch(face_geometry_prep_kernel, dim3(blocks_for(ms, bs)), dim3(bs), 0, st, ms, Fs.p, Vs.p, areas.p, L.p);
        Tmp<int> ic(ns + 1), hc(ns + 1), ioff(ns + 1), hoff(ns + 1), icur(ns), hcur(ns);
        fcw::launch(fill_int_kernel, dim3(blocks_for(ns + 1, bs)), dim3(bs), 0, st, ic.p, ns + 1, 0);
        fcw::launch(fill_int_kernel, dim3(blocks_for(ns + 1, bs)), dim3(bs), 0, st, hc.p, ns + 1, 0);
        fcw::launch(fill_int_kernel, dim3(blocks_for(ns, bs)), dim3(bs), 0, st, icur.p, ns, 0);
        fcw::launch(fill_int_kernel, dim3(blocks_for(ns, bs)), dim3(bs), 0, st, hcur.p, ns, 0);
        fcw::launch(count_incidences_kernel, dim3(blocks_for(ms, bs)), dim3(bs), 0, st, ms, Fs.p, ic.p, hc.p);
        const int ni = scan(st, ic.p, ns + 1, ioff.p, total.p, sums.p);
        const int nh = scan(st, hc.p, ns + 1, hoff.p, total.p, sums.p);
        Tmp<int> ikey(ni), hnb(nh), ucnt(ns + 1), uoff(ns + 1), bflag(ns), bpos(ns);
        fcw::launch(fill_lists_kernel, dim3(blocks_for(ms, bs)), dim3(bs), 0, st, ms, Fs.p, ioff.p, icur.p, ikey.p, hoff.p, hcur.p, hnb.p);
        fcw::launch(fill_int_kernel, dim3(1), dim3(1), 0, st, ucnt.p + ns, 1, 0);
        fcw::launch(vertex_lists_kernel, dim3(blocks_for(ns, bs)), dim3(bs), 0, st, ns, ms, ioff.p, ikey.p, areas.p, va.p, hoff.p, hnb.p,
                                                               ucnt.p, bflag.p);
        const int nu = scan(st, ucnt.p, ns + 1, uoff.p, total.p, sums.p);
        Tmp<int> cols(nu), bnd(ns);
        fcw::launch(unique_neighbours_kernel, dim3(blocks_for(ns, bs)), dim3(bs), 0, st, ns, hoff.p, hnb.p, uoff.p, cols.p);
        const int nb = scan(st, bflag.p, ns, bpos.p, total.p, sums.p);
        fcw::launch(compact_flags_kernel, dim3(blocks_for(ns, bs)), dim3(bs), 0, st, ns, bflag.p, bpos.p, bnd.p);
        out.n_boundary = nb;
        out.n_adj = nu;
        // breadth-first order over every component
        Tmp<int> order(ns), visited(ns), cnt(std::max(1, ns)), off(std::max(1, ns)), best(1);
        Tmp<unsigned long long> claim(ns);
        fcw::launch(fill_int_kernel, dim3(blocks_for(ns, bs)), dim3(bs), 0, st, visited.p, ns, 0);
        fcw::launch(fill_u64_kernel, dim3(blocks_for(ns, bs)), dim3(bs), 0, st, claim.p, ns, ~0ull);
        int placed = 0;
        int s = (start >= 0 && start < ns) ? start : 0;
        while (placed < ns) {
            fcw::launch(start_component_kernel, dim3(1), dim3(1), 0, st, order.p, placed, s, visited.p, claim.p);
            int f0 = placed, fn = 1;
            placed += 1;
            while (fn > 0) {
                fcw::launch(bfs_claim_kernel, dim3(blocks_for(fn, bs)), dim3(bs), 0, st, order.p, f0, fn, uoff.p, cols.p, visited.p, claim.p);
                fcw::launch(bfs_count_kernel, dim3(blocks_for(fn, bs)), dim3(bs), 0, st, order.p, f0, fn, uoff.p, cols.p, visited.p, claim.p,
                                                                    cnt.p);
                const int nn = scan(st, cnt.p, fn, off.p, total.p, sums.p);
                if (nn > 0) {
                    fcw::launch(bfs_emit_kernel, dim3(blocks_for(fn, bs)), dim3(bs), 0, st, order.p, f0, fn, uoff.p, cols.p, visited.p,
                                                                       claim.p, off.p);
                    fcw::launch(bfs_mark_kernel, dim3(blocks_for(nn, bs)), dim3(bs), 0, st, order.p, f0 + fn, nn, visited.p, claim.p);
                }
                f0 += fn;
                fn = nn;
                placed += nn;
            }
            if (placed >= ns) break;
            int h_best = ns;
            check(hipMemcpyAsync(best.p, &h_best, sizeof(int), hipMemcpyHostToDevice, st), "best");
            fcw::launch(first_unvisited_kernel, dim3(blocks_for(ns, bs)), dim3(bs), 0, st, ns, visited.p, best.p);
            check(hipMemcpyAsync(&h_best, best.p, sizeof(int), hipMemcpyDeviceToHost, st), "best back");
            check(hipStreamSynchronize(st), "components");
            if (h_best >= ns) throw std::string("breadth-first order lost vertices");
            s = h_best;
        }
        check(hipGetLastError(), "kernels");
        // results to the caller's buffers
        check(hipMemcpyAsync(out.sub_to_orig, s2o.p, sizeof(int) * ns, hipMemcpyDeviceToHost, st), "s2o");
        check(hipMemcpyAsync(out.orig_to_sub, o2s.p, sizeof(int) * n, hipMemcpyDeviceToHost, st), "o2s");
        check(hipMemcpyAsync(out.faces, Fs.p, sizeof(int) * 3 * (size_t)ms, hipMemcpyDeviceToHost, st), "faces");
        check(hipMemcpyAsync(out.vertices, Vs.p, sizeof(double) * 3 * (size_t)ns, hipMemcpyDeviceToHost, st), "V");
        check(hipMemcpyAsync(out.boundary, bnd.p, sizeof(int) * nb, hipMemcpyDeviceToHost, st), "boundary");
        check(hipMemcpyAsync(out.order, order.p, sizeof(int) * ns, hipMemcpyDeviceToHost, st), "order");
        check(hipMemcpyAsync(out.face_areas, areas.p, sizeof(double) * ms, hipMemcpyDeviceToHost, st), "areas");
        check(hipMemcpyAsync(out.face_L, L.p, sizeof(double) * ms, hipMemcpyDeviceToHost, st), "L");
        check(hipMemcpyAsync(out.vertex_areas, va.p, sizeof(double) * ns, hipMemcpyDeviceToHost, st), "va");
        check(hipMemcpyAsync(out.adj_ptr, uoff.p, sizeof(int) * (ns + 1), hipMemcpyDeviceToHost, st), "adj ptr");
        check(hipMemcpyAsync(out.adj, cols.p, sizeof(int) * nu, hipMemcpyDeviceToHost, st), "adj");
        check(hipStreamSynchronize(st), "copy back");
        return 0;
    } catch (const std::string& e) {
        err = e;
        return 1;
    }
}

}  // namespace fcw
