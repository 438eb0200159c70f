// conn.hip -- connectivity on the resident CSR, gfx950.
//
// Replaces:
//   connected_components!, connected_components   src/connectivity.jl:13-95
//   neighborhood, neighborhood_dists               src/connectivity.jl:634-702
//     (and is_connected, :131-134, on its labels)
//   period                                         src/connectivity.jl:495-516
//     (the gcd over the non-tree edges of the BFS tree from vertex 1; the
//     BFS and the strong-connectivity test run on bfs.hip)
//
// Components.  connected_components! labels every vertex with the first
// vertex of its component in the order of the outer loop, which is the
// smallest vertex id of the component (all_neighbors: weak components for a
// directed graph).  Here a union-find over the edges computes the same
// labels: a hook always sets the larger of two roots to point at the
// smaller, so the smallest id of a component is never hooked and ends as its
// root.  Finds split paths with atomicMin (a pointer only ever moves to a
// smaller id of the same set); hooks are a compare-and-swap on a root.  The
// result does not depend on the order of the hooks.
//
// Neighborhood.  _neighborhood (connectivity.jl:685-702) with unit lengths
// returns the FIFO queue of a breadth-first search cut after level d: level
// 0 is v, and level L+1 lists, for each vertex u of level L in queue order,
// the not yet seen neighbours of u in adjacency order.  Here each level is
// built from the previous one without a queue: a vertex w first seen at
// level L+1 belongs to the earliest level-L vertex (by queue position) that
// has it as a neighbour (atomicMin of that position); every level-L vertex
// counts the vertices it owns, an exclusive scan of the counts gives each
// its output offset, and it writes them in adjacency order.
//
// Period.  For a strongly connected digraph with levels L from vertex 1,
// period(g) = gcd over the edges (s, d) outside the BFS tree of
// L[s] - L[d] + 1 (connectivity.jl:503-516).  A tree edge gives 0 and gcd(x,
// 0) = x, and a self-loop gives 1, so the gcd over every edge is the same
// number (0 when there is no edge outside the tree); each block reduces its
// edges and the host combines the blocks.
#include "lgmi_internal.h"

#include <algorithm>
#include <numeric>
#include <vector>

namespace lgmi {

namespace {

constexpr int kConnBlock = 256;
constexpr int32_t kCcHeavy = 64;   // rows longer than this take a wavefront

__device__ inline int32_t ld(const int32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Every write to `parent` while hooks run is a read-modify-write atomic (the
// hooks' compare-and-swap, the path splitting's atomicMin): measured on
// gfx950, plain (or atomic) stores into lines that other work-groups update
// with compare-and-swap lost some of those updates.  atomicMin keeps the
// invariant parent[x] <= x within x's set, so splitting stays valid when it
// races with a hook.
__device__ int32_t find_root(int32_t* parent, int32_t x) {
    int32_t p = ld(&parent[x]);
    while (p != x) {
        const int32_t gp = ld(&parent[p]);
        if (gp != p) atomicMin(&parent[x], gp);   // path splitting
        x = p;
        p = gp;
    }
    return x;
}

__device__ void unite(int32_t* parent, int32_t a, int32_t b) {
    for (;;) {
        a = find_root(parent, a);
        b = find_root(parent, b);
        if (a == b) return;
        if (a < b) {
            const int32_t t = a;
            a = b;
            b = t;
        }
        // a > b: hook the larger root under the smaller
        int32_t expected = a;
        if (__hip_atomic_compare_exchange_strong(&parent[a], &expected, b, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
    }
}

__global__ void cc_init(int32_t* parent, int32_t nv) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x) parent[v] = v;
}

// Light rows: one thread per vertex; an undirected edge is taken once (from
// its larger end), a directed edge once (from its tail).
__global__ __launch_bounds__(kConnBlock) void cc_hook_light(const int32_t* rp, const int32_t* col, int32_t nv,
                                                            int directed, int32_t* parent) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x) {
        const int32_t b = rp[v], e = rp[v + 1];
        if (e - b > kCcHeavy) continue;
        for (int32_t k = b; k < e; k++) {
            const int32_t u = col[k];
            if (directed ? u != v : u < v) unite(parent, v, u);
        }
    }
}

// Heavy rows: one wavefront per vertex, 64 entries at a time.
__global__ __launch_bounds__(kConnBlock) void cc_hook_heavy(const int32_t* rp, const int32_t* col, int32_t nv,
                                                            int directed, int32_t* parent) {
    const int lane = threadIdx.x & 63;
    const int32_t w0 = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nw = (gridDim.x * blockDim.x) >> 6;
    for (int32_t v = w0; v < nv; v += nw) {
        const int32_t b = rp[v], e = rp[v + 1];
        if (e - b <= kCcHeavy) continue;
        for (int32_t k = b + lane; k < e; k += 64) {
            const int32_t u = col[k];
            if (directed ? u != v : u < v) unite(parent, v, u);
        }
    }
}

__global__ void cc_finish(int32_t* parent, int32_t nv) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x) {
        const int32_t r = find_root(parent, v);
        atomicMin(&parent[v], r);
    }
}

__device__ inline int64_t gcd64(int64_t a, int64_t b) {
    while (b) {
        const int64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

// gcd over the edges (s, d) of L[s] - L[d] + 1, per block.
__global__ __launch_bounds__(kConnBlock) void period_gcd(const int32_t* rp, const int32_t* col, const int32_t* dist,
                                                         int32_t nv, int64_t* out) {
    __shared__ int64_t s_g[kConnBlock / 64];
    int64_t g = 0;
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x) {
        const int64_t ls = dist[v];
        for (int32_t k = rp[v]; k < rp[v + 1] && g != 1; k++) {
            int64_t x = ls - (int64_t)dist[col[k]] + 1;
            if (x < 0) x = -x;
            g = gcd64(g, x);
        }
    }
    for (int o = 32; o >= 1; o >>= 1) g = gcd64(g, (int64_t)__shfl_xor((long long)g, o));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) s_g[wid] = g;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t r = 0;
        for (int i = 0; i < kConnBlock / 64; i++) r = gcd64(r, s_g[i]);
        out[blockIdx.x] = r;
    }
}

// ---- neighborhood: one wavefront per vertex of the current level ----

// Queue positions (beg + i) are unique over all levels, so a claim left
// from an earlier level never matches a vertex of the current one.
__global__ __launch_bounds__(kConnBlock) void nb_claim(const int32_t* rp, const int32_t* col, const int32_t* lvl,
                                                       int32_t n, int32_t beg, const int32_t* dist, int32_t* first) {
    const int lane = threadIdx.x & 63;
    const int32_t w0 = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nw = (gridDim.x * blockDim.x) >> 6;
    for (int32_t i = w0; i < n; i += nw) {
        const int32_t u = lvl[i];
        for (int32_t k = rp[u] + lane; k < rp[u + 1]; k += 64) {
            const int32_t w = col[k];
            if (dist[w] < 0) atomicMin(&first[w], beg + i);
        }
    }
}

__global__ __launch_bounds__(kConnBlock) void nb_count(const int32_t* rp, const int32_t* col, const int32_t* lvl,
                                                       int32_t n, int32_t beg, const int32_t* first, int32_t* cnt) {
    const int lane = threadIdx.x & 63;
    const int32_t w0 = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nw = (gridDim.x * blockDim.x) >> 6;
    for (int32_t i = w0; i < n; i += nw) {
        const int32_t u = lvl[i];
        int32_t c = 0;
        for (int32_t off = rp[u]; off < rp[u + 1]; off += 64) {
            const int32_t k = off + lane;
            const bool own = k < rp[u + 1] && first[col[k]] == beg + i;
            c += __popcll(__ballot(own));
        }
        if (lane == 0) cnt[i] = c;
    }
}

__global__ __launch_bounds__(kConnBlock) void nb_write(const int32_t* rp, const int32_t* col, const int32_t* lvl,
                                                       int32_t n, int32_t beg, const int32_t* first,
                                                       const int32_t* off_out, int32_t* next, int32_t* dist,
                                                       int32_t level) {
    const int lane = threadIdx.x & 63;
    const int32_t w0 = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nw = (gridDim.x * blockDim.x) >> 6;
    for (int32_t i = w0; i < n; i += nw) {
        const int32_t u = lvl[i];
        int32_t pos = off_out[i];
        for (int32_t off = rp[u]; off < rp[u + 1]; off += 64) {
            const int32_t k = off + lane;
            int32_t w = -1;
            bool own = false;
            if (k < rp[u + 1]) {
                w = col[k];
                own = first[w] == beg + i;
            }
            const uint64_t m = __ballot(own);
            if (own) {
                const int32_t r = (int32_t)__popcll(m & ((1ull << lane) - 1));
                next[pos + r] = w;
                dist[w] = level;
            }
            pos += (int32_t)__popcll(m);
        }
    }
}

__global__ void nb_init(int32_t* dist, int32_t* first, int32_t nv, int32_t v, int32_t* q) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gridDim.x * blockDim.x) {
        dist[i] = i == v ? 0 : -1;
        first[i] = 0x7fffffff;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) q[0] = v;
}

// Exclusive scan of n int32 in blocks of 1024 (4 per thread); the block
// totals go to sums.
constexpr int kScanBlock = 256;
constexpr int kScanPer = 4;
constexpr int kScanTile = kScanBlock * kScanPer;

__global__ __launch_bounds__(kScanBlock) void scan_tiles(const int32_t* in, int32_t* out, int32_t n, int32_t* sums) {
    __shared__ int32_t s_wave[kScanBlock / 64];
    const int tid = threadIdx.x, lane = tid & 63, wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)tid * kScanPer;
    int32_t x[kScanPer];
    int32_t t = 0;
    for (int j = 0; j < kScanPer; j++) {
        x[j] = base + j < n ? in[base + j] : 0;
        t += x[j];
    }
    int32_t incl = t;
    for (int d = 1; d < 64; d <<= 1) {
        const int32_t y = __shfl_up(incl, d);
        if (lane >= d) incl += y;
    }
    if (lane == 63) s_wave[wid] = incl;
    __syncthreads();
    int32_t wave_off = 0;
    for (int w = 0; w < wid; w++) wave_off += s_wave[w];
    int32_t run = wave_off + incl - t;
    for (int j = 0; j < kScanPer; j++) {
        if (base + j < n) out[base + j] = run;
        run += x[j];
    }
    if (tid == kScanBlock - 1) sums[blockIdx.x] = wave_off + incl;
}

__global__ __launch_bounds__(kScanBlock) void scan_add(int32_t* out, int32_t n, const int32_t* tile_off) {
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    const int32_t add = tile_off[blockIdx.x];
    for (int j = threadIdx.x; j < kScanTile; j += kScanBlock)
        if (base + j < n) out[base + j] += add;
}

}  // namespace

// out[i] = in[0] + ... + in[i-1]; returns the total (device buffers, host
// synchronisation at the end).
int64_t exclusive_scan(hipStream_t s, const int32_t* in, int32_t* out, int32_t n) {
    if (n <= 0) return 0;
    const int32_t tiles = (n + kScanTile - 1) / kScanTile;
    int32_t* sums = dev_alloc_n<int32_t>((size_t)tiles);
    int32_t* offs = nullptr;
    try {
        scan_tiles<<<tiles, kScanBlock, 0, s>>>(in, out, n, sums);
        LGMI_HIP(hipGetLastError());
        if (tiles > 1) {
            offs = dev_alloc_n<int32_t>((size_t)tiles);
            exclusive_scan(s, sums, offs, tiles);
            scan_add<<<tiles, kScanBlock, 0, s>>>(out, n, offs);
            LGMI_HIP(hipGetLastError());
        }
        int32_t last_out = 0, last_in = 0;
        LGMI_HIP(hipMemcpyAsync(&last_out, out + (n - 1), sizeof(int32_t), hipMemcpyDeviceToHost, s));
        LGMI_HIP(hipMemcpyAsync(&last_in, in + (n - 1), sizeof(int32_t), hipMemcpyDeviceToHost, s));
        LGMI_HIP(hipStreamSynchronize(s));
        dev_free(sums);
        dev_free(offs);
        return (int64_t)last_out + last_in;
    } catch (...) {
        dev_free(sums);
        dev_free(offs);
        throw;
    }
}

int32_t neighborhood(lgmi_graph* g, int32_t v, int64_t d, int dir_in, int32_t* verts, int32_t* dists) {
    const int32_t nv = g->nv;
    if (d < 0) return 0;
    const DevCSR& c = dir_in ? g->in : g->out;
    hipStream_t s = g->stream;
    int32_t *dist = nullptr, *first = nullptr, *q = nullptr, *cnt = nullptr, *off = nullptr;
    int32_t total = 1;
    try {
        dist = dev_alloc_n<int32_t>((size_t)nv);
        first = dev_alloc_n<int32_t>((size_t)nv);
        q = dev_alloc_n<int32_t>((size_t)nv);
        cnt = dev_alloc_n<int32_t>((size_t)nv);
        off = dev_alloc_n<int32_t>((size_t)nv);
        const int grid = std::max(1, std::min(g->cu_count * 8, (nv + kConnBlock - 1) / kConnBlock));
        nb_init<<<grid, kConnBlock, 0, s>>>(dist, first, nv, v, q);
        LGMI_HIP(hipGetLastError());
        int32_t beg = 0, n = 1;
        for (int64_t level = 1; level <= d && n > 0; level++) {
            const int32_t* lvl = q + beg;
            const int wg = std::max(1, std::min(g->cu_count * 8, (n + kConnBlock / 64 - 1) / (kConnBlock / 64)));
            nb_claim<<<wg, kConnBlock, 0, s>>>(c.row_ptr, c.col, lvl, n, beg, dist, first);
            nb_count<<<wg, kConnBlock, 0, s>>>(c.row_ptr, c.col, lvl, n, beg, first, cnt);
            LGMI_HIP(hipGetLastError());
            const int64_t m = exclusive_scan(s, cnt, off, n);
            if (m == 0) break;
            nb_write<<<wg, kConnBlock, 0, s>>>(c.row_ptr, c.col, lvl, n, beg, first, off, q + beg + n, dist,
                                               (int32_t)level);
            LGMI_HIP(hipGetLastError());
            beg += n;
            n = (int32_t)m;
            total = beg + n;
        }
        LGMI_HIP(hipMemcpyAsync(verts, q, sizeof(int32_t) * (size_t)total, hipMemcpyDeviceToHost, s));
        if (dists) {
            // distances in queue order
            std::vector<int32_t> dv((size_t)nv);
            LGMI_HIP(hipMemcpyAsync(dv.data(), dist, sizeof(int32_t) * (size_t)nv, hipMemcpyDeviceToHost, s));
            LGMI_HIP(hipStreamSynchronize(s));
            for (int32_t i = 0; i < total; i++) dists[i] = dv[(size_t)verts[i]];
        }
        LGMI_HIP(hipStreamSynchronize(s));
    } catch (...) {
        dev_free(dist); dev_free(first); dev_free(q); dev_free(cnt); dev_free(off);
        throw;
    }
    dev_free(dist); dev_free(first); dev_free(q); dev_free(cnt); dev_free(off);
    return total;
}

void connected_components(lgmi_graph* g, int32_t* label) {
    const int32_t nv = g->nv;
    if (nv == 0) return;
    hipStream_t s = g->stream;
    int32_t* parent = dev_alloc_n<int32_t>((size_t)nv);
    try {
        const int grid = std::max(1, std::min(g->cu_count * 8, (nv + kConnBlock - 1) / kConnBlock));
        const int hgrid = std::max(1, std::min(g->cu_count * 8, (nv + kConnBlock / 64 - 1) / (kConnBlock / 64)));
        cc_init<<<grid, kConnBlock, 0, s>>>(parent, nv);
        cc_hook_light<<<grid, kConnBlock, 0, s>>>(g->out.row_ptr, g->out.col, nv, g->directed, parent);
        cc_hook_heavy<<<hgrid, kConnBlock, 0, s>>>(g->out.row_ptr, g->out.col, nv, g->directed, parent);
        cc_finish<<<grid, kConnBlock, 0, s>>>(parent, nv);
        LGMI_HIP(hipGetLastError());
        LGMI_HIP(hipMemcpyAsync(label, parent, sizeof(int32_t) * (size_t)nv, hipMemcpyDeviceToHost, s));
        LGMI_HIP(hipStreamSynchronize(s));
    } catch (...) {
        dev_free(parent);
        throw;
    }
    dev_free(parent);
}

int64_t period_from_levels(lgmi_graph* g, const int32_t* d_dist) {
    const int32_t nv = g->nv;
    if (nv == 0) return 0;
    hipStream_t s = g->stream;
    const int grid = std::max(1, std::min(g->cu_count * 4, (nv + kConnBlock - 1) / kConnBlock));
    int64_t* d_out = dev_alloc_n<int64_t>((size_t)grid);
    std::vector<int64_t> h((size_t)grid);
    try {
        period_gcd<<<grid, kConnBlock, 0, s>>>(g->out.row_ptr, g->out.col, d_dist, nv, d_out);
        LGMI_HIP(hipGetLastError());
        LGMI_HIP(hipMemcpyAsync(h.data(), d_out, sizeof(int64_t) * (size_t)grid, hipMemcpyDeviceToHost, s));
        LGMI_HIP(hipStreamSynchronize(s));
    } catch (...) {
        dev_free(d_out);
        throw;
    }
    dev_free(d_out);
    int64_t r = 0;
    for (int64_t x : h) r = std::gcd(r, x);
    return r;
}

}  // namespace lgmi
