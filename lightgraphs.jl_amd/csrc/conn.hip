// conn.hip -- connectivity on the resident CSR, gfx950.
//
// Replaces:
//   connected_components!, connected_components   src/connectivity.jl:13-95
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

}  // namespace

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
