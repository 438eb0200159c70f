// bc.hip -- exact Brandes betweenness, 64 sources per lane group, gfx950.
//
// Replaces, for unweighted graphs (distmx = DefaultDistance):
//   betweenness_centrality      src/centrality/betweenness.jl:46-78
//   dijkstra_shortest_paths(...; allpaths=true, trackvertices=true)
//                               src/shortestpaths/dijkstra.jl:59-141
//   _accumulate_basic!          betweenness.jl:81-107
//   _accumulate_endpoints!      betweenness.jl:109-134
//   _rescale!                   betweenness.jl:136-160
//   Parallel.betweenness_centrality (:distributed and :threads)
//                               src/Parallel/centrality/betweenness.jl:1-75
//
// With unit weights the reference's Dijkstra dequeues vertices in
// non-decreasing distance, so its path counts equal the breadth-first
// shortest-path counts.  Here 64 sources advance together: for every vertex
// a uint64 holds one bit per source lane (seen[v]; F_L[v] = lanes for which
// v is at distance L).  A wavefront takes one vertex at a time and its 64
// lanes are the 64 sources, so the FP64 rows sig[v][0..63] are read and
// written as 512-byte coalesced rows.
//
// Forward (level L -> L+1, pull over in-neighbours): for every vertex with
// an unseen lane, OR the frontier masks of its in-neighbours and, per lane,
// add sigma of the neighbours on the frontier.  Sigma is integer-valued in
// FP64, so the sum is exact whatever the order (below 2^53).
// Backward (level L = Lmax .. 1, pull over out-neighbours):
//   delta_i[v] = sigma_i[v] * sum_{w succ} (1 + delta_i[w]) / sigma_i[w]
// is formed with no atomics; the coefficient (1+delta)/sigma replaces sigma
// in the same slot once the vertex is final.  The per-vertex sum over lanes
// uses a fixed butterfly, and levels, slots and groups are added in a fixed
// order, so repeated runs are bit-identical.  Only the floating-point
// summation order differs from the reference (ulp level).
//
// Several lane groups ("slots") run in the same launches (blockIdx.y).
#include "lgmi_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace lgmi {

struct BcCtrl {
    int32_t level;
    int32_t done;
    int32_t maxlevel;
    int32_t found;
    uint32_t blocks_done;
    int32_t overflow;
    int32_t pad[2];
};

struct BcWork {
    int32_t nv = 0;
    int32_t nslots = 0;
    int32_t cap_levels = 0;
    uint64_t* seen = nullptr;    // [slot][nv]
    uint64_t* F = nullptr;       // [slot][cap][nv]
    double* sig = nullptr;       // [slot][nv][64]
    double* bc_acc = nullptr;    // [slot][nv]
    double* bc_out = nullptr;    // [nv]
    double* endp = nullptr;      // [nv] endpoints term
    uint64_t* active = nullptr;  // [slot]
    int32_t* lane_src = nullptr; // [slot][64]
    BcCtrl* ctrl = nullptr;      // [slot]
    BcCtrl* h_ctrl = nullptr;    // pinned mirror
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
};

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

struct BcArgs {
    const int32_t* in_rp;
    const int32_t* in_col;
    const int32_t* out_rp;
    const int32_t* out_col;
    int32_t nv;
    int32_t nchunks;
    int32_t cap;
    uint64_t* seen;
    uint64_t* F;
    double* sig;
    double* bc_acc;
    const uint64_t* active;
    BcCtrl* ctrl;
};

__global__ void bc_clear_acc(double* acc, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        acc[i] = 0.0;
}

// Clears seen and F_0 of every slot; one grid over nv, blockIdx.y = slot.
__global__ void bc_init(BcArgs a) {
    const int slot = blockIdx.y;
    uint64_t* seen = a.seen + (size_t)slot * a.nv;
    uint64_t* F0 = a.F + (size_t)slot * a.cap * a.nv;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < a.nv; v += (int64_t)gridDim.x * blockDim.x) {
        seen[v] = 0;
        F0[v] = 0;
    }
}

// Marks the sources of every slot (one block per slot, lane = source lane).
__global__ void bc_seed(BcArgs a, const int32_t* lane_src) {
    const int slot = blockIdx.x;
    const int lane = threadIdx.x;
    const int32_t s = lane_src[slot * 64 + lane];
    uint64_t* seen = a.seen + (size_t)slot * a.nv;
    uint64_t* F0 = a.F + (size_t)slot * a.cap * a.nv;
    double* sig = a.sig + (size_t)slot * a.nv * 64;
    if (s >= 0) {
        atomicOr((unsigned long long*)&F0[s], 1ull << lane);
        atomicOr((unsigned long long*)&seen[s], 1ull << lane);
        sig[(size_t)s * 64 + lane] = 1.0;
    }
    if (lane == 0) {
        BcCtrl* c = &a.ctrl[slot];
        c->level = 0;
        c->done = a.active[slot] == 0 ? 1 : 0;
        c->maxlevel = 0;
        c->found = 0;
        c->blocks_done = 0;
        c->overflow = 0;
    }
}

__global__ __launch_bounds__(kBlock) void bc_forward(BcArgs a) {
    const int slot = blockIdx.y;
    BcCtrl* c = &a.ctrl[slot];
    __shared__ int32_t s_L, s_done, s_found;
    __shared__ uint32_t s_ticket;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (tid == 0) {
        s_done = c->done;
        s_L = c->level;
        s_found = 0;
        if (!s_done && s_L + 1 >= a.cap) {
            c->overflow = 1;
            s_done = 1;
        }
    }
    __syncthreads();
    if (s_done) return;
    const int32_t L = s_L;
    const int32_t nv = a.nv;
    const uint64_t act = a.active[slot];
    const uint64_t* FL = a.F + ((size_t)slot * a.cap + L) * nv;
    uint64_t* Fn = a.F + ((size_t)slot * a.cap + L + 1) * nv;
    uint64_t* seen = a.seen + (size_t)slot * nv;
    double* sig = a.sig + (size_t)slot * nv * 64;
    bool any = false;
    for (int32_t ch = blockIdx.x * kWaves + wid; ch < a.nchunks; ch += gridDim.x * kWaves) {
        const int32_t vj = ch * 64 + lane;
        uint64_t sj = 0;
        int32_t bj = 0, ej = 0;
        bool cand = false;
        if (vj < nv) {
            sj = seen[vj];
            if (~sj & act) {
                bj = a.in_rp[vj];
                ej = a.in_rp[vj + 1];
                cand = ej > bj;
            }
        }
        uint64_t newF = 0;
        uint64_t cm = __ballot(cand);
        while (cm) {
            const int j = __ffsll((unsigned long long)cm) - 1;
            cm &= cm - 1;
            const int32_t v = ch * 64 + j;
            const uint64_t unseen = ~__shfl(sj, j) & act;
            const int32_t b = __shfl(bj, j), e = __shfl(ej, j);
            double acc = 0.0;
            uint64_t got = 0;
            for (int32_t off = b; off < e; off += 64) {
                const int32_t k = off + lane;
                uint64_t mk = 0;
                int32_t uk = 0;
                if (k < e) {
                    uk = a.in_col[k];
                    mk = FL[uk] & unseen;
                }
                uint64_t hits = __ballot(mk != 0);
                while (hits) {
                    const int t0 = __ffsll((unsigned long long)hits) - 1;
                    hits &= hits - 1;
                    const int32_t u0 = __shfl(uk, t0);
                    const uint64_t m0 = __shfl(mk, t0);
                    got |= m0;
                    if (hits) {
                        const int t1 = __ffsll((unsigned long long)hits) - 1;
                        hits &= hits - 1;
                        const int32_t u1 = __shfl(uk, t1);
                        const uint64_t m1 = __shfl(mk, t1);
                        got |= m1;
                        const double x0 = ((m0 >> lane) & 1) ? sig[(size_t)u0 * 64 + lane] : 0.0;
                        const double x1 = ((m1 >> lane) & 1) ? sig[(size_t)u1 * 64 + lane] : 0.0;
                        acc += x0;
                        acc += x1;
                    } else {
                        if ((m0 >> lane) & 1) acc += sig[(size_t)u0 * 64 + lane];
                    }
                }
            }
            if ((got >> lane) & 1) sig[(size_t)v * 64 + lane] = acc;
            if (lane == j) newF = got;
        }
        if (vj < nv) {
            Fn[vj] = newF;
            if (newF) seen[vj] = sj | newF;
        }
        any |= newF != 0;
    }
    if (__ballot(any) && lane == 0) s_found = 1;
    __syncthreads();
    if (tid == 0) {
        if (s_found) atomicOr(&c->found, 1);
        __threadfence();
        s_ticket = atomicAdd(&c->blocks_done, 1u);
    }
    __syncthreads();
    if (tid == 0 && s_ticket == gridDim.x - 1) {
        __threadfence();
        const int32_t f = __hip_atomic_load(&c->found, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        c->blocks_done = 0;
        c->found = 0;
        if (f) {
            c->level = L + 1;
        } else {
            c->done = 1;
            c->maxlevel = L;
        }
        __threadfence();
    }
}

__global__ __launch_bounds__(kBlock) void bc_backward(BcArgs a, int32_t k) {
    const int slot = blockIdx.y;
    const int32_t L = a.ctrl[slot].maxlevel - k;
    if (L < 1) return;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int32_t nv = a.nv;
    const uint64_t* FL = a.F + ((size_t)slot * a.cap + L) * nv;
    const uint64_t* Fn = a.F + ((size_t)slot * a.cap + L + 1) * nv;
    double* sig = a.sig + (size_t)slot * nv * 64;
    double* acc_bc = a.bc_acc + (size_t)slot * nv;
    for (int32_t ch = blockIdx.x * kWaves + wid; ch < a.nchunks; ch += gridDim.x * kWaves) {
        const int32_t vj = ch * 64 + lane;
        uint64_t Aj = 0;
        int32_t bj = 0, ej = 0;
        if (vj < nv) {
            Aj = FL[vj];
            if (Aj) { bj = a.out_rp[vj]; ej = a.out_rp[vj + 1]; }
        }
        double bcj = 0.0;
        uint64_t am = __ballot(Aj != 0);
        while (am) {
            const int j = __ffsll((unsigned long long)am) - 1;
            am &= am - 1;
            const int32_t v = ch * 64 + j;
            const uint64_t A = __shfl(Aj, j);
            const int32_t b = __shfl(bj, j), e = __shfl(ej, j);
            double acc = 0.0;
            for (int32_t off = b; off < e; off += 64) {
                const int32_t kk = off + lane;
                uint64_t mk = 0;
                int32_t wk = 0;
                if (kk < e) {
                    wk = a.out_col[kk];
                    mk = Fn[wk] & A;
                }
                uint64_t hits = __ballot(mk != 0);
                while (hits) {
                    const int t0 = __ffsll((unsigned long long)hits) - 1;
                    hits &= hits - 1;
                    const int32_t w0 = __shfl(wk, t0);
                    const uint64_t m0 = __shfl(mk, t0);
                    if (hits) {
                        const int t1 = __ffsll((unsigned long long)hits) - 1;
                        hits &= hits - 1;
                        const int32_t w1 = __shfl(wk, t1);
                        const uint64_t m1 = __shfl(mk, t1);
                        const double x0 = ((m0 >> lane) & 1) ? sig[(size_t)w0 * 64 + lane] : 0.0;
                        const double x1 = ((m1 >> lane) & 1) ? sig[(size_t)w1 * 64 + lane] : 0.0;
                        acc += x0;
                        acc += x1;
                    } else {
                        if ((m0 >> lane) & 1) acc += sig[(size_t)w0 * 64 + lane];
                    }
                }
            }
            double contrib = 0.0;
            if ((A >> lane) & 1) {
                const double s = sig[(size_t)v * 64 + lane];
                const double d = s * acc;
                sig[(size_t)v * 64 + lane] = (1.0 + d) / s;
                contrib = d;
            }
            for (int o = 32; o >= 1; o >>= 1) contrib += __shfl_xor(contrib, o);
            if (lane == j) bcj = contrib;
        }
        if (Aj) acc_bc[vj] += bcj;
    }
}

// bc[v] = (sum over slots of bc_acc, in slot order, + endpoints term) * scale
__global__ void bc_finish(const double* acc, int32_t nslots, int32_t nv, const double* endp, int do_scale,
                          double scale, double* out) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int k = 0; k < nslots; k++) s += acc[(size_t)k * nv + v];
        if (endp) s += endp[v];
        if (do_scale) s *= scale;
        out[v] = s;
    }
}

}  // namespace

void bc_free(BcWork* w) {
    if (!w) return;
    dev_free(w->seen); dev_free(w->F); dev_free(w->sig); dev_free(w->bc_acc); dev_free(w->bc_out);
    dev_free(w->endp); dev_free(w->active); dev_free(w->lane_src); dev_free(w->ctrl);
    if (w->h_ctrl) (void)hipHostFree(w->h_ctrl);
    for (auto& e : w->ev) if (e) (void)hipEventDestroy(e);
    delete w;
}

static size_t free_device_bytes() {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) { (void)hipGetLastError(); return 0; }
    return fr;
}

// (Re)allocates the per-slot state for `nslots` slots and `cap` levels.
static void bc_alloc(lgmi_graph* g, int32_t nslots, int32_t cap) {
    BcWork* w = g->bc;
    if (w && w->nslots >= nslots && w->cap_levels >= cap && w->nv == g->nv) return;
    bc_free(w);
    g->bc = nullptr;
    w = new BcWork();
    try {
        w->nv = g->nv;
        w->nslots = nslots;
        w->cap_levels = cap;
        const size_t nv = (size_t)std::max(g->nv, 1);
        w->seen = dev_alloc_n<uint64_t>(nv * nslots);
        w->F = dev_alloc_n<uint64_t>(nv * nslots * (size_t)cap);
        w->sig = dev_alloc_n<double>(nv * nslots * 64);
        w->bc_acc = dev_alloc_n<double>(nv * nslots);
        w->bc_out = dev_alloc_n<double>(nv);
        w->endp = dev_alloc_n<double>(nv);
        w->active = dev_alloc_n<uint64_t>(nslots);
        w->lane_src = dev_alloc_n<int32_t>((size_t)nslots * 64);
        w->ctrl = dev_alloc_n<BcCtrl>(nslots);
        LGMI_HIP(hipHostMalloc((void**)&w->h_ctrl, sizeof(BcCtrl) * nslots, hipHostMallocDefault));
        for (auto& e : w->ev) LGMI_HIP(hipEventCreate(&e));
    } catch (...) {
        bc_free(w);
        throw;
    }
    g->bc = w;
}

// Grows the level capacity keeping the first `keep` levels of every slot.
static void bc_grow_levels(lgmi_graph* g, int32_t new_cap, int32_t keep) {
    BcWork* w = g->bc;
    const size_t nv = (size_t)std::max(g->nv, 1);
    uint64_t* F2 = dev_alloc_n<uint64_t>(nv * w->nslots * (size_t)new_cap);
    for (int s = 0; s < w->nslots; s++)
        LGMI_HIP(hipMemcpyAsync(F2 + (size_t)s * new_cap * nv, w->F + (size_t)s * w->cap_levels * nv,
                                sizeof(uint64_t) * nv * (size_t)keep, hipMemcpyDeviceToDevice, g->stream));
    LGMI_HIP(hipStreamSynchronize(g->stream));
    dev_free(w->F);
    w->F = F2;
    w->cap_levels = new_cap;
}

void bc_run(lgmi_graph* g, const int32_t* srcs, int32_t nsrc, int64_t k_scale, int skip_degree0, int normalize,
            int endpoints, double* bc, lgmi_bc_stats* stats) {
    const int32_t nv = g->nv;
    hipStream_t st = g->stream;
    // degree(g, s) > 0 filter (betweenness.jl:58) on host copies of the row pointers
    std::vector<int32_t> rp_out((size_t)nv + 1), rp_in;
    LGMI_HIP(hipMemcpy(rp_out.data(), g->out.row_ptr, sizeof(int32_t) * ((size_t)nv + 1), hipMemcpyDeviceToHost));
    if (g->directed) {
        rp_in.resize((size_t)nv + 1);
        LGMI_HIP(hipMemcpy(rp_in.data(), g->in.row_ptr, sizeof(int32_t) * ((size_t)nv + 1), hipMemcpyDeviceToHost));
    }
    auto degree = [&](int32_t v) -> int64_t {
        int64_t d = rp_out[v + 1] - rp_out[v];
        if (g->directed) d += rp_in[v + 1] - rp_in[v];
        return d;
    };
    std::vector<int32_t> act;
    act.reserve((size_t)nsrc);
    for (int32_t i = 0; i < nsrc; i++)
        if (!skip_degree0 || degree(srcs[i]) > 0) act.push_back(srcs[i]);
    const int64_t nact = (int64_t)act.size();
    const int32_t ngroups = (int32_t)((nact + 63) / 64);

    // slots: as many groups at once as fit a memory budget
    const int32_t cap0 = 32;
    const size_t per_slot = (size_t)std::max(nv, 1) * (8 + 8 * (size_t)cap0 + 512 + 8);
    size_t budget = free_device_bytes() / 3;
    if (budget == 0) budget = (size_t)4 << 30;
    int32_t nslots = (int32_t)std::max<size_t>(1, std::min<size_t>((size_t)std::max(ngroups, 1), budget / per_slot));
    nslots = std::min(nslots, 1024);
    bc_alloc(g, nslots, cap0);
    BcWork* w = g->bc;
    nslots = std::min(w->nslots, std::max(ngroups, 1));

    BcArgs a{};
    a.in_rp = g->in.row_ptr; a.in_col = g->in.col;
    a.out_rp = g->out.row_ptr; a.out_col = g->out.col;
    a.nv = nv;
    a.nchunks = (nv + 63) / 64;
    a.seen = w->seen; a.F = w->F; a.sig = w->sig; a.bc_acc = w->bc_acc;
    a.active = w->active; a.ctrl = w->ctrl; a.cap = w->cap_levels;

    const int cu = std::max(1, g->cu_count);
    const int target_blocks = cu * 8;
    const int bx = std::max(1, std::min((a.nchunks + kWaves - 1) / kWaves, std::max(1, target_blocks / nslots)));
    const int ibx = std::max(1, std::min((nv + 255) / 256, std::max(1, target_blocks / nslots)));

    LGMI_HIP(hipEventRecord(w->ev[0], st));
    bc_clear_acc<<<std::max(1, std::min(target_blocks, (int)(((size_t)nv * nslots + 255) / 256))), 256, 0, st>>>(
        w->bc_acc, (int64_t)nv * nslots);
    float fwd_ms = 0.f, bwd_ms = 0.f;
    int32_t max_levels = 0;
    std::vector<int32_t> lane_src((size_t)nslots * 64);
    std::vector<uint64_t> active((size_t)nslots);
    for (int32_t g0 = 0; g0 < ngroups; g0 += nslots) {
        const int32_t ns = std::min(nslots, ngroups - g0);
        for (int32_t s = 0; s < nslots; s++) {
            active[s] = 0;
            for (int l = 0; l < 64; l++) {
                const int64_t idx = (int64_t)(g0 + s) * 64 + l;
                const bool on = s < ns && idx < nact;
                lane_src[(size_t)s * 64 + l] = on ? act[(size_t)idx] : -1;
                if (on) active[s] |= 1ull << l;
            }
        }
        LGMI_HIP(hipMemcpyAsync(w->lane_src, lane_src.data(), sizeof(int32_t) * lane_src.size(), hipMemcpyHostToDevice, st));
        LGMI_HIP(hipMemcpyAsync(w->active, active.data(), sizeof(uint64_t) * active.size(), hipMemcpyHostToDevice, st));
        a.F = w->F; a.cap = w->cap_levels;
        bc_init<<<dim3(ibx, ns), 256, 0, st>>>(a);
        bc_seed<<<ns, 64, 0, st>>>(a, w->lane_src);
        LGMI_HIP(hipGetLastError());
        LGMI_HIP(hipEventRecord(w->ev[1], st));
        // forward: level steps in batches until every slot is done
        int32_t lcur = 0;
        int batch = 8;
        for (;;) {
            if (lcur + batch + 2 >= w->cap_levels) {
                int32_t nc = w->cap_levels;
                while (lcur + batch + 2 >= nc) nc *= 2;
                // a finished slot also needs its empty F_{maxlevel+1}
                bc_grow_levels(g, nc, std::min(lcur + 2, w->cap_levels));
                a.F = w->F; a.cap = w->cap_levels;
            }
            for (int i = 0; i < batch; i++) bc_forward<<<dim3(bx, ns), kBlock, 0, st>>>(a);
            LGMI_HIP(hipGetLastError());
            LGMI_HIP(hipMemcpyAsync(w->h_ctrl, w->ctrl, sizeof(BcCtrl) * ns, hipMemcpyDeviceToHost, st));
            LGMI_HIP(hipStreamSynchronize(st));
            bool all = true;
            lcur = 0;
            for (int s = 0; s < ns; s++) {
                if (w->h_ctrl[s].overflow) throw Error(LGMI_E_HIP, "betweenness level capacity exceeded");
                all = all && w->h_ctrl[s].done;
                lcur = std::max(lcur, w->h_ctrl[s].done ? w->h_ctrl[s].maxlevel : w->h_ctrl[s].level);
            }
            if (all) break;
            if (lcur > nv + 2) throw Error(LGMI_E_HIP, "betweenness forward pass did not terminate");
            batch = std::min(batch * 2, 256);
        }
        int32_t lmax = 0;
        for (int s = 0; s < ns; s++) lmax = std::max(lmax, w->h_ctrl[s].maxlevel);
        max_levels = std::max(max_levels, lmax);
        LGMI_HIP(hipEventRecord(w->ev[2], st));
        for (int32_t k = 0; k < lmax; k++) bc_backward<<<dim3(bx, ns), kBlock, 0, st>>>(a, k);
        LGMI_HIP(hipGetLastError());
        LGMI_HIP(hipEventRecord(w->ev[3], st));
        LGMI_HIP(hipEventSynchronize(w->ev[3]));
        float t1 = 0.f, t2 = 0.f;
        LGMI_HIP(hipEventElapsedTime(&t1, w->ev[1], w->ev[2]));
        LGMI_HIP(hipEventElapsedTime(&t2, w->ev[2], w->ev[3]));
        fwd_ms += t1;
        bwd_ms += t2;
    }
    // endpoints: every processed source s adds length(S)-1 = nv-1 to bc[s]
    // and 1 to every other vertex (betweenness.jl:122,130)
    const double* endp = nullptr;
    if (endpoints && nact > 0) {
        std::vector<double> e((size_t)nv, (double)nact);
        std::vector<int64_t> cnt((size_t)nv, 0);
        for (int32_t s : act) cnt[(size_t)s]++;
        for (int32_t v = 0; v < nv; v++)
            if (cnt[v]) e[v] = (double)(nact - cnt[v]) + (double)(nv - 1) * (double)cnt[v];
        LGMI_HIP(hipMemcpyAsync(w->endp, e.data(), sizeof(double) * (size_t)nv, hipMemcpyHostToDevice, st));
        LGMI_HIP(hipStreamSynchronize(st));
        endp = w->endp;
    }
    // _rescale! (betweenness.jl:136-160), same operation order
    int do_scale = 0;
    double scale = 0.0;
    if (normalize) {
        if (nv > 2) { do_scale = 1; scale = 1.0 / (double)(((int64_t)nv - 1) * ((int64_t)nv - 2)); }
    } else if (!g->directed) {
        do_scale = 1;
        scale = 1.0 / 2.0;
    }
    if (do_scale && k_scale > 0) scale = scale * (double)nv / (double)k_scale;
    if (nv > 0) {
        bc_finish<<<std::max(1, std::min(target_blocks, (nv + 255) / 256)), 256, 0, st>>>(
            w->bc_acc, nslots, nv, endp, do_scale, scale, w->bc_out);
        LGMI_HIP(hipGetLastError());
        LGMI_HIP(hipMemcpyAsync(bc, w->bc_out, sizeof(double) * (size_t)nv, hipMemcpyDeviceToHost, st));
    }
    LGMI_HIP(hipEventRecord(w->ev[3], st));
    LGMI_HIP(hipEventSynchronize(w->ev[3]));
    if (stats) {
        float tot = 0.f;
        LGMI_HIP(hipEventElapsedTime(&tot, w->ev[0], w->ev[3]));
        stats->groups = ngroups;
        stats->max_levels = max_levels;
        stats->device_ms = tot;
        stats->forward_ms = fwd_ms;
        stats->backward_ms = bwd_ms;
    }
}

}  // namespace lgmi
