// bfs.hip -- direction-optimising single/multi-source BFS for gfx950.
//
// Replaces the level loop of gdistances! (src/traversals/bfs.jl:95-124),
// _bfs_parents (bfs.jl:38-66) and the threaded Parallel.gdistances! /
// Parallel.bfs_tree! (src/Parallel/traversals/gdistances.jl:36-116,
// src/Parallel/traversals/bfs.jl:49-92).
//
// One kernel launch advances the search by one level L -> L+1, either
//  * top-down: expand the frontier queue Q_L along traversal edges; the
//    vertices of a work-group are binned by degree into workgroup-wide,
//    wavefront-wide and per-lane (scan-based gather) processing; new
//    vertices are claimed with an atomicOr on the next-frontier bitmap and
//    appended to Q_{L+1} with one atomic per wavefront; parents take the
//    minimum over all frontier neighbours with atomicMin; or
//  * bottom-up: every unvisited vertex scans its reverse adjacency in
//    ascending order and stops at the first neighbour in the frontier
//    bitmap F_L, which is the minimum-id parent.
// Both directions give identical distances and parents: the reference's
// sorted frontier makes parent[v] the smallest u with dist[u] = dist[v]-1
// and an edge u -> v, which is what either direction computes.
//
// The last work-group to finish a step (ticket counter) folds the step's
// counters into the control block and picks the next direction: Beamer's
// alpha/beta rule, a learned MLP, or a forced direction.  The host only
// enqueues step launches; it synchronises once per batch of steps.
#include "lgmi_internal.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace lgmi {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;
constexpr int32_t kWgThreshold = 4096;   // top-down: whole work-group per vertex
constexpr int32_t kWaveThreshold = 64;   // top-down: one wavefront per vertex
constexpr int32_t kBuCoop = 48;          // bottom-up: wavefront-cooperative scan

struct StepArgs {
    const int32_t* td_rp;     // traversal adjacency (top-down expands it)
    const int32_t* td_col;
    const int32_t* bu_rp;     // reverse adjacency (bottom-up scans it)
    const int32_t* bu_col;
    const uint32_t* bu_nz;    // bit v: row v of the reverse adjacency is non-empty
    int32_t nv;
    int32_t nwords;           // 2 * ceil(nv / 64)
    int32_t* dist;
    uint32_t* parent;         // 0xffffffff = unset (nullptr: no parents)
    int32_t* q0;
    int32_t* q1;
    uint32_t* bm0;
    uint32_t* bm1;
    uint32_t* bm2;
    uint32_t* vis;            // union of F_0 .. F_{L-1} (F_L folded in during step L)
    BfsCtrl* ctrl;
    LevelRec* log;
    int32_t max_log;
    const PolicyWeights* policy;
};

__device__ inline uint32_t* bm_at(const StepArgs& a, int k) {
    return k == 0 ? a.bm0 : (k == 1 ? a.bm1 : a.bm2);
}

__device__ inline int hist_bin(int32_t deg) {
    // bins of powers of four: 0, 1-3, 4-15, 16-63, 64-255, 256-1023, >=1024
    if (deg <= 0) return 0;
    int b = (31 - __clz(deg)) / 2 + 1;
    return b > 6 ? 6 : b;
}

// Wavefront-aggregated append: one atomicAdd per wavefront and a prefix
// count of the active lanes.  Must be reached by all active lanes.
__device__ inline void wave_append(bool pred, int32_t v, int32_t* q, int32_t* counter) {
    uint64_t m = __ballot(pred);
    if (m == 0) return;
    const int lane = __lane_id();
    const int leader = __ffsll((unsigned long long)m) - 1;
    int32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (int32_t)__popcll(m));
    base = __shfl(base, leader);
    if (pred) q[base + __popcll(m & ((1ull << lane) - 1ull))] = v;
}

__device__ inline float flog2p(double x) { return (float)log2(1.0 + (x > 0 ? x : 0.0)); }

// Features of the state at the start of a step (see LGMI_POLICY_FEATURES).
__device__ void compute_features(const BfsCtrl* c, float* f) {
    const double nf = c->n_front, mf = (double)c->m_front, mu = (double)c->m_unvisited;
    const double nu = (double)c->n_unvisited;
    f[0] = flog2p(nf);
    f[1] = flog2p(mf);
    f[2] = flog2p(mu);
    f[3] = flog2p(nu);
    f[4] = flog2p((double)c->level);
    f[5] = (float)(nf / (double)(c->nv > 0 ? c->nv : 1));
    f[6] = flog2p(mf) - flog2p(mu);
    f[7] = flog2p(nf) - flog2p((double)c->n_prev);
    f[8] = (float)c->prev_dir;
    uint32_t tot = 0;
    for (int b = 0; b < 7; b++) tot += c->front_hist[b];
    for (int b = 0; b < 7; b++) f[9 + b] = tot ? (float)c->front_hist[b] / (float)tot : 0.f;
}

// Direction for the next step.  Runs in the last block; `h` is LDS scratch.
__device__ int choose_direction(const StepArgs& a, BfsCtrl* c, const float* feat, float* h) {
    if (c->forced >= 0) return c->forced;
    if (c->policy_mode == LGMI_POLICY_TOPDOWN) return 0;
    if (c->policy_mode == LGMI_POLICY_LEARNED && a.policy != nullptr) {
        const PolicyWeights* w = a.policy;
        const int t = threadIdx.x;
        if (t < kHidden) {
            float acc = w->b1[t];
            for (int i = 0; i < kFeat; i++) acc += w->w1[t * kFeat + i] * ((feat[i] - w->mu[i]) / w->sd[i]);
            h[t] = acc > 0.f ? acc : 0.f;
        }
        __syncthreads();
        __shared__ int s_dec;
        if (t == 0) {
            float z0 = w->b2[0], z1 = w->b2[1];
            for (int j = 0; j < kHidden; j++) { z0 += w->w2[j] * h[j]; z1 += w->w2[kHidden + j] * h[j]; }
            s_dec = z1 > z0 ? 1 : 0;
        }
        __syncthreads();
        return s_dec;
    }
    // Beamer alpha/beta (GAP form): leave top-down when the frontier's edges
    // exceed the unvisited edges / alpha; leave bottom-up when the frontier
    // shrinks and holds fewer than nv / beta vertices.
    if (c->prev_dir == 0) {
        return ((double)c->m_front > (double)c->m_unvisited / (double)c->alpha) ? 1 : 0;
    }
    const bool shrinking = c->n_front < c->n_prev;
    const bool small = (double)c->n_front <= (double)c->nv / (double)c->beta;
    return (shrinking && small) ? 0 : 1;
}

template <bool PARENTS>
__device__ inline bool td_visit(const StepArgs& a, int32_t u, int32_t v, int32_t L,
                                const uint32_t* FL, uint32_t* Fn) {
    const int w = v >> 5;
    const uint32_t bit = 1u << (v & 31);
    if ((a.vis[w] | FL[w]) & bit) return false;       // distance <= L already
    if (PARENTS) atomicMin(&a.parent[v], (uint32_t)u);
    if (Fn[w] & bit) return false;
    const uint32_t old = atomicOr(&Fn[w], bit);
    if (old & bit) return false;
    a.dist[v] = L + 1;
    return true;
}

template <bool PARENTS>
__global__ __launch_bounds__(kBlock) void bfs_step(StepArgs a) {
    __shared__ int32_t s_level, s_dir, s_done, s_nfront;
    __shared__ uint32_t s_ticket;
    __shared__ unsigned long long s_out, s_in, s_scan, s_cand;
    __shared__ uint32_t s_hist[kHistBins];
    __shared__ int32_t s_big_n;
    __shared__ int32_t s_big_u[kBlock], s_big_beg[kBlock], s_big_deg[kBlock];
    __shared__ int32_t s_wu[kWavesPerBlock][64], s_wbeg[kWavesPerBlock][64], s_wscan[kWavesPerBlock][64];
    __shared__ float s_h[kHidden];
    __shared__ float s_feat[kFeat];

    BfsCtrl* c = a.ctrl;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wid = tid >> 6;
    if (tid == 0) {
        s_done = c->done;
        s_level = c->level;
        s_dir = c->dir;
        s_nfront = c->n_front;
        s_out = 0;
        s_in = 0;
        s_scan = 0;
        s_cand = 0;
        if (!s_done) atomicMin((unsigned long long*)&c->t_step_start, (unsigned long long)wall_clock64());
    }
    if (tid < kHistBins) s_hist[tid] = 0;
    __syncthreads();
    if (s_done) return;

    const int32_t L = s_level;
    const uint32_t* FL = bm_at(a, L % 3);
    uint32_t* Fn = bm_at(a, (L + 1) % 3);
    uint32_t* Fc = bm_at(a, (L + 2) % 3);
    const int32_t* q_in = (L & 1) ? a.q1 : a.q0;
    int32_t* q_out = (L & 1) ? a.q0 : a.q1;
    unsigned long long my_out = 0, my_in = 0, my_scan = 0, my_cand = 0;
    const bool same_adj = (a.td_rp == a.bu_rp);

    auto account = [&](bool claimed, int32_t v) {
        if (claimed) {
            const int32_t dout = a.td_rp[v + 1] - a.td_rp[v];
            const int32_t din = same_adj ? dout : (a.bu_rp[v + 1] - a.bu_rp[v]);
            my_out += (unsigned long long)dout;
            my_in += (unsigned long long)din;
            atomicAdd(&s_hist[hist_bin(dout)], 1u);
        }
        wave_append(claimed, v, q_out, &c->next_count);
    };

    if (s_dir == 0) {
        // ---------------- top-down ----------------
        // fold F_L into vis, clear the bitmap of step L+2
        for (int w = blockIdx.x * kBlock + tid; w < a.nwords; w += gridDim.x * kBlock) {
            a.vis[w] |= FL[w];
            Fc[w] = 0;
        }
        const int32_t n_in = s_nfront;
        for (int64_t base = (int64_t)blockIdx.x * kBlock; base < n_in; base += (int64_t)gridDim.x * kBlock) {
            const int64_t idx = base + tid;
            int32_t u = -1, beg = 0, deg = 0;
            if (idx < n_in) {
                u = q_in[idx];
                beg = a.td_rp[u];
                deg = a.td_rp[u + 1] - beg;
            }
            // workgroup bin
            if (tid == 0) s_big_n = 0;
            __syncthreads();
            if (deg >= kWgThreshold) {
                const int k = atomicAdd(&s_big_n, 1);
                s_big_u[k] = u; s_big_beg[k] = beg; s_big_deg[k] = deg;
                deg = 0;
            }
            __syncthreads();
            const int nbig = s_big_n;
            for (int k = 0; k < nbig; k++) {
                const int32_t bu = s_big_u[k], bb = s_big_beg[k], bd = s_big_deg[k];
                for (int32_t e0 = 0; e0 < bd; e0 += kBlock) {
                    const int32_t e = e0 + tid;
                    bool claimed = false;
                    int32_t v = 0;
                    if (e < bd) { v = a.td_col[bb + e]; claimed = td_visit<PARENTS>(a, bu, v, L, FL, Fn); }
                    account(claimed, v);
                }
            }
            // wavefront bin
            uint64_t m = __ballot(deg >= kWaveThreshold);
            while (m) {
                const int j = __ffsll((unsigned long long)m) - 1;
                m &= m - 1;
                const int32_t wu = __shfl(u, j), wb = __shfl(beg, j), wd = __shfl(deg, j);
                for (int32_t e0 = 0; e0 < wd; e0 += 64) {
                    const int32_t e = e0 + lane;
                    bool claimed = false;
                    int32_t v = 0;
                    if (e < wd) { v = a.td_col[wb + e]; claimed = td_visit<PARENTS>(a, wu, v, L, FL, Fn); }
                    account(claimed, v);
                }
            }
            if (deg >= kWaveThreshold) deg = 0;
            // per-lane bin: scan-based gather over the wavefront's small rows
            int32_t incl = deg;
            for (int d = 1; d < 64; d <<= 1) {
                const int32_t y = __shfl_up(incl, d);
                if (lane >= d) incl += y;
            }
            const int32_t total = __shfl(incl, 63);
            s_wu[wid][lane] = u;
            s_wbeg[wid][lane] = beg;
            s_wscan[wid][lane] = incl - deg;
            __syncthreads();
            for (int32_t off = 0; off < total; off += 64) {
                const int32_t e = off + lane;
                bool claimed = false;
                int32_t v = 0;
                if (e < total) {
                    int lo = 0, hi = 63;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (s_wscan[wid][mid] <= e) lo = mid; else hi = mid - 1;
                    }
                    const int32_t su = s_wu[wid][lo];
                    v = a.td_col[s_wbeg[wid][lo] + (e - s_wscan[wid][lo])];
                    claimed = td_visit<PARENTS>(a, su, v, L, FL, Fn);
                }
                account(claimed, v);
            }
            __syncthreads();
        }
    } else {
        // ---------------- bottom-up ----------------
        const int32_t nchunks = a.nwords >> 1;
        for (int32_t ch = blockIdx.x * kWavesPerBlock + wid; ch < nchunks; ch += gridDim.x * kWavesPerBlock) {
            const int32_t v = ch * 64 + lane;
            const int w = v >> 5;
            const uint32_t bit = 1u << (lane & 31);
            const uint32_t fw = FL[w];
            const uint32_t visw = a.vis[w] | fw;
            const bool cand = (v < a.nv) && !(visw & bit) && (a.bu_nz[w] & bit);
            int32_t beg = 0, end = 0;
            if (cand) { beg = a.bu_rp[v]; end = a.bu_rp[v + 1]; }
            int32_t par = -1;
            const bool big = cand && (end - beg) > kBuCoop;
            if (cand) my_cand++;
            if (cand && !big) {
                for (int32_t e = beg; e < end; e += 4) {
                    my_scan += (unsigned long long)min(4, end - e);
                    const int32_t u0 = a.bu_col[e];
                    const int32_t u1 = (e + 1 < end) ? a.bu_col[e + 1] : u0;
                    const int32_t u2 = (e + 2 < end) ? a.bu_col[e + 2] : u0;
                    const int32_t u3 = (e + 3 < end) ? a.bu_col[e + 3] : u0;
                    const uint32_t b0 = FL[u0 >> 5], b1 = FL[u1 >> 5], b2 = FL[u2 >> 5], b3 = FL[u3 >> 5];
                    if ((b0 >> (u0 & 31)) & 1u) { par = u0; break; }
                    if ((b1 >> (u1 & 31)) & 1u) { par = u1; break; }
                    if ((b2 >> (u2 & 31)) & 1u) { par = u2; break; }
                    if ((b3 >> (u3 & 31)) & 1u) { par = u3; break; }
                }
            }
            uint64_t mb = __ballot(big);
            while (mb) {
                const int j = __ffsll((unsigned long long)mb) - 1;
                mb &= mb - 1;
                const int32_t jb = __shfl(beg, j), je = __shfl(end, j);
                int32_t found_u = -1;
                for (int32_t off = jb; off < je; off += 64) {
                    const int32_t e = off + lane;
                    bool hit = false;
                    int32_t uu = 0;
                    if (e < je) { uu = a.bu_col[e]; hit = (FL[uu >> 5] >> (uu & 31)) & 1u; }
                    const uint64_t hm = __ballot(hit);
                    if (lane == j) my_scan += (unsigned long long)min(64, je - off);
                    if (hm) { found_u = __shfl(uu, __ffsll((unsigned long long)hm) - 1); break; }
                }
                if (lane == j) par = found_u;
            }
            const bool found = par >= 0;
            if (found) {
                a.dist[v] = L + 1;
                if (PARENTS) a.parent[v] = (uint32_t)par;
            }
            const uint64_t fm = __ballot(found);
            if ((lane & 31) == 0) {
                Fn[w] = (lane == 0) ? (uint32_t)fm : (uint32_t)(fm >> 32);
                a.vis[w] = visw;
                Fc[w] = 0;
            }
            account(found, v);
        }
    }

    // ---------------- block reduction and step finalisation ----------------
    for (int d = 32; d >= 1; d >>= 1) {
        my_out += __shfl_down(my_out, d);
        my_in += __shfl_down(my_in, d);
        my_scan += __shfl_down(my_scan, d);
        my_cand += __shfl_down(my_cand, d);
    }
    if (lane == 0) {
        atomicAdd(&s_out, my_out);
        atomicAdd(&s_in, my_in);
        atomicAdd(&s_scan, my_scan);
        atomicAdd(&s_cand, my_cand);
    }
    __syncthreads();
    if (tid == 0) {
        if (s_out) atomicAdd(&c->next_out, s_out);
        if (s_in) atomicAdd(&c->next_in, s_in);
        if (s_scan) atomicAdd(&c->next_scanned, s_scan);
        if (s_cand) atomicAdd(&c->next_cand, s_cand);
    }
    if (tid < kHistBins && s_hist[tid]) atomicAdd(&c->hist[tid], s_hist[tid]);
    __threadfence();
    __syncthreads();
    if (tid == 0) s_ticket = atomicAdd(&c->blocks_done, 1u);
    __syncthreads();
    if (s_ticket != gridDim.x - 1) return;

    // Last block: every other block has published its counters.
    __threadfence();
    if (tid == 0) {
        const uint64_t now = wall_clock64();
        const int32_t n_next = __hip_atomic_load(&c->next_count, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long nout = __hip_atomic_load(&c->next_out, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long nin = __hip_atomic_load(&c->next_in, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (L < a.max_log) {
            LevelRec* r = &a.log[L];
            r->dir = s_dir;
            r->n_front = c->n_front;
            r->m_front = c->m_front;
            r->m_unvisited = c->m_unvisited;
            r->n_next = n_next;
            r->scanned = s_dir ? (int64_t)__hip_atomic_load(&c->next_scanned, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)
                               : c->m_front;
            r->cand = (int64_t)__hip_atomic_load(&c->next_cand, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            r->t_start = c->t_step_start;
            r->t_end = now;
            r->t_prev_end = c->t_prev_end;
        }
        c->t_prev_end = now;
        c->t_step_start = ~0ull;
        c->n_prev = c->n_front;
        c->n_front = n_next;
        c->m_front = (int64_t)nout;
        c->m_unvisited -= (int64_t)nin;
        c->n_unvisited -= n_next;
        c->prev_dir = s_dir;
        c->level = L + 1;
        for (int b = 0; b < kHistBins; b++) {
            c->front_hist[b] = __hip_atomic_load(&c->hist[b], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            c->hist[b] = 0;
        }
        c->next_count = 0;
        c->next_out = 0;
        c->next_in = 0;
        c->next_scanned = 0;
        c->next_cand = 0;
        c->blocks_done = 0;
        if (n_next == 0) c->done = 1;
        else compute_features(c, s_feat);
        s_done = c->done;
    }
    __syncthreads();
    if (s_done) return;
    const int nd = choose_direction(a, c, s_feat, s_h);
    if (tid == 0) {
        c->dir = nd;
        if (L + 1 < a.max_log) {
            for (int i = 0; i < kFeat; i++) a.log[L + 1].feat[i] = s_feat[i];
        }
        __threadfence();
    }
}

__global__ void bfs_init(int32_t* dist, uint32_t* parent, int32_t nv, uint32_t* bm0, uint32_t* bm1,
                         uint32_t* bm2, uint32_t* vis, int32_t nwords) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride) {
        dist[i] = -1;
        if (parent) parent[i] = 0xffffffffu;
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += stride) {
        bm0[i] = 0; bm1[i] = 0; bm2[i] = 0; vis[i] = 0;
    }
}

// Marks the sources (already in q0 unless single_src >= 0) and fills the
// control block.  One block.
__global__ __launch_bounds__(kBlock) void bfs_seed(StepArgs a, int32_t nsrc, int32_t single_src,
                                                   int64_t nnz_scan, int32_t forced, int32_t mode,
                                                   float alpha, float beta) {
    __shared__ unsigned long long s_out, s_in;
    __shared__ uint32_t s_hist[kHistBins];
    const int tid = threadIdx.x;
    if (tid == 0) { s_out = 0; s_in = 0; }
    if (tid < kHistBins) s_hist[tid] = 0;
    __syncthreads();
    if (single_src >= 0 && tid == 0) a.q0[0] = single_src;
    __syncthreads();
    unsigned long long mo = 0, mi = 0;
    for (int i = tid; i < nsrc; i += kBlock) {
        const int32_t s = a.q0[i];
        a.dist[s] = 0;
        if (a.parent) a.parent[s] = (uint32_t)s;
        atomicOr(&a.bm0[s >> 5], 1u << (s & 31));
        const int32_t dout = a.td_rp[s + 1] - a.td_rp[s];
        mo += (unsigned long long)dout;
        mi += (unsigned long long)(a.bu_rp[s + 1] - a.bu_rp[s]);
        atomicAdd(&s_hist[hist_bin(dout)], 1u);
    }
    atomicAdd(&s_out, mo);
    atomicAdd(&s_in, mi);
    __syncthreads();
    if (tid == 0) {
        BfsCtrl* c = a.ctrl;
        c->level = 0;
        c->dir = forced >= 0 ? forced : 0;
        c->done = nsrc == 0 ? 1 : 0;
        c->n_front = nsrc;
        c->n_prev = 0;
        c->prev_dir = 0;
        c->m_front = (int64_t)s_out;
        c->m_unvisited = nnz_scan - (int64_t)s_in;
        c->n_unvisited = (int64_t)a.nv - nsrc;
        c->next_count = 0;
        c->blocks_done = 0;
        c->next_out = 0;
        c->next_in = 0;
        c->next_scanned = 0;
        c->next_cand = 0;
        for (int b = 0; b < kHistBins; b++) { c->hist[b] = 0; c->front_hist[b] = s_hist[b]; }
        c->t_prev_end = wall_clock64();
        c->t_step_start = ~0ull;
        c->forced = forced;
        c->policy_mode = mode;
        c->alpha = alpha;
        c->beta = beta;
        c->nv = a.nv;
        c->nnz_scan = nnz_scan;
        if (a.max_log > 0) {
            float f[kFeat];
            compute_features(c, f);
            for (int i = 0; i < kFeat; i++) a.log[0].feat[i] = f[i];
        }
    }
}

}  // namespace

BfsWork::~BfsWork() {}

void bfs_free(BfsWork* w) {
    if (!w) return;
    dev_free(w->dist); dev_free(w->parent); dev_free(w->q[0]); dev_free(w->q[1]);
    dev_free(w->bm[0]); dev_free(w->bm[1]); dev_free(w->bm[2]); dev_free(w->vis);
    dev_free(w->ctrl); dev_free(w->log); dev_free(w->policy);
    if (w->h_ctrl) (void)hipHostFree(w->h_ctrl);
    if (w->ev0) (void)hipEventDestroy(w->ev0);
    if (w->ev1) (void)hipEventDestroy(w->ev1);
    for (auto e : w->step_ev) (void)hipEventDestroy(e);
    delete w;
}

void bfs_prepare(lgmi_graph* g) {
    if (g->bfs) return;
    BfsWork* w = new BfsWork();
    try {
        w->nv = g->nv;
        w->nwords = 2 * ((g->nv + 63) / 64);
        w->dist = dev_alloc_n<int32_t>(g->nv);
        w->parent = dev_alloc_n<int32_t>(g->nv);
        // a multi-source call may list up to nsrc >= nv sources; q0 grows on demand
        w->q[0] = dev_alloc_n<int32_t>((size_t)g->nv + 1);
        w->q[1] = dev_alloc_n<int32_t>((size_t)g->nv + 1);
        for (int k = 0; k < 3; k++) w->bm[k] = dev_alloc_n<uint32_t>(w->nwords);
        w->vis = dev_alloc_n<uint32_t>(w->nwords);
        w->ctrl = dev_alloc_n<BfsCtrl>(1);
        w->max_log = 4096;
        w->log = dev_alloc_n<LevelRec>(w->max_log);
        w->policy = dev_alloc_n<PolicyWeights>(1);
        LGMI_HIP(hipHostMalloc((void**)&w->h_ctrl, sizeof(BfsCtrl), hipHostMallocDefault));
        LGMI_HIP(hipEventCreate(&w->ev0));
        LGMI_HIP(hipEventCreate(&w->ev1));
        LGMI_HIP(hipMemsetAsync(w->ctrl, 0, sizeof(BfsCtrl), g->stream));
        LGMI_HIP(hipMemsetAsync(w->log, 0, sizeof(LevelRec) * w->max_log, g->stream));
    } catch (...) {
        bfs_free(w);
        throw;
    }
    g->bfs = w;
}

static StepArgs make_args(lgmi_graph* g, int dir_in, bool parents) {
    BfsWork* w = g->bfs;
    StepArgs a{};
    const DevCSR& td = dir_in ? g->in : g->out;
    const DevCSR& bu = dir_in ? g->out : g->in;
    a.td_rp = td.row_ptr; a.td_col = td.col;
    a.bu_rp = bu.row_ptr; a.bu_col = bu.col; a.bu_nz = bu.nz_bm;
    a.nv = g->nv;
    a.nwords = w->nwords;
    a.dist = w->dist;
    a.parent = parents ? reinterpret_cast<uint32_t*>(w->parent) : nullptr;
    a.q0 = w->q[0]; a.q1 = w->q[1];
    a.bm0 = w->bm[0]; a.bm1 = w->bm[1]; a.bm2 = w->bm[2];
    a.vis = w->vis;
    a.ctrl = w->ctrl;
    a.log = w->log;
    a.max_log = w->max_log;
    a.policy = w->policy;
    return a;
}

static uint64_t g_policy_uploaded_version = 0;

void policy_upload(lgmi_graph* g) {
    std::lock_guard<std::mutex> lk(policy_mutex());
    PolicyHost& p = policy_host();
    if (p.have_mlp) {
        LGMI_HIP(hipMemcpyAsync(g->bfs->policy, &p.w, sizeof(PolicyWeights), hipMemcpyHostToDevice, g->stream));
    }
    g_policy_uploaded_version = p.version;
}

void bfs_run(lgmi_graph* g, const int32_t* h_srcs, int32_t nsrc, int32_t single_src, int dir_in,
             int want_parents, int forced_dir, lgmi_bfs_stats* stats) {
    bfs_prepare(g);
    BfsWork* w = g->bfs;
    if (nsrc > g->nv) throw Error(LGMI_E_ARG, "source list longer than nv (de-duplicate first)");
    int mode;
    float alpha, beta;
    bool have_mlp;
    {
        std::lock_guard<std::mutex> lk(policy_mutex());
        PolicyHost& p = policy_host();
        mode = p.mode; alpha = p.alpha; beta = p.beta; have_mlp = p.have_mlp;
    }
    if (mode == LGMI_POLICY_LEARNED && have_mlp) policy_upload(g);
    if (mode == LGMI_POLICY_LEARNED && !have_mlp) mode = LGMI_POLICY_HEURISTIC;

    StepArgs a = make_args(g, dir_in, want_parents != 0);
    if (!(mode == LGMI_POLICY_LEARNED)) a.policy = nullptr;
    const int64_t nnz_scan = g->nnz;  // the reverse CSR has as many entries as the forward one
    hipStream_t s = g->stream;
    const int grid = std::max(1, g->cu_count) * 8;
    if (single_src < 0 && nsrc > 0)
        LGMI_HIP(hipMemcpyAsync(w->q[0], h_srcs, sizeof(int32_t) * (size_t)nsrc, hipMemcpyHostToDevice, s));
    LGMI_HIP(hipEventRecord(w->ev0, s));
    const int init_grid = std::max(1, std::min<int>(grid, (g->nv + 255) / 256));
    bfs_init<<<init_grid, 256, 0, s>>>(w->dist, a.parent, g->nv, w->bm[0], w->bm[1], w->bm[2], w->vis, w->nwords);
    bfs_seed<<<1, kBlock, 0, s>>>(a, single_src >= 0 ? 1 : nsrc, single_src, nnz_scan, forced_dir, mode, alpha, beta);
    LGMI_HIP(hipGetLastError());
    // Enqueue steps in batches; a step after the last level exits at once.
    int batch = std::max(4, std::min(w->last_levels + 1, 64));
    int launched = 0;
    auto step_event = [&](int i) {
        while ((int)w->step_ev.size() <= i) {
            hipEvent_t e;
            LGMI_HIP(hipEventCreate(&e));
            w->step_ev.push_back(e);
        }
        LGMI_HIP(hipEventRecord(w->step_ev[(size_t)i], s));
    };
    if (w->time_steps) step_event(0);
    for (;;) {
        for (int i = 0; i < batch; i++) {
            if (want_parents) bfs_step<true><<<grid, kBlock, 0, s>>>(a);
            else bfs_step<false><<<grid, kBlock, 0, s>>>(a);
            if (w->time_steps) step_event(launched + i + 1);
        }
        launched += batch;
        LGMI_HIP(hipGetLastError());
        LGMI_HIP(hipMemcpyAsync(w->h_ctrl, w->ctrl, sizeof(BfsCtrl), hipMemcpyDeviceToHost, s));
        LGMI_HIP(hipEventRecord(w->ev1, s));
        LGMI_HIP(hipEventSynchronize(w->ev1));
        if (w->h_ctrl->done) break;
        if (launched > g->nv + 8) throw Error(LGMI_E_HIP, "BFS did not terminate");
        batch = std::min(batch * 2, 1024);
    }
    w->last_levels = w->h_ctrl->level;
    w->last_dir_in = dir_in;
    w->last_parents = want_parents;
    w->last_launches = launched;
    w->last_steps_event_ms = 0.0;
    w->last_launch_ms.assign((size_t)launched, 0.f);
    if (w->time_steps) {
        for (int i = 0; i < launched; i++) {
            float ms = 0.f;
            LGMI_HIP(hipEventElapsedTime(&ms, w->step_ev[(size_t)i], w->step_ev[(size_t)i + 1]));
            w->last_launch_ms[(size_t)i] = ms;
            w->last_steps_event_ms += ms;
        }
    }
    if (stats) {
        float ms = 0.f;
        LGMI_HIP(hipEventElapsedTime(&ms, w->ev0, w->ev1));
        stats->levels = w->h_ctrl->level;
        stats->device_ms = ms;
        stats->td_steps = 0;
        stats->bu_steps = 0;
        stats->reached = 0;
        stats->edges_reached = 0;
        stats->launches = launched;
        stats->pad = 0;
        stats->steps_event_ms = w->last_steps_event_ms;
        const int nrec = std::min(w->h_ctrl->level, w->max_log);
        if (nrec > 0) {
            std::vector<LevelRec> recs((size_t)nrec);
            LGMI_HIP(hipMemcpy(recs.data(), w->log, sizeof(LevelRec) * (size_t)nrec, hipMemcpyDeviceToHost));
            int64_t reached = 0;
            for (const auto& r : recs) {
                if (r.dir) stats->bu_steps++; else stats->td_steps++;
                reached += r.n_front;
            }
            stats->reached = (int32_t)reached;
        }
    }
}

void bfs_fetch(lgmi_graph* g, int32_t* dist, int32_t* parent) {
    BfsWork* w = g->bfs;
    if (!w) throw Error(LGMI_E_ARG, "no BFS has run on this graph");
    if (dist) LGMI_HIP(hipMemcpyAsync(dist, w->dist, sizeof(int32_t) * (size_t)g->nv, hipMemcpyDeviceToHost, g->stream));
    if (parent) {
        if (!w->last_parents) throw Error(LGMI_E_ARG, "the last BFS ran without parents");
        LGMI_HIP(hipMemcpyAsync(parent, w->parent, sizeof(int32_t) * (size_t)g->nv, hipMemcpyDeviceToHost, g->stream));
    }
    LGMI_HIP(hipStreamSynchronize(g->stream));
}

}  // namespace lgmi

// ----------------------------------------------------------------------
// step records export
// ----------------------------------------------------------------------
extern "C" int lgmi_bfs_steps(lgmi_graph_t g, int32_t max_steps, lgmi_bfs_step* out) {
    using namespace lgmi;
    try {
        if (!g || max_steps < 0 || (max_steps > 0 && !out)) throw Error(LGMI_E_ARG, "bad arguments");
        std::lock_guard<std::mutex> lk(g->mu);
        BfsWork* w = g->bfs;
        if (!w) return 0;
        const int n = std::min({max_steps, w->last_levels, w->max_log});
        if (n <= 0) return 0;
        std::vector<LevelRec> recs((size_t)n);
        LGMI_HIP(hipMemcpy(recs.data(), w->log, sizeof(LevelRec) * (size_t)n, hipMemcpyDeviceToHost));
        int dev = 0, khz = 0;
        LGMI_HIP(hipGetDevice(&dev));
        LGMI_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        const double us_per_tick = khz > 0 ? 1000.0 / (double)khz : 0.01;
        for (int i = 0; i < n; i++) {
            const LevelRec& r = recs[(size_t)i];
            lgmi_bfs_step& o = out[i];
            std::memset(&o, 0, sizeof(o));
            for (int k = 0; k < kFeat; k++) o.feat[k] = r.feat[k];
            o.dir = r.dir;
            o.n_front = r.n_front;
            o.m_front = r.m_front;
            o.m_unvisited = r.m_unvisited;
            o.n_next = r.n_next;
            o.scanned = r.scanned;
            o.cand = r.cand;
            o.kernel_us = (float)((double)(r.t_end - r.t_start) * us_per_tick);
            o.step_us = (float)((double)(r.t_end - r.t_prev_end) * us_per_tick);
            o.event_ms = (size_t)i < w->last_launch_ms.size() ? w->last_launch_ms[(size_t)i] : 0.f;
        }
        return n;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return LGMI_E_HIP;
    }
}

extern "C" int lgmi_bfs_time_steps(lgmi_graph_t g, int enable) {
    using namespace lgmi;
    try {
        if (!g) throw Error(LGMI_E_ARG, "bad arguments");
        std::lock_guard<std::mutex> lk(g->mu);
        bfs_prepare(g);
        g->bfs->time_steps = enable != 0;
        return LGMI_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return LGMI_E_HIP;
    }
}
