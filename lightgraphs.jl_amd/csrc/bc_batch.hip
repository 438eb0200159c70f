// bc_batch.hip -- all-source exact betweenness for many independent small
// undirected graphs in one persistent launch (gfx950).
//
// Each graph gets betweenness_centrality(g) with vs = vertices(g),
// normalize / endpoints as given (src/centrality/betweenness.jl:46-160):
// Brandes forward/backward per source with unit distances, which is what
// the reference's Dijkstra pass computes for DefaultDistance
// (src/shortestpaths/dijkstra.jl:59-141).
//
// Work decomposition: a task is (graph, contiguous source range); graphs
// whose n * (nnz + n) exceeds the per-task target are split (by default the
// largest graph into about 128 tasks).  The graphs are dealt to eight task
// queues, one per XCD, largest first to the least-loaded queue; persistent
// 1024-thread work-groups claim tasks from their XCD's queue (decreasing
// work) with an atomic counter, then from the other queues, so that the
// work-groups sharing an L2 work on the same few graphs at a time.  A task
// runs in one of two modes:
//  * block mode: the whole work-group runs one single-source search at a
//    time, with the per-vertex search state in LDS (160 KB holds 20 000
//    vertices), or in global scratch for larger graphs;
//  * wave mode (n <= wave_max_n, an autotuned threshold): each of the 16
//    wavefronts runs its own source with its own LDS slice, synchronising
//    only inside the wavefront.
//
// The search state is one FP64 word per vertex, read by sign:
//    0   unvisited;
//   > 0  on a finished level, holding sigma (the shortest-path count);
//   < 0  forward: discovered on the level being built, holding -sigma;
//        backward: finished, holding -(1 + delta) / sigma.
// In an undirected graph a vertex on level l only has neighbours on levels
// l-1, l and l+1, so the sign alone tells predecessors from successors.
//
// Forward, level by level: every frontier row is expanded (a group of KG
// lanes per row, KG = 4, 8 or 16 as tuned; a whole wavefront for rows
// longer than 8 * KG) and pushes
// its sigma into each neighbour whose state is <= 0 with an LDS FP64
// atomic add; the add that finds 0 appends the neighbour to the queue
// (wavefront-aggregated counter).  Sigma is integer-valued, so the sums are
// exact and independent of the order of the adds while sigma < 2^53.  After
// a barrier the new level's states are negated to positive.
// Backward, deepest level first: every vertex of level l sums the negative
// states of its neighbours (exactly its successors' coefficients) in column
// order per lane and a fixed butterfly across lanes, forms delta = sigma *
// sum, adds it to the task's partial betweenness, and writes
// -(1 + delta) / sigma after a barrier.
// Partial arrays are per task (per wavefront in wave mode, folded in a fixed
// order), and a second kernel adds the tasks of a graph in task order and
// applies _rescale!, so results do not depend on scheduling.
#include "lgmi_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <numeric>
#include <vector>

namespace lgmi {

namespace {

constexpr int kBB = 1024;           // threads per work-group
constexpr int kBW = kBB / 64;       // wavefronts per work-group
// Task queues, one per XCD: work-group b (placed on XCD b mod 8 by the
// dispatcher's round-robin) takes from queue b mod kXq first, so that the
// work-groups sharing an L2 work on the same few graphs and share their
// adjacency there; a work-group whose queue is empty takes from the others.
constexpr int kXq = 8;

struct BatchArgs {
    const int32_t* nv;        // [ngraphs]
    const int64_t* row_off;
    const int64_t* col_off;
    const int32_t* rp_cat;
    const int32_t* col_cat;
    const int32_t* task_graph;
    const int32_t* task_s0;
    const int32_t* task_s1;
    const int64_t* task_part; // offset of the task's partial array
    int32_t ntasks;
    int32_t wave_max_n;
    int32_t lds_max_n;        // the state fits in LDS for n <= this (block mode)
    int32_t max_n;
    double* part;             // task partial betweenness
    int32_t* task_ctr;        // [kXq][32]: one counter per queue, 128 bytes apart
    const int32_t* qoff;      // [kXq + 1] queue q holds tasks qoff[q] .. qoff[q + 1] - 1
    // per-block scratch (global); wave mode slices it per wavefront
    int32_t* queue;           // [blocks][scr]
    double* coef;             // [blocks][scr] new backward states, by queue position
    int32_t* loff;            // [blocks][scr + 2 * kBW] level offsets into the queue
    double* gsig;             // [blocks][max_n] (block mode without LDS room)
    double* wpart;            // [blocks][kBW][wave_max_n]
    int64_t scr;              // >= max(max_n, kBW * wave_max_n)
};

// Team = one wavefront (TEAM = 64) or the whole work-group (TEAM = kBB).
template <int TEAM>
__device__ inline void team_sync() {
    if (TEAM == kBB) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
}

// State load.  In global scratch the states are also changed by atomics,
// which are performed in L2, so the load bypasses the CU cache there.
template <bool LDS>
__device__ inline double state(const double* sig, int32_t v) {
    if (LDS) return sig[v];
    return __hip_atomic_load(sig + v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Push sigma su into neighbour v; true for the one add that discovers v.
template <bool LDS>
__device__ inline bool push_edge(double* sig, int32_t v, double su) {
    if (state<LDS>(sig, v) <= 0.0) return atomicAdd(sig + v, -su) == 0.0;
    return false;
}

// Wavefront-aggregated queue append of the lanes' claimed candidates (SL
// slots per lane): one counter add per wavefront.  Every lane calls it.
template <int SL>
__device__ inline void append_n(const bool* c, const int32_t* v, int* counter, int32_t* queue, int lane) {
    uint64_t m[SL];
    int tot = 0;
#pragma unroll
    for (int k = 0; k < SL; k++) {
        m[k] = __ballot(c[k]);
        tot += __popcll(m[k]);
    }
    if (tot) {
        int base = 0;
        if (lane == 0) base = atomicAdd(counter, tot);
        base = __shfl(base, 0);
        const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
        for (int k = 0; k < SL; k++) {
            if (c[k]) queue[base + __popcll(m[k] & lt)] = v[k];
            base += __popcll(m[k]);
        }
    }
}

template <int W>
__device__ inline double lane_sum(double x) {
#pragma unroll
    for (int o = 1; o < W; o <<= 1) x += __shfl_xor(x, o);
    return x;
}

// Rows of one pass are taken kRows = 64 / KG at a time per wavefront.  The
// loads are software-pipelined over three row blocks: while block k is
// expanded, the row-pointer pair (and, backward, the partial betweenness)
// of block k+1 and the queue entries of block k+2 are already in flight, so
// one memory round trip is exposed per block instead of three.
struct RowBlock {
    int32_t u;       // vertex (-1: none)
    int32_t b, e1;   // adjacency range
    double pu;       // backward: part[u] before this source's delta
};

template <bool BWD>
__device__ inline void row_fetch(RowBlock& rb, const int32_t* __restrict__ rp, const double* part) {
    rb.b = rb.u >= 0 ? rp[rb.u] : 0;
    rb.e1 = rb.u >= 0 ? rp[rb.u + 1] : 0;
    if (BWD) rb.pu = rb.u >= 0 ? part[rb.u] : 0.0;
}

// One single-source Brandes search by a team; `t` = thread index in the
// team, sig = zeroed state array of the graph, counter = the team's LDS
// queue counter.  Leaves sig zeroed.
template <int TEAM, bool LDS, int KG, int SL>
__device__ void brandes_source(const int32_t* __restrict__ rp, const int32_t* __restrict__ col, int32_t s,
                               double* sig, int32_t* queue, int32_t* loff, double* coef, double* part,
                               int* counter, int t) {
    constexpr int NW = TEAM / 64;
    constexpr int kRows = 64 / KG;
    constexpr int kHeavy = 2 * KG * SL; // longer rows use a whole wavefront
    const int lane = t & 63, w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int gl = lane & (KG - 1), grp = lane / KG;
    const int32_t step = NW * kRows;
    if (t == 0) {
        sig[s] = 1.0;
        queue[0] = s;
        loff[0] = 0;
        loff[1] = 1;
        *counter = 1;
    }
    team_sync<TEAM>();
    int32_t L = 0, lo = 0, hi = 1;
    for (;;) {
        {
            int32_t r0 = lo + w * kRows;
            RowBlock cur, nxt;
            cur.u = r0 + grp < hi ? queue[r0 + grp] : -1;
            nxt.u = r0 + step + grp < hi ? queue[r0 + step + grp] : -1;
            row_fetch<false>(cur, rp, part);
            for (; r0 < hi; r0 += step) {
                const int32_t i2 = r0 + 2 * step + grp;
                const int32_t u2 = i2 < hi ? queue[i2] : -1;
                row_fetch<false>(nxt, rp, part);
                const int32_t u = cur.u;
                int32_t b = cur.b, e1 = cur.e1;
                const double su = u >= 0 ? state<LDS>(sig, u) : 0.0;
                const bool heavy = e1 - b > kHeavy;
                uint64_t hm = __ballot(heavy && gl == 0);
                while (hm) {
                    const int j = __ffsll((unsigned long long)hm) - 1;
                    hm &= hm - 1;
                    const int32_t hb = __shfl(b, j), he = __shfl(e1, j);
                    const double hs = __shfl(su, j);
                    for (int32_t e = hb + lane; __ballot(e < he); e += 128) {
                        const bool a0 = e < he, a1 = e + 64 < he;
                        const int32_t v0 = a0 ? col[e] : 0;
                        const int32_t v1 = a1 ? col[e + 64] : 0;
                        const bool cc[2] = {a0 && push_edge<LDS>(sig, v0, hs), a1 && push_edge<LDS>(sig, v1, hs)};
                        const int32_t vv[2] = {v0, v1};
                        append_n<2>(cc, vv, counter, queue, lane);
                    }
                }
                if (heavy) e1 = b;
                for (int32_t e = b + gl; __ballot(e < e1); e += SL * KG) {
                    int32_t vv[SL];
                    bool cc[SL];
#pragma unroll
                    for (int k = 0; k < SL; k++) vv[k] = e + k * KG < e1 ? col[e + k * KG] : 0;
#pragma unroll
                    for (int k = 0; k < SL; k++) cc[k] = e + k * KG < e1 && push_edge<LDS>(sig, vv[k], su);
                    append_n<SL>(cc, vv, counter, queue, lane);
                }
                cur = nxt;
                nxt.u = u2;
            }
        }
        team_sync<TEAM>();
        const int32_t qn = *counter;
        if (qn == hi) break;
        for (int32_t i = hi + t; i < qn; i += TEAM) {
            const int32_t v = queue[i];
            sig[v] = -state<LDS>(sig, v);
        }
        if (t == 0) loff[L + 2] = qn;
        L++;
        lo = hi;
        hi = qn;
        team_sync<TEAM>();
    }
    // backward, deepest level first; level 0 (the source) adds nothing
    for (int32_t l = L; l >= 1; l--) {
        const int32_t a0 = loff[l], a1 = loff[l + 1];
        int32_t r0 = a0 + w * kRows;
        RowBlock cur, nxt;
        cur.u = r0 + grp < a1 ? queue[r0 + grp] : -1;
        nxt.u = r0 + step + grp < a1 ? queue[r0 + step + grp] : -1;
        row_fetch<true>(cur, rp, part);
        for (; r0 < a1; r0 += step) {
            const int32_t i2 = r0 + 2 * step + grp;
            const int32_t u2 = i2 < a1 ? queue[i2] : -1;
            row_fetch<true>(nxt, rp, part);
            const int32_t i = r0 + grp;
            const int32_t u = cur.u;
            int32_t b = cur.b, e1 = cur.e1;
            const bool heavy = e1 - b > kHeavy;
            uint64_t hm = __ballot(heavy && gl == 0);
            while (hm) {
                const int j = __ffsll((unsigned long long)hm) - 1;
                hm &= hm - 1;
                const int32_t hb = __shfl(b, j), he = __shfl(e1, j);
                double acc = 0.0;
                for (int32_t e = hb + lane; e < he; e += 128) {
                    const int32_t w1 = e + 64 < he ? col[e + 64] : -1;
                    const double c0 = state<LDS>(sig, col[e]);
                    const double c1 = w1 >= 0 ? state<LDS>(sig, w1) : 0.0;
                    if (c0 < 0.0) acc -= c0;
                    if (c1 < 0.0) acc -= c1;
                }
                acc = lane_sum<64>(acc);
                if (lane == j) {
                    const double su = state<LDS>(sig, u);
                    const double d = su * acc;
                    part[u] = cur.pu + d;
                    coef[i] = (1.0 + d) / su;
                }
            }
            if (heavy) e1 = b;
            double acc = 0.0;
            for (int32_t e = b + gl; e < e1; e += SL * KG) {
                int32_t ww[SL];
#pragma unroll
                for (int k = 0; k < SL; k++) ww[k] = e + k * KG < e1 ? col[e + k * KG] : -1;
#pragma unroll
                for (int k = 0; k < SL; k++) {
                    const double c = ww[k] >= 0 ? state<LDS>(sig, ww[k]) : 0.0;
                    if (c < 0.0) acc -= c;
                }
            }
            acc = lane_sum<KG>(acc);
            if (u >= 0 && !heavy && gl == 0) {
                const double su = state<LDS>(sig, u);
                const double d = su * acc;
                part[u] = cur.pu + d;
                coef[i] = (1.0 + d) / su;
            }
            cur = nxt;
            nxt.u = u2;
        }
        team_sync<TEAM>();
        for (int32_t i = a0 + t; i < a1; i += TEAM) sig[queue[i]] = -coef[i];
        team_sync<TEAM>();
    }
    for (int32_t i = t; i < hi; i += TEAM) sig[queue[i]] = 0.0;
    team_sync<TEAM>();
}

template <int KG, int SL>
__global__ __launch_bounds__(kBB) void bc_batch_kernel(BatchArgs a) {
    extern __shared__ double s_sig[];
    __shared__ int s_task;
    __shared__ int s_qc[kBW];
    const int tid = threadIdx.x, lane = tid & 63, wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    int32_t* queue_b = a.queue + (size_t)blockIdx.x * a.scr;
    double* coef_b = a.coef + (size_t)blockIdx.x * a.scr;
    int32_t* loff_b = a.loff + (size_t)blockIdx.x * (a.scr + 2 * kBW);
    uint32_t empty = 0;   // queues this work-group has found empty (thread 0)
    for (;;) {
        if (tid == 0) {
            int task = -1;
            for (int k = 0; k < kXq && task < 0; k++) {
                const int q = (int)((blockIdx.x + (unsigned)k) % kXq);
                if ((empty >> q) & 1u) continue;
                const int len = a.qoff[q + 1] - a.qoff[q];
                const int t = len > 0 ? atomicAdd(&a.task_ctr[q * 32], 1) : len;
                if (t < len) task = a.qoff[q] + t;
                else empty |= 1u << q;
            }
            s_task = task;
        }
        __syncthreads();
        const int task = s_task;
        __syncthreads();
        if (task < 0) break;
        const int32_t gi = a.task_graph[task];
        const int32_t n = a.nv[gi];
        const int32_t s0 = a.task_s0[task], s1 = a.task_s1[task];
        const int32_t* rp = a.rp_cat + a.row_off[gi];
        const int32_t* col = a.col_cat + a.col_off[gi];
        double* part = a.part + a.task_part[task];
        if (n <= a.wave_max_n) {
            // ---------------- wave mode ----------------
            const int32_t wn = a.wave_max_n;
            double* sig = s_sig + (size_t)wid * wn;
            int32_t* queue = queue_b + (size_t)wid * wn;
            double* coef = coef_b + (size_t)wid * wn;
            int32_t* loff = loff_b + (size_t)wid * (wn + 2);
            double* wp = a.wpart + ((size_t)blockIdx.x * kBW + wid) * wn;
            for (int32_t v = lane; v < n; v += 64) {
                sig[v] = 0.0;
                wp[v] = 0.0;
            }
            team_sync<64>();
            for (int32_t s = s0 + wid; s < s1; s += kBW)
                brandes_source<64, true, KG, SL>(rp, col, s, sig, queue, loff, coef, wp, &s_qc[wid], lane);
            __syncthreads();
            // fold the wavefront partials into the task partial, fixed order
            for (int32_t v = tid; v < n; v += kBB) {
                double acc = 0.0;
                for (int q = 0; q < kBW; q++) acc += a.wpart[((size_t)blockIdx.x * kBW + q) * wn + v];
                part[v] = acc;
            }
            __syncthreads();
        } else if (n <= a.lds_max_n) {
            // ---------------- block mode, state in LDS ----------------
            for (int32_t v = tid; v < n; v += kBB) {
                s_sig[v] = 0.0;
                part[v] = 0.0;
            }
            __syncthreads();
            for (int32_t s = s0; s < s1; s++)
                brandes_source<kBB, true, KG, SL>(rp, col, s, s_sig, queue_b, loff_b, coef_b, part, &s_qc[0], tid);
            __syncthreads();
        } else {
            // ---------------- block mode, state in global scratch ----------------
            double* gs = a.gsig + (size_t)blockIdx.x * a.max_n;
            for (int32_t v = tid; v < n; v += kBB) {
                gs[v] = 0.0;
                part[v] = 0.0;
            }
            __syncthreads();
            for (int32_t s = s0; s < s1; s++)
                brandes_source<kBB, false, KG, SL>(rp, col, s, gs, queue_b, loff_b, coef_b, part, &s_qc[0], tid);
            __syncthreads();
        }
    }
}

// bc of every graph: sum of its tasks' partials in task order times the
// _rescale! scale.  One thread per output slot; slots in gaps between the
// graphs' ranges are left alone.
__global__ void bc_batch_finish(int32_t ngraphs, const int32_t* nv, const int64_t* bc_off, const int32_t* gtask0,
                                const int32_t* gtask_n, const int32_t* gtasks, const int64_t* task_part,
                                const double* part, const double* scale, double* out, int64_t total) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        // find the graph of output i (binary search over bc_off)
        int lo = 0, hi = ngraphs - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (bc_off[mid] <= i) lo = mid; else hi = mid - 1;
        }
        const int g = lo;
        const int64_t v = i - bc_off[g];
        if (v >= nv[g]) continue;
        double s = 0.0;
        for (int k = 0; k < gtask_n[g]; k++) s += part[task_part[gtasks[gtask0[g] + k]] + v];
        out[i] = s * scale[g];
    }
}

struct BatchTuning {
    int32_t wave_max_n = -1;        // -1: pick from LDS capacity
    double task_work_target = 0.0;  // 0: automatic
    int32_t lds_max_n = -1;         // -1: LDS capacity
    int32_t lane_group = 2;         // lanes per light row: 1, 2, 4 or 8
    int32_t slots = 8;              // adjacency entries per lane per round: 2, 4, 8 or 16
};

template <int KG>
const void* batch_kernel_sl(int sl) {
    if constexpr (KG == 1) {
        if (sl == 16) return (const void*)bc_batch_kernel<KG, 16>;
    }
    if (sl == 8) return (const void*)bc_batch_kernel<KG, 8>;
    if (sl == 2) return (const void*)bc_batch_kernel<KG, 2>;
    return (const void*)bc_batch_kernel<KG, 4>;
}

const void* batch_kernel(int kg, int sl) {
    if (kg == 1) return batch_kernel_sl<1>(sl);
    if (kg == 2) return batch_kernel_sl<2>(sl);
    if (kg == 8) return batch_kernel_sl<8>(sl);
    return batch_kernel_sl<4>(sl);
}

BatchTuning& tuning() {
    static BatchTuning t;
    return t;
}

}  // namespace
}  // namespace lgmi

using namespace lgmi;

extern "C" int lgmi_betweenness_batch_tuning(int32_t wave_max_n, double task_work_target, int32_t lds_max_n,
                                             int32_t lane_group, int32_t slots) {
    try {
        if (wave_max_n < -1 || !(task_work_target >= 0) || lds_max_n < -1 ||
            (lane_group != 0 && lane_group != 1 && lane_group != 2 && lane_group != 4 && lane_group != 8) ||
            (slots != 0 && slots != 2 && slots != 4 && slots != 8 && !(slots == 16 && lane_group == 1)))
            throw Error(LGMI_E_ARG, "bad tuning values");
        tuning().lane_group = lane_group ? lane_group : 2;
        tuning().slots = slots ? slots : 8;
        tuning().wave_max_n = wave_max_n;
        tuning().task_work_target = task_work_target;
        tuning().lds_max_n = lds_max_n;
        return LGMI_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    }
}

extern "C" int lgmi_betweenness_batch(int32_t ngraphs, const int32_t* nv, const int64_t* row_off,
                                      const int32_t* row_ptr_cat, const int64_t* col_off, const int32_t* col_cat,
                                      const int64_t* bc_off, int normalize, int endpoints, double* bc_cat,
                                      double* device_ms) {
    try {
        if (ngraphs < 0 || (ngraphs > 0 && (!nv || !row_off || !row_ptr_cat || !col_off || !col_cat || !bc_off || !bc_cat)))
            throw Error(LGMI_E_ARG, "bad arguments");
        if (device_ms) *device_ms = 0.0;
        if (ngraphs == 0) return LGMI_OK;
        // ---- host-side validation and layout ----
        int32_t max_n = 0;
        int64_t rp_total = 0, col_total = 0, out_total = 0;
        for (int32_t g = 0; g < ngraphs; g++) {
            const int32_t n = nv[g];
            if (n < 0) throw Error(LGMI_E_ARG, "nv < 0");
            const int32_t* rp = row_ptr_cat + row_off[g];
            const int32_t* col = col_cat + col_off[g];
            if (rp[0] != 0) throw Error(LGMI_E_ARG, "row_ptr must start at 0");
            for (int32_t v = 0; v < n; v++) {
                if (rp[v + 1] < rp[v]) throw Error(LGMI_E_ARG, "row_ptr not monotone");
                for (int32_t e = rp[v]; e < rp[v + 1]; e++) {
                    if (col[e] < 0 || col[e] >= n) throw Error(LGMI_E_BOUNDS, "neighbour id out of range");
                    if (e > rp[v] && col[e - 1] >= col[e]) throw Error(LGMI_E_ARG, "rows must be sorted and unique");
                }
            }
            if (g > 0 && bc_off[g] < bc_off[g - 1] + nv[g - 1])
                throw Error(LGMI_E_ARG, "bc_off must be ascending and non-overlapping");
            max_n = std::max(max_n, n);
            rp_total = std::max<int64_t>(rp_total, row_off[g] + n + 1);
            col_total = std::max<int64_t>(col_total, col_off[g] + rp[n]);
            out_total = std::max<int64_t>(out_total, bc_off[g] + n);
        }
        ensure_device();
        int dev = 0;
        LGMI_HIP(hipGetDevice(&dev));
        int cu = device_cu_count();
        int lds_block = 0;
        LGMI_HIP(hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        const size_t static_lds = 64;
        const size_t dyn_lds = (size_t)std::max(0, lds_block - 1024);
        const int32_t lds_cap = (int32_t)(dyn_lds / sizeof(double));
        const int32_t lds_max_n = tuning().lds_max_n >= 0 ? std::min(tuning().lds_max_n, lds_cap) : lds_cap;
        int32_t wave_max_n = tuning().wave_max_n >= 0 ? tuning().wave_max_n : (int32_t)std::min<int64_t>(
            lds_cap / kBW, 1024);
        wave_max_n = std::min<int32_t>(wave_max_n, lds_cap / kBW);
        (void)static_lds;

        // ---- tasks: split graphs by work, largest first ----
        std::vector<double> work((size_t)ngraphs);
        for (int32_t g = 0; g < ngraphs; g++) {
            const int32_t* rp = row_ptr_cat + row_off[g];
            work[(size_t)g] = (double)nv[g] * (double)(rp[nv[g]] + nv[g]);
        }
        const int nblocks = std::max(1, cu);
        // Task size: by default the largest graph is cut into about 128 tasks,
        // so that the 32 work-groups of its queue's XCD all work on it, with
        // the tasks' partial arrays held to 8 GB (each task keeps one FP64
        // value per vertex of its graph; the target doubles until they fit).
        double target = tuning().task_work_target;
        if (!(target > 0)) {
            double wmax = 1.0;
            for (int32_t g = 0; g < ngraphs; g++) wmax = std::max(wmax, work[(size_t)g]);
            target = std::max(1.0, wmax / 128.0);
            auto part_bytes = [&](double t) {
                double b = 0.0;
                for (int32_t g = 0; g < ngraphs; g++)
                    b += std::max(1.0, std::min((double)nv[g], std::ceil(work[(size_t)g] / t))) * (double)nv[g] * 8.0;
                return b;
            };
            while (part_bytes(target) > 8e9) target *= 2.0;
        }
        struct Task { int32_t g, s0, s1; double w; };
        std::vector<Task> tasks;
        for (int32_t g = 0; g < ngraphs; g++) {
            const int32_t n = nv[g];
            if (n == 0) continue;
            int32_t chunks = (int32_t)std::min<double>(n, std::ceil(work[(size_t)g] / target));
            if (n <= wave_max_n) chunks = std::min<int32_t>(chunks, std::max(1, n / kBW));
            chunks = std::max(1, chunks);
            for (int32_t c = 0; c < chunks; c++) {
                const int32_t s0 = (int32_t)((int64_t)n * c / chunks), s1 = (int32_t)((int64_t)n * (c + 1) / chunks);
                if (s1 > s0) tasks.push_back({g, s0, s1, work[(size_t)g] * (double)(s1 - s0) / (double)n});
            }
        }
        // graphs to queues: largest work first, each to the least-loaded
        // queue; a queue's tasks in decreasing work
        std::vector<int32_t> gq((size_t)ngraphs, 0);
        {
            std::vector<int32_t> gord((size_t)ngraphs);
            std::iota(gord.begin(), gord.end(), 0);
            std::stable_sort(gord.begin(), gord.end(),
                             [&](int32_t x, int32_t y) { return work[(size_t)x] > work[(size_t)y]; });
            double load[kXq] = {};
            for (int32_t g : gord) {
                int q = 0;
                for (int k = 1; k < kXq; k++)
                    if (load[k] < load[q]) q = k;
                gq[(size_t)g] = q;
                load[q] += work[(size_t)g];
            }
        }
        std::vector<int32_t> order(tasks.size());
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
            const int32_t qx = gq[(size_t)tasks[x].g], qy = gq[(size_t)tasks[y].g];
            if (qx != qy) return qx < qy;
            return tasks[x].w > tasks[y].w;
        });
        std::vector<int32_t> qoff(kXq + 1, 0);
        for (const Task& t : tasks) qoff[(size_t)gq[(size_t)t.g] + 1]++;
        for (int q = 0; q < kXq; q++) qoff[(size_t)q + 1] += qoff[(size_t)q];
        const int32_t ntasks = (int32_t)tasks.size();
        std::vector<int32_t> tg(ntasks), ts0(ntasks), ts1(ntasks);
        std::vector<int64_t> tpart(ntasks);
        int64_t part_total = 0;
        for (int32_t i = 0; i < ntasks; i++) {
            const Task& t = tasks[order[i]];
            tg[i] = t.g; ts0[i] = t.s0; ts1[i] = t.s1;
            tpart[i] = part_total;
            part_total += nv[t.g];
        }
        // per graph: its tasks in source order (the order their partials are summed)
        std::vector<int32_t> gtask0(ngraphs, 0), gtask_n(ngraphs, 0), gtasks(ntasks);
        {
            std::vector<std::vector<std::pair<int32_t, int32_t>>> per(ngraphs);
            for (int32_t i = 0; i < ntasks; i++) per[tg[i]].push_back({ts0[i], i});
            int32_t k = 0;
            for (int32_t g = 0; g < ngraphs; g++) {
                std::sort(per[g].begin(), per[g].end());
                gtask0[g] = k;
                gtask_n[g] = (int32_t)per[g].size();
                for (auto& pr : per[g]) gtasks[k++] = pr.second;
            }
        }
        // _rescale! per graph (k = n)
        std::vector<double> scale(ngraphs);
        for (int32_t g = 0; g < ngraphs; g++) {
            const int64_t n = nv[g];
            double sc = 1.0;
            bool do_scale = false;
            if (normalize) {
                if (n > 2) { do_scale = true; sc = 1.0 / (double)((n - 1) * (n - 2)); }
            } else {
                do_scale = true;
                sc = 1.0 / 2.0;
            }
            if (do_scale && n > 0) sc = sc * (double)n / (double)n;
            scale[g] = do_scale ? sc : 1.0;
        }

        // ---- device buffers ----
        hipStream_t st;
        LGMI_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        std::vector<void*> bufs;
        auto dalloc = [&](size_t bytes) {
            void* p = dev_alloc(bytes);
            bufs.push_back(p);
            return p;
        };
        auto up = [&](const void* h, size_t bytes) {
            void* p = dalloc(bytes);
            if (bytes) LGMI_HIP(hipMemcpyAsync(p, h, bytes, hipMemcpyHostToDevice, st));
            return p;
        };
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipGraph_t graph = nullptr;
        hipGraphExec_t gexec = nullptr;
        try {
            BatchArgs a{};
            a.nv = (const int32_t*)up(nv, sizeof(int32_t) * ngraphs);
            a.row_off = (const int64_t*)up(row_off, sizeof(int64_t) * ngraphs);
            a.col_off = (const int64_t*)up(col_off, sizeof(int64_t) * ngraphs);
            a.rp_cat = (const int32_t*)up(row_ptr_cat, sizeof(int32_t) * (size_t)rp_total);
            a.col_cat = (const int32_t*)up(col_cat, sizeof(int32_t) * (size_t)std::max<int64_t>(col_total, 1));
            a.task_graph = (const int32_t*)up(tg.data(), sizeof(int32_t) * ntasks);
            a.task_s0 = (const int32_t*)up(ts0.data(), sizeof(int32_t) * ntasks);
            a.task_s1 = (const int32_t*)up(ts1.data(), sizeof(int32_t) * ntasks);
            a.task_part = (const int64_t*)up(tpart.data(), sizeof(int64_t) * ntasks);
            a.ntasks = ntasks;
            a.wave_max_n = wave_max_n;
            a.lds_max_n = lds_max_n;
            a.max_n = max_n;
            a.part = (double*)dalloc(sizeof(double) * (size_t)std::max<int64_t>(part_total, 1));
            a.task_ctr = (int32_t*)dalloc(sizeof(int32_t) * kXq * 32);
            LGMI_HIP(hipMemsetAsync(a.task_ctr, 0, sizeof(int32_t) * kXq * 32, st));
            a.qoff = (const int32_t*)up(qoff.data(), sizeof(int32_t) * (kXq + 1));
            a.scr = std::max<int64_t>({(int64_t)max_n, (int64_t)kBW * wave_max_n, 1});
            a.queue = (int32_t*)dalloc(sizeof(int32_t) * (size_t)nblocks * a.scr);
            a.coef = (double*)dalloc(sizeof(double) * (size_t)nblocks * a.scr);
            a.loff = (int32_t*)dalloc(sizeof(int32_t) * (size_t)nblocks * (a.scr + 2 * kBW));
            a.gsig = (double*)dalloc(sizeof(double) * (size_t)nblocks * (max_n > lds_max_n ? max_n : 1));
            a.wpart = (double*)dalloc(sizeof(double) * (size_t)nblocks * kBW * std::max(wave_max_n, 1));
            double* d_out = (double*)dalloc(sizeof(double) * (size_t)std::max<int64_t>(out_total, 1));
            const int64_t* d_bcoff = (const int64_t*)up(bc_off, sizeof(int64_t) * ngraphs);
            const int32_t* d_gt0 = (const int32_t*)up(gtask0.data(), sizeof(int32_t) * ngraphs);
            const int32_t* d_gtn = (const int32_t*)up(gtask_n.data(), sizeof(int32_t) * ngraphs);
            const int32_t* d_gts = (const int32_t*)up(gtasks.data(), sizeof(int32_t) * std::max(ntasks, 1));
            const double* d_scale = (const double*)up(scale.data(), sizeof(double) * ngraphs);
            LGMI_HIP(hipStreamSynchronize(st));
            const void* kern = batch_kernel(tuning().lane_group, tuning().slots);
            LGMI_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn_lds));
            LGMI_HIP(hipEventCreate(&e0));
            LGMI_HIP(hipEventCreate(&e1));
            // The persistent launch and the reduction over tasks are captured
            // in one hipGraph (the task counters are cleared above, outside
            // the graph) and replayed once.
            void* kargs[] = {&a};
            LGMI_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            hipError_t ce = hipLaunchKernel(kern, dim3(nblocks), dim3(kBB), kargs, dyn_lds, st);
            if (ce == hipSuccess) {
                bc_batch_finish<<<std::max(1, std::min(cu * 8, (int)((out_total + 255) / 256))), 256, 0, st>>>(
                    ngraphs, a.nv, d_bcoff, d_gt0, d_gtn, d_gts, a.task_part, a.part, d_scale, d_out, out_total);
                ce = hipGetLastError();
            }
            const hipError_t ee = hipStreamEndCapture(st, &graph);
            LGMI_HIP(ce);
            LGMI_HIP(ee);
            LGMI_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
            LGMI_HIP(hipEventRecord(e0, st));
            LGMI_HIP(hipGraphLaunch(gexec, st));
            LGMI_HIP(hipEventRecord(e1, st));
            std::vector<double> hout((size_t)out_total);
            LGMI_HIP(hipMemcpyAsync(hout.data(), d_out, sizeof(double) * (size_t)out_total, hipMemcpyDeviceToHost,
                                    st));
            LGMI_HIP(hipStreamSynchronize(st));
            for (int32_t g = 0; g < ngraphs; g++)
                std::copy(hout.begin() + bc_off[g], hout.begin() + bc_off[g] + nv[g], bc_cat + bc_off[g]);
            float ms = 0.f;
            LGMI_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (device_ms) *device_ms = ms;
        } catch (...) {
            (void)hipStreamSynchronize(st);
            if (gexec) (void)hipGraphExecDestroy(gexec);
            if (graph) (void)hipGraphDestroy(graph);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            for (void* p : bufs) dev_free(p);
            (void)hipStreamDestroy(st);
            throw;
        }
        (void)hipGraphExecDestroy(gexec);
        (void)hipGraphDestroy(graph);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        for (void* p : bufs) dev_free(p);
        (void)hipStreamDestroy(st);
        // endpoints: every source with degree > 0 adds n-1 to itself and 1 to
        // each other vertex (betweenness.jl:122,130); added before scaling
        if (endpoints) {
            for (int32_t g = 0; g < ngraphs; g++) {
                const int32_t n = nv[g];
                const int32_t* rp = row_ptr_cat + row_off[g];
                int64_t nact = 0;
                for (int32_t v = 0; v < n; v++) nact += rp[v + 1] > rp[v];
                for (int32_t v = 0; v < n; v++) {
                    const bool act = rp[v + 1] > rp[v];
                    const double add = (double)(nact - (act ? 1 : 0)) + (act ? (double)(n - 1) : 0.0);
                    bc_cat[bc_off[g] + v] += add * scale[g];
                }
            }
        }
        return LGMI_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return LGMI_E_HIP;
    }
}
