// bfs.hip -- direction-optimising single/multi-source BFS for gfx950.
//
// Replaces the level loop of gdistances! (src/traversals/bfs.jl:95-124),
// _bfs_parents (bfs.jl:38-66) and the threaded Parallel.gdistances! /
// Parallel.bfs_tree! (src/Parallel/traversals/gdistances.jl:36-116,
// src/Parallel/traversals/bfs.jl:49-92).
//
// State in HBM: dist[nv], parent[nv] (Int32), three frontier bitmaps F_L
// rotating by level (read F_L, write F_{L+1}, clear F_{L+2}) and the visited
// bitmap vis = F_0 | ... | F_{L-1} (F_L is folded in during step L).  No
// vertex queues: every step sweeps bitmap tiles of 64 words (2048 vertices),
// one wavefront per tile; the set bits of interest (frontier vertices for
// top-down, unvisited vertices with a non-empty row for bottom-up) are
// compacted into batches of 64 vertices with a prefix popcount over the
// lanes and a rank-select inside the word, so every lane of a batch works.
//
// One launch of bfs_step advances the search by one level L -> L+1:
//  * top-down: rows of 32..1024 entries are expanded by the whole
//    wavefront, shorter rows by a scan-based gather that packs several rows
//    into the 64 lanes, and rows longer than 1024 entries through
//    precomputed 1024-entry work items spread over all work-groups.  A new
//    vertex is claimed with an atomicOr on F_{L+1}; its parent is the
//    minimum over its frontier neighbours (atomicMin after a plain-load
//    filter).
//  * bottom-up: every candidate scans its reverse adjacency in ascending
//    order, eight entries per round trip, and stops at the first neighbour
//    in F_L, which is the minimum-id parent; the tile's F_{L+1} words are
//    gathered in LDS and written whole.
// Both directions give identical distances and parents: the reference's
// sorted frontier makes parent[v] the smallest u with dist[u] = dist[v]-1
// and an edge u -> v, which is what either direction computes.
//
// Every block writes its counters (new vertices, their degrees, entries
// scanned, a degree histogram) to its own slot of a partials array; the
// one-block bfs_finalize kernel that follows each step sums them in block
// order, updates the control block and picks the next direction (Beamer
// alpha/beta, the learned MLP, or a forced direction).  The host enqueues
// step/finalise pairs and synchronises once per batch of levels.
#include "lgmi_internal.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace lgmi {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int32_t kWaveRow = 32;    // top-down: rows this long or longer are expanded by a wavefront
constexpr int32_t kBuCoop = 64;     // bottom-up: rows longer than this are scanned by a wavefront
constexpr int kFinBlock = 1024;     // finalise kernel threads
constexpr int kGraphLevels = 16;    // step/finalise pairs per captured graph

struct StepArgs {
    const int32_t* td_rp;     // traversal adjacency (top-down expands it)
    const int32_t* td_col;
    const int32_t* td_item_v; // heavy-row items of the traversal adjacency
    const int32_t* td_item_beg;
    int32_t td_items;
    const int32_t* bu_rp;     // reverse adjacency (bottom-up scans it)
    const int32_t* bu_col;
    const uint32_t* bu_nz;    // bit v: row v of the reverse adjacency is non-empty
    int32_t nv;
    int32_t nwords;           // 2 * ceil(nv / 64)
    int32_t* dist;
    uint32_t* parent;         // 0xffffffff = unset (nullptr: no parents)
    uint32_t* bm0;
    uint32_t* bm1;
    uint32_t* bm2;
    uint32_t* vis;
    BfsCtrl* ctrl;
    unsigned long long* part; // [gridDim.x][kPart]
    LevelRec* log;
    int32_t max_log;
    const PolicyWeights* policy;
};

__device__ inline uint32_t* bm_at(const StepArgs& a, int k) {
    return k == 0 ? a.bm0 : (k == 1 ? a.bm1 : a.bm2);
}

__device__ inline int hist_bin(int32_t deg) {
    // powers of four: 0, 1-3, 4-15, 16-63, 64-255, 256-1023, >= 1024
    if (deg <= 0) return 0;
    const int b = (31 - __clz(deg)) / 2 + 1;
    return b > 6 ? 6 : b;
}

__device__ inline float flog2p(double x) { return (float)log2(1.0 + (x > 0 ? x : 0.0)); }

// Position of the r-th (0-based) set bit of x (r < popcount(x)).
__device__ inline int select_bit(uint32_t x, int r) {
    int pos = 0, c;
    c = __popc(x & 0xffffu); if (r >= c) { r -= c; x >>= 16; pos += 16; }
    c = __popc(x & 0xffu);   if (r >= c) { r -= c; x >>= 8;  pos += 8; }
    c = __popc(x & 0xfu);    if (r >= c) { r -= c; x >>= 4;  pos += 4; }
    c = __popc(x & 0x3u);    if (r >= c) { r -= c; x >>= 2;  pos += 2; }
    c = (int)(x & 1u);       if (r >= c) { pos += 1; }
    return pos;
}

// Lane `lane` of a batch starting at set-bit index `base` of the tile's
// words (lane l holds word cw with exclusive popcount prefix excl) gets
// the vertex of set bit base+lane, or -1 past the end.  Must be called by
// all 64 lanes.  wi/bi return the word index in the tile and the bit.
__device__ inline int32_t batch_vertex(uint32_t cw, int32_t excl, int32_t total, int32_t base, int32_t word0,
                                       int lane, int& wi, int& bi) {
    const int32_t g = base + lane;
    int lo = 0;
    #pragma unroll
    for (int step = 32; step >= 1; step >>= 1) {
        const int probe = lo + step;
        const int32_t pv = __shfl(excl, probe & 63);
        if (probe < 64 && pv <= g) lo = probe;
    }
    const uint32_t word = __shfl(cw, lo);
    const int32_t r = g - __shfl(excl, lo);
    wi = lo;
    bi = 0;
    if (g >= total) return -1;
    bi = select_bit(word, r);
    return (word0 + lo) * 32 + bi;
}

// Features of the state at the start of a step (LGMI_POLICY_FEATURES).
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

template <bool PARENTS>
__device__ inline bool td_visit(const StepArgs& a, int32_t u, int32_t v, int32_t L, const uint32_t* FL,
                                uint32_t* Fn) {
    const int w = v >> 5;
    const uint32_t bit = 1u << (v & 31);
    if ((a.vis[w] | FL[w]) & bit) return false;       // distance <= L already
    if (PARENTS) {
        // parent only decreases, so a stale value <= u already rules u out
        if (a.parent[v] > (uint32_t)u) atomicMin(&a.parent[v], (uint32_t)u);
    }
    if (Fn[w] & bit) return false;
    const uint32_t old = atomicOr(&Fn[w], bit);
    if (old & bit) return false;
    a.dist[v] = L + 1;
    return true;
}

struct Counters {
    unsigned long long n = 0, out = 0, in = 0, scan = 0, cand = 0;
};

template <bool PARENTS>
__global__ __launch_bounds__(kBlock) void bfs_step(StepArgs a) {
    __shared__ int32_t s_level, s_dir, s_done;
    __shared__ uint32_t s_hist[kHistBins];
    __shared__ unsigned long long s_red[kWaves][5];
    __shared__ uint32_t s_fn[kWaves][64];

    const BfsCtrl* c = a.ctrl;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (tid == 0) {
        s_done = c->done;
        s_level = c->level;
        s_dir = c->dir;
        if (blockIdx.x == 0 && !s_done) a.ctrl->t_step_start = wall_clock64();
    }
    if (tid < kHistBins) s_hist[tid] = 0;
    __syncthreads();
    if (s_done) return;

    const int32_t L = s_level;
    const uint32_t* FL = bm_at(a, L % 3);
    uint32_t* Fn = bm_at(a, (L + 1) % 3);
    uint32_t* Fc = bm_at(a, (L + 2) % 3);
    const bool same_adj = (a.td_rp == a.bu_rp);
    const int32_t ntiles = (a.nwords + 63) >> 6;
    const int gw = blockIdx.x * kWaves + wid, nw = gridDim.x * kWaves;
    Counters k;

    auto account = [&](bool claimed, int32_t v) {
        if (claimed) {
            const int32_t dout = a.td_rp[v + 1] - a.td_rp[v];
            const int32_t din = same_adj ? dout : (a.bu_rp[v + 1] - a.bu_rp[v]);
            k.n++;
            k.out += (unsigned long long)dout;
            k.in += (unsigned long long)din;
            atomicAdd(&s_hist[hist_bin(dout)], 1u);
        }
    };

    if (s_dir == 0) {
        // ---------------- top-down ----------------
        // heavy rows: 1024-entry items; a wavefront tests 64 items at once
        // (strided by the number of wavefronts, so the items of one hub
        // land on different wavefronts) and expands those on the frontier
        for (int64_t base = 0; (int64_t)gw + base * nw < a.td_items; base += 64) {
            const int64_t it = (int64_t)gw + (base + lane) * nw;
            int32_t iu = -1;
            if (it < a.td_items) {
                iu = a.td_item_v[it];
                if (!((FL[iu >> 5] >> (iu & 31)) & 1u)) iu = -1;
            }
            uint64_t m = __ballot(iu >= 0);
            while (m) {
                const int j = __ffsll((unsigned long long)m) - 1;
                m &= m - 1;
                const int32_t u = __shfl(iu, j);
                const int32_t b = a.td_item_beg[(int64_t)gw + (base + j) * nw];
                const int32_t e = min(b + kHeavyRow, a.td_rp[u + 1]);
                k.scan += (lane == 0) ? (unsigned long long)(e - b) : 0ull;
                for (int32_t x = b + lane; x < e; x += 64) {
                    const int32_t v = a.td_col[x];
                    account(td_visit<PARENTS>(a, u, v, L, FL, Fn), v);
                }
            }
        }
        // frontier tiles
        for (int32_t t = gw; t < ntiles; t += nw) {
            const int32_t word0 = t * 64;
            const int32_t w = word0 + lane;
            uint32_t fw = 0;
            if (w < a.nwords) {
                fw = FL[w];
                a.vis[w] |= fw;           // fold F_L into vis
                Fc[w] = 0;                // clear F_{L+2}
            }
            const int32_t pc = __popc(fw);
            int32_t incl = pc;
            for (int d = 1; d < 64; d <<= 1) {
                const int32_t y = __shfl_up(incl, d);
                if (lane >= d) incl += y;
            }
            const int32_t excl = incl - pc;
            const int32_t total = __shfl(incl, 63);
            for (int32_t base = 0; base < total; base += 64) {
                int wi, bi;
                const int32_t u = batch_vertex(fw, excl, total, base, word0, lane, wi, bi);
                int32_t beg = 0, deg = 0;
                if (u >= 0) {
                    beg = a.td_rp[u];
                    deg = a.td_rp[u + 1] - beg;
                    if (deg > kHeavyRow) deg = 0;     // expanded through its items
                }
                k.scan += (unsigned long long)deg;
                // wavefront-wide rows
                uint64_t m = __ballot(deg >= kWaveRow);
                while (m) {
                    const int j = __ffsll((unsigned long long)m) - 1;
                    m &= m - 1;
                    const int32_t wu = __shfl(u, j), wb = __shfl(beg, j), wd = __shfl(deg, j);
                    for (int32_t x = lane; x < wd; x += 64) {
                        const int32_t v = a.td_col[wb + x];
                        account(td_visit<PARENTS>(a, wu, v, L, FL, Fn), v);
                    }
                }
                if (deg >= kWaveRow) deg = 0;
                // short rows: scan-based gather over the batch
                int32_t rin = deg;
                for (int d = 1; d < 64; d <<= 1) {
                    const int32_t y = __shfl_up(rin, d);
                    if (lane >= d) rin += y;
                }
                const int32_t rex = rin - deg;
                const int32_t rtot = __shfl(rin, 63);
                for (int32_t off = 0; off < rtot; off += 64) {
                    const int32_t e = off + lane;
                    int lo = 0;
                    #pragma unroll
                    for (int step = 32; step >= 1; step >>= 1) {
                        const int probe = lo + step;
                        const int32_t pv = __shfl(rex, probe & 63);
                        if (probe < 64 && pv <= e) lo = probe;
                    }
                    const int32_t ob = __shfl(beg, lo), oe = __shfl(rex, lo), ou = __shfl(u, lo);
                    if (e < rtot) {
                        const int32_t v = a.td_col[ob + (e - oe)];
                        account(td_visit<PARENTS>(a, ou, v, L, FL, Fn), v);
                    }
                }
            }
        }
    } else {
        // ---------------- bottom-up ----------------
        for (int32_t t = gw; t < ntiles; t += nw) {
            const int32_t word0 = t * 64;
            const int32_t w = word0 + lane;
            uint32_t cw = 0;
            if (w < a.nwords) {
                const uint32_t visw = a.vis[w] | FL[w];
                cw = ~visw & a.bu_nz[w];  // unvisited with a non-empty reverse row
                a.vis[w] = visw;          // fold F_L into vis
                Fc[w] = 0;                // clear F_{L+2}
            }
            s_fn[wid][lane] = 0;
            const int32_t pc = __popc(cw);
            int32_t incl = pc;
            for (int d = 1; d < 64; d <<= 1) {
                const int32_t y = __shfl_up(incl, d);
                if (lane >= d) incl += y;
            }
            const int32_t excl = incl - pc;
            const int32_t total = __shfl(incl, 63);
            k.cand += (lane == 0) ? (unsigned long long)total : 0ull;
            for (int32_t base = 0; base < total; base += 64) {
                int wi, bi;
                const int32_t v = batch_vertex(cw, excl, total, base, word0, lane, wi, bi);
                int32_t beg = 0, end = 0;
                if (v >= 0) { beg = a.bu_rp[v]; end = a.bu_rp[v + 1]; }
                int32_t par = -1;
                const bool big = v >= 0 && (end - beg) > kBuCoop;
                if (v >= 0 && !big) {
                    for (int32_t e = beg; e < end; e += 8) {
                        k.scan += (unsigned long long)min(8, end - e);
                        int32_t uu[8];
                        uint32_t bw[8];
                        #pragma unroll
                        for (int q = 0; q < 8; q++) uu[q] = (e + q < end) ? a.bu_col[e + q] : -1;
                        #pragma unroll
                        for (int q = 0; q < 8; q++) bw[q] = uu[q] >= 0 ? FL[uu[q] >> 5] : 0u;
                        #pragma unroll
                        for (int q = 0; q < 8; q++) {
                            if (par < 0 && uu[q] >= 0 && ((bw[q] >> (uu[q] & 31)) & 1u)) par = uu[q];
                        }
                        if (par >= 0) break;
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
                        if (lane == j) k.scan += (unsigned long long)min(64, je - off);
                        if (hm) { found_u = __shfl(uu, __ffsll((unsigned long long)hm) - 1); break; }
                    }
                    if (lane == j) par = found_u;
                }
                if (par >= 0) {
                    a.dist[v] = L + 1;
                    if (PARENTS) a.parent[v] = (uint32_t)par;
                    atomicOr(&s_fn[wid][wi], 1u << bi);
                }
                account(par >= 0, v);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (w < a.nwords) Fn[w] = s_fn[wid][lane];
            __builtin_amdgcn_wave_barrier();
        }
    }

    // block reduction -> this block's partials slot (plain stores)
    unsigned long long r[5] = {k.n, k.out, k.in, k.scan, k.cand};
    #pragma unroll
    for (int i = 0; i < 5; i++)
        for (int d = 32; d >= 1; d >>= 1) r[i] += __shfl_down(r[i], d);
    if (lane == 0)
        for (int i = 0; i < 5; i++) s_red[wid][i] = r[i];
    __syncthreads();
    if (tid < 5) {
        unsigned long long s = 0;
        for (int q = 0; q < kWaves; q++) s += s_red[q][tid];
        a.part[(size_t)blockIdx.x * kPart + tid] = s;
    } else if (tid >= P_HIST && tid < P_HIST + kHistBins) {
        a.part[(size_t)blockIdx.x * kPart + tid] = s_hist[tid - P_HIST];
    }
}

// The learned policy's weights as lane j of wavefront 0 needs them (hidden
// unit j), loaded at the start of the finalise kernel so that their latency
// overlaps the partial-sum loads.  The device copy has the feature
// standardisation folded into W1 and b1 (done on the host when the policy
// is set), so inference is 16 FMAs per lane and a butterfly sum.
struct PolicyRegs {
    float4 w1[kFeat / 4];
    float b1, w2a, w2b, b2a, b2b;
};

__device__ inline void load_policy(const PolicyWeights* w, int t, PolicyRegs& r) {
    const float4* row = reinterpret_cast<const float4*>(w->w1 + t * kFeat);
    #pragma unroll
    for (int q = 0; q < kFeat / 4; q++) r.w1[q] = row[q];
    r.b1 = w->b1[t];
    r.w2a = w->w2[t];
    r.w2b = w->w2[kHidden + t];
    r.b2a = w->b2[0];
    r.b2b = w->b2[1];
}

// Direction for the next step (called by every thread of the finalise block).
__device__ int choose_direction(const StepArgs& a, const BfsCtrl* c, const float* feat, const PolicyRegs& pr) {
    if (c->forced >= 0) return c->forced;
    if (c->policy_mode == LGMI_POLICY_TOPDOWN) return 0;
    if (c->policy_mode == LGMI_POLICY_LEARNED && a.policy != nullptr) {
        // wavefront 0: lane j = hidden unit j; logits by a butterfly sum
        const int t = threadIdx.x;
        __shared__ int s_dec;
        if (t < kHidden) {
            const float* xs = feat;
            float acc = pr.b1;
            #pragma unroll
            for (int q = 0; q < kFeat / 4; q++) {
                const float4 r = pr.w1[q];
                acc += r.x * xs[4 * q] + r.y * xs[4 * q + 1] + r.z * xs[4 * q + 2] + r.w * xs[4 * q + 3];
            }
            const float hj = acc > 0.f ? acc : 0.f;
            float a0 = pr.w2a * hj, a1 = pr.w2b * hj;
            for (int o = 32; o >= 1; o >>= 1) {
                a0 += __shfl_xor(a0, o);
                a1 += __shfl_xor(a1, o);
            }
            if (t == 0) s_dec = (a1 + pr.b2b) > (a0 + pr.b2a) ? 1 : 0;
        }
        __syncthreads();
        return s_dec;
    }
    // Beamer alpha/beta (GAP form): leave top-down when the frontier's edges
    // exceed the unvisited edges / alpha; leave bottom-up when the frontier
    // shrinks and holds at most nv / beta vertices.
    if (c->prev_dir == 0)
        return ((double)c->m_front > (double)c->m_unvisited / (double)c->alpha) ? 1 : 0;
    const bool shrinking = c->n_front < c->n_prev;
    const bool small = (double)c->n_front <= (double)c->nv / (double)c->beta;
    return (shrinking && small) ? 0 : 1;
}

// One block: sums the step's partials in a fixed order, updates the control
// block and the step log, and picks the next direction.
__global__ __launch_bounds__(kFinBlock) void bfs_finalize(StepArgs a, int32_t nblocks) {
    __shared__ unsigned long long s_sum[kFinBlock / kPart][kPart];
    __shared__ int32_t s_done;
    __shared__ float s_feat[kFeat];
    BfsCtrl* c = a.ctrl;
    const int tid = threadIdx.x;
    if (tid == 0) s_done = c->done;
    __syncthreads();
    if (s_done) return;
    const uint64_t now = wall_clock64();
    PolicyRegs pr;
    if (a.policy != nullptr && tid < kHidden && c->forced < 0 && c->policy_mode == LGMI_POLICY_LEARNED)
        load_policy(a.policy, tid, pr);
    // thread t reads slot t % kPart of blocks t / kPart, + 64, ... (coalesced rows)
    constexpr int kRows = kFinBlock / kPart;
    const int slot = tid % kPart, row = tid / kPart;
    // integer counters: eight independent loads in flight per thread
    unsigned long long acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int b = row;
    for (; b + 7 * kRows < nblocks; b += 8 * kRows) {
        #pragma unroll
        for (int q = 0; q < 8; q++) acc[q] += a.part[(size_t)(b + q * kRows) * kPart + slot];
    }
    for (; b < nblocks; b += kRows) acc[0] += a.part[(size_t)b * kPart + slot];
    unsigned long long s = 0;
    #pragma unroll
    for (int q = 0; q < 8; q++) s += acc[q];
    s_sum[row][slot] = s;
    __syncthreads();
    for (int h = kRows / 2; h >= 1; h >>= 1) {
        if (row < h) s_sum[row][slot] += s_sum[row + h][slot];
        __syncthreads();
    }
    if (tid == 0) {
        const unsigned long long* tot = s_sum[0];
        const int32_t L = c->level;
        const int32_t n_next = (int32_t)tot[P_NEXT];
        if (L < a.max_log) {
            LevelRec* r = &a.log[L];
            r->dir = c->dir;
            r->n_front = c->n_front;
            r->m_front = c->m_front;
            r->m_unvisited = c->m_unvisited;
            r->n_next = n_next;
            r->scanned = (int64_t)tot[P_SCAN];
            r->cand = (int64_t)tot[P_CAND];
            r->t_start = c->t_step_start;
            r->t_end = now;
            r->t_prev_end = c->t_prev_end;
        }
        c->t_prev_end = now;
        c->n_prev = c->n_front;
        c->n_front = n_next;
        c->m_front = (int64_t)tot[P_OUT];
        c->m_unvisited -= (int64_t)tot[P_IN];
        c->n_unvisited -= n_next;
        c->prev_dir = c->dir;
        c->level = L + 1;
        for (int b = 0; b < kHistBins; b++) c->front_hist[b] = (uint32_t)tot[P_HIST + b];
        if (n_next == 0) c->done = 1;
        else compute_features(c, s_feat);
        s_done = c->done;
    }
    __syncthreads();
    if (s_done) return;
    const int nd = choose_direction(a, c, s_feat, pr);
    if (tid == 0) {
        c->dir = nd;
        const int32_t L1 = c->level;
        if (L1 < a.max_log)
            for (int i = 0; i < kFeat; i++) a.log[L1].feat[i] = s_feat[i];
    }
}

__device__ void init_ctrl(const StepArgs& a, int32_t nsrc, unsigned long long m_out, unsigned long long m_in,
                          const uint32_t* hist, int64_t nnz_scan, int32_t forced, int32_t mode, float alpha,
                          float beta) {
    BfsCtrl* c = a.ctrl;
    c->level = 0;
    c->dir = forced >= 0 ? forced : 0;
    c->done = nsrc == 0 ? 1 : 0;
    c->n_front = nsrc;
    c->n_prev = 0;
    c->prev_dir = 0;
    c->m_front = (int64_t)m_out;
    c->m_unvisited = nnz_scan - (int64_t)m_in;
    c->n_unvisited = (int64_t)a.nv - nsrc;
    for (int b = 0; b < kHistBins; b++) c->front_hist[b] = hist[b];
    c->t_prev_end = wall_clock64();
    c->t_step_start = c->t_prev_end;
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

// Clears the search state.  With src >= 0 it also seeds that single source
// and fills the control block (no separate seed launch).
__global__ void bfs_init(StepArgs a, int32_t src, int64_t nnz_scan, int32_t forced, int32_t mode, float alpha,
                         float beta) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.nv; i += stride) {
        a.dist[i] = i == src ? 0 : -1;
        if (a.parent) a.parent[i] = i == src ? (uint32_t)src : 0xffffffffu;
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.nwords; i += stride) {
        a.bm0[i] = (src >= 0 && i == (src >> 5)) ? (1u << (src & 31)) : 0u;
        a.bm1[i] = 0; a.bm2[i] = 0; a.vis[i] = 0;
    }
    if (src >= 0 && blockIdx.x == 0 && threadIdx.x == 0) {
        const int32_t dout = a.td_rp[src + 1] - a.td_rp[src];
        const int32_t din = a.bu_rp[src + 1] - a.bu_rp[src];
        uint32_t hist[kHistBins] = {0, 0, 0, 0, 0, 0, 0, 0};
        hist[hist_bin(dout)] = 1;
        init_ctrl(a, 1, (unsigned long long)dout, (unsigned long long)din, hist, nnz_scan, forced, mode, alpha, beta);
    }
}

// Marks a list of distinct sources and fills the control block.  One block.
__global__ __launch_bounds__(kBlock) void bfs_seed(StepArgs a, const int32_t* srcs, int32_t nsrc,
                                                   int64_t nnz_scan, int32_t forced, int32_t mode, float alpha,
                                                   float beta) {
    __shared__ unsigned long long s_out, s_in;
    __shared__ uint32_t s_hist[kHistBins];
    const int tid = threadIdx.x;
    if (tid == 0) { s_out = 0; s_in = 0; }
    if (tid < kHistBins) s_hist[tid] = 0;
    __syncthreads();
    unsigned long long mo = 0, mi = 0;
    for (int i = tid; i < nsrc; i += kBlock) {
        const int32_t s = srcs[i];
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
    if (tid == 0) init_ctrl(a, nsrc, s_out, s_in, s_hist, nnz_scan, forced, mode, alpha, beta);
}

// Undoes one step: every vertex of F_{L+1} (`Fn`) returns to unvisited and
// Fn is cleared.  Used by the direction probe to run the other direction on
// the identical state.
__global__ void bfs_undo(StepArgs a, uint32_t* Fn) {
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < a.nwords;
         w += (int64_t)gridDim.x * blockDim.x) {
        uint32_t bits = Fn[w];
        while (bits) {
            const int b = __ffs(bits) - 1;
            bits &= bits - 1;
            const int64_t v = w * 32 + b;
            a.dist[v] = -1;
            if (a.parent) a.parent[v] = 0xffffffffu;
        }
        Fn[w] = 0;
    }
}

}  // namespace

BfsWork::~BfsWork() {}

void bfs_free(BfsWork* w) {
    if (!w) return;
    dev_free(w->dist); dev_free(w->parent); dev_free(w->srcbuf); dev_free(w->part);
    dev_free(w->bm[0]); dev_free(w->bm[1]); dev_free(w->bm[2]); dev_free(w->vis);
    dev_free(w->ctrl); dev_free(w->log); dev_free(w->policy);
    if (w->h_ctrl) (void)hipHostFree(w->h_ctrl);
    if (w->ev0) (void)hipEventDestroy(w->ev0);
    if (w->ev1) (void)hipEventDestroy(w->ev1);
    for (auto e : w->step_ev) (void)hipEventDestroy(e);
    for (auto& x : w->gexec)
        if (x) (void)hipGraphExecDestroy(x);
    delete w;
}

void bfs_prepare(lgmi_graph* g) {
    if (g->bfs) return;
    BfsWork* w = new BfsWork();
    try {
        w->nv = g->nv;
        w->nwords = 2 * ((g->nv + 63) / 64);
        w->grid = std::max(1, g->cu_count) * 8;
        w->dist = dev_alloc_n<int32_t>(g->nv);
        w->parent = dev_alloc_n<int32_t>(g->nv);
        w->srcbuf = dev_alloc_n<int32_t>((size_t)g->nv + 1);
        w->part = dev_alloc_n<unsigned long long>((size_t)w->grid * kPart);
        for (int k = 0; k < 3; k++) w->bm[k] = dev_alloc_n<uint32_t>(std::max(w->nwords, 2));
        w->vis = dev_alloc_n<uint32_t>(std::max(w->nwords, 2));
        w->ctrl = dev_alloc_n<BfsCtrl>(1);
        w->max_log = 16384;   // covers a 4096 x 4096 grid (up to 8190 levels)
        w->log = dev_alloc_n<LevelRec>(w->max_log);
        w->policy = dev_alloc_n<PolicyWeights>(1);
        LGMI_HIP(hipHostMalloc((void**)&w->h_ctrl, sizeof(BfsCtrl), hipHostMallocDefault));
        LGMI_HIP(hipEventCreate(&w->ev0));
        LGMI_HIP(hipEventCreate(&w->ev1));
        LGMI_HIP(hipMemsetAsync(w->ctrl, 0, sizeof(BfsCtrl), g->stream));
        LGMI_HIP(hipMemsetAsync(w->log, 0, sizeof(LevelRec) * w->max_log, g->stream));
        LGMI_HIP(hipMemsetAsync(w->part, 0, sizeof(unsigned long long) * (size_t)w->grid * kPart, g->stream));
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
    a.td_item_v = td.item_v; a.td_item_beg = td.item_beg; a.td_items = td.n_items;
    a.bu_rp = bu.row_ptr; a.bu_col = bu.col; a.bu_nz = bu.nz_bm;
    a.nv = g->nv;
    a.nwords = w->nwords;
    a.dist = w->dist;
    a.parent = parents ? reinterpret_cast<uint32_t*>(w->parent) : nullptr;
    a.bm0 = w->bm[0]; a.bm1 = w->bm[1]; a.bm2 = w->bm[2];
    a.vis = w->vis;
    a.ctrl = w->ctrl;
    a.part = w->part;
    a.log = w->log;
    a.max_log = w->max_log;
    a.policy = w->policy;
    return a;
}

void policy_upload(lgmi_graph* g) {
    std::lock_guard<std::mutex> lk(policy_mutex());
    PolicyHost& p = policy_host();
    if (p.have_mlp)
        LGMI_HIP(hipMemcpyAsync(g->bfs->policy, &p.w, sizeof(PolicyWeights), hipMemcpyHostToDevice, g->stream));
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
    if (mode != LGMI_POLICY_LEARNED) a.policy = nullptr;
    const int64_t nnz_scan = g->nnz;  // the reverse CSR has as many entries as the forward one
    hipStream_t s = g->stream;
    const int grid = w->grid;
    if (single_src < 0 && nsrc > 0)
        LGMI_HIP(hipMemcpyAsync(w->srcbuf, h_srcs, sizeof(int32_t) * (size_t)nsrc, hipMemcpyHostToDevice, s));
    LGMI_HIP(hipEventRecord(w->ev0, s));
    const int init_grid = std::max(1, std::min<int>(grid, (g->nv + 255) / 256));
    bfs_init<<<init_grid, 256, 0, s>>>(a, single_src, nnz_scan, forced_dir, mode, alpha, beta);
    if (single_src < 0)
        bfs_seed<<<1, kBlock, 0, s>>>(a, w->srcbuf, nsrc, nnz_scan, forced_dir, mode, alpha, beta);
    LGMI_HIP(hipGetLastError());
    // Enqueue steps in batches; a step after the last level exits at once.
    int batch = std::max(4, std::min(w->last_levels + 1, 64));
    int launched = 0;
    auto event_at = [&](size_t i) {
        while (w->step_ev.size() <= i) {
            hipEvent_t e;
            LGMI_HIP(hipEventCreate(&e));
            w->step_ev.push_back(e);
        }
        LGMI_HIP(hipEventRecord(w->step_ev[i], s));
    };
    static const bool graphs_on = [] {
        const char* e = std::getenv("LGMI_BFS_GRAPHS");
        return !(e && e[0] == '0');
    }();
    for (;;) {
        if (graphs_on && !w->time_steps && launched >= kGraphLevels) {
            // deep search: replay captured step/finalise sequences
            const int key = (want_parents ? 1 : 0) | (dir_in ? 2 : 0) | (a.policy ? 4 : 0);
            if (!w->gexec[key]) {
                hipGraph_t gr = nullptr;
                LGMI_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                for (int i = 0; i < kGraphLevels; i++) {
                    if (want_parents) bfs_step<true><<<grid, kBlock, 0, s>>>(a);
                    else bfs_step<false><<<grid, kBlock, 0, s>>>(a);
                    bfs_finalize<<<1, kFinBlock, 0, s>>>(a, grid);
                }
                LGMI_HIP(hipStreamEndCapture(s, &gr));
                hipGraphExec_t ex = nullptr;
                const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                LGMI_HIP(e);
                w->gexec[key] = ex;
            }
            const int ng = std::max(1, batch / kGraphLevels);
            for (int i = 0; i < ng; i++) LGMI_HIP(hipGraphLaunch(w->gexec[key], s));
            launched += ng * kGraphLevels;
        } else {
            for (int i = 0; i < batch; i++) {
                const size_t li = (size_t)(launched + i);
                if (w->time_steps) event_at(2 * li);
                if (want_parents) bfs_step<true><<<grid, kBlock, 0, s>>>(a);
                else bfs_step<false><<<grid, kBlock, 0, s>>>(a);
                if (w->time_steps) event_at(2 * li + 1);
                bfs_finalize<<<1, kFinBlock, 0, s>>>(a, grid);
            }
            launched += batch;
        }
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
            LGMI_HIP(hipEventElapsedTime(&ms, w->step_ev[2 * (size_t)i], w->step_ev[2 * (size_t)i + 1]));
            w->last_launch_ms[(size_t)i] = ms;
            w->last_steps_event_ms += ms;
        }
    }
    if (stats) {
        float ms = 0.f;
        LGMI_HIP(hipEventElapsedTime(&ms, w->ev0, w->ev1));
        std::memset(stats, 0, sizeof(*stats));
        stats->levels = w->h_ctrl->level;
        stats->device_ms = ms;
        stats->launches = launched;
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
            for (int q = 0; q < kFeat; q++) o.feat[q] = r.feat[q];
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

// ----------------------------------------------------------------------
// direction probe: both directions timed on the identical level state
// ----------------------------------------------------------------------
extern "C" int lgmi_bfs_probe(lgmi_graph_t g, int32_t src, int dir_in, const int32_t* levels, int32_t nlevels,
                              lgmi_probe_rec* out) {
    using namespace lgmi;
    try {
        if (!g || nlevels < 0 || (nlevels > 0 && (!levels || !out))) throw Error(LGMI_E_ARG, "bad arguments");
        std::lock_guard<std::mutex> lk(g->mu);
        if (src < 0 || src >= g->nv) throw Error(LGMI_E_BOUNDS, "source vertex out of range");
        for (int32_t i = 1; i < nlevels; i++)
            if (levels[i] <= levels[i - 1]) throw Error(LGMI_E_ARG, "levels must be strictly increasing");
        bfs_prepare(g);
        BfsWork* w = g->bfs;
        StepArgs a = make_args(g, dir_in, true);
        a.policy = nullptr;
        hipStream_t s = g->stream;
        const int grid = w->grid;
        const int init_grid = std::max(1, std::min<int>(grid, (g->nv + 255) / 256));
        bfs_init<<<init_grid, 256, 0, s>>>(a, src, g->nnz, -1, LGMI_POLICY_HEURISTIC, 15.0f, 18.0f);
        LGMI_HIP(hipGetLastError());
        hipEvent_t ev[4];
        for (auto& e : ev) LGMI_HIP(hipEventCreate(&e));
        auto sync_ctrl = [&]() {
            LGMI_HIP(hipMemcpyAsync(w->h_ctrl, w->ctrl, sizeof(BfsCtrl), hipMemcpyDeviceToHost, s));
            LGMI_HIP(hipStreamSynchronize(s));
        };
        auto run_forced = [&](int dir, hipEvent_t e0, hipEvent_t e1) {
            BfsCtrl c = *w->h_ctrl;
            c.dir = dir;
            LGMI_HIP(hipMemcpyAsync(w->ctrl, &c, sizeof(BfsCtrl), hipMemcpyHostToDevice, s));
            LGMI_HIP(hipEventRecord(e0, s));
            bfs_step<true><<<grid, kBlock, 0, s>>>(a);
            LGMI_HIP(hipEventRecord(e1, s));
            bfs_finalize<<<1, kFinBlock, 0, s>>>(a, grid);
            LGMI_HIP(hipGetLastError());
            LGMI_HIP(hipStreamSynchronize(s));
        };
        int nout = 0;
        try {
            sync_ctrl();
            for (int32_t li = 0; li < nlevels && !w->h_ctrl->done; li++) {
                const int32_t target = levels[li];
                while (!w->h_ctrl->done && w->h_ctrl->level < target) {
                    const int n = std::min(target - w->h_ctrl->level, 256);
                    for (int i = 0; i < n; i++) {
                        bfs_step<true><<<grid, kBlock, 0, s>>>(a);
                        bfs_finalize<<<1, kFinBlock, 0, s>>>(a, grid);
                    }
                    LGMI_HIP(hipGetLastError());
                    sync_ctrl();
                }
                if (w->h_ctrl->done || w->h_ctrl->level != target) break;
                const BfsCtrl saved = *w->h_ctrl;
                const int32_t L = saved.level;
                uint32_t* Fn = w->bm[(L + 1) % 3];
                // Each direction runs kReps times on the restored state, in
                // alternating order; the minimum is kept, so one slow launch
                // (a cold start after the host synchronised) does not decide
                // the label.
                constexpr int kReps = 3;
                float best[2] = {1e30f, 1e30f};
                const int first = ((L + src) & 1) ? 1 : 0;
                for (int rep = 0; rep < kReps; rep++) {
                    for (int q = 0; q < 2; q++) {
                        const int dir = (first + q + rep) & 1;
                        *w->h_ctrl = saved;
                        run_forced(dir, ev[0], ev[1]);
                        float t = 0.f;
                        LGMI_HIP(hipEventElapsedTime(&t, ev[0], ev[1]));
                        best[dir] = std::min(best[dir], t);
                        if (!(rep == kReps - 1 && q == 1)) {
                            bfs_undo<<<grid, 256, 0, s>>>(a, Fn);
                            LGMI_HIP(hipGetLastError());
                        }
                    }
                }
                const float t_first = best[first], t_second = best[1 - first];
                LevelRec rec;
                LGMI_HIP(hipMemcpy(&rec, w->log + std::min(L, w->max_log - 1), sizeof(LevelRec), hipMemcpyDeviceToHost));
                lgmi_probe_rec& o = out[nout++];
                std::memset(&o, 0, sizeof(o));
                for (int q = 0; q < kFeat; q++) o.feat[q] = rec.feat[q];
                o.td_us = 1000.f * (first == 0 ? t_first : t_second);
                o.bu_us = 1000.f * (first == 0 ? t_second : t_first);
                o.level = L;
                o.n_front = saved.n_front;
                sync_ctrl();
            }
        } catch (...) {
            for (auto& e : ev) (void)hipEventDestroy(e);
            throw;
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        w->last_levels = 0;
        w->last_parents = 0;
        return nout;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return LGMI_E_HIP;
    }
}
