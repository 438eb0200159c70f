// bfs.hip -- direction-optimising single/multi-source BFS for gfx950.
//
// Replaces the level loop of gdistances! (src/traversals/bfs.jl:95-124),
// _bfs_parents (bfs.jl:38-66) and the threaded Parallel.gdistances! /
// Parallel.bfs_tree! (src/Parallel/traversals/gdistances.jl:36-116,
// src/Parallel/traversals/bfs.jl:49-92).
//
// State in HBM: dist[nv], parent[nv] (Int32), three frontier bitmaps F_k
// rotating by level (read F_k, write F_{k+1}, clear F_{k+2}) and the visited
// bitmap vis = F_0 | ... | F_{k-1} (F_k is folded in during step k).  No
// vertex queues: every step sweeps bitmap tiles of 512 vertices, one
// wavefront per tile and 8 bits per lane; the set bits of interest (frontier
// vertices for top-down, unvisited vertices with a non-empty row for
// bottom-up) are compacted into batches of 64 vertices with a prefix
// popcount over the lanes and a rank-select inside the byte, so every lane
// of a batch works.
//
// One launch of bfs_level per level k.  Its prologue replaces a separate
// reduction kernel and any cross-work-group atomics: every work-group loads
// the per-block counters the launch for level k-1 wrote (one round of
// independent loads), sums them in LDS, derives the level state S_k from
// S_{k-1} and picks the direction of step k (Beamer alpha/beta, the learned
// MLP, or a forced direction).  All work-groups compute the same S_k;
// work-group 0 also stores it, with the level log.  Level parity selects
// the counter and state buffers, so no launch overwrites what a work-group
// of the same launch still reads.  The launch that finds F_k empty stores
// k as done_level; later launches of the batch exit at once.
//
// Step k -> k+1:
//  * top-down: rows of 32..256 entries are expanded by the whole wavefront,
//    shorter rows by a scan-based gather that packs several rows into the
//    64 lanes, rows longer than 256 entries through precomputed 256-entry
//    work items dealt round-robin over all wavefronts; each lane handles 4
//    entries per round with all their loads in flight together.  A vertex's
//    parent is the minimum over its frontier neighbours (atomicMin after a
//    plain-load filter); the atomicMin that finds the parent unset claims
//    the vertex.  Device-scope atomics run at about 26 G operations/s on this
//    GPU, several times below plain loads and stores, so the step issues at
//    most one per fresh edge plus one per new vertex.
//  * bottom-up: every candidate scans its sorted reverse row for the first
//    neighbour in F_k, which is its minimum-id parent.  Stage one tests the
//    first entry of every candidate, read from a contiguous first-entry
//    array (it is the parent for most vertices on the large levels); the
//    rest wait in a per-wavefront LDS queue and scan
//    four entries per round trip for up to four round trips, after which a
//    whole wavefront scans each row still open, 64 entries per round trip.
//    The tile's F_{k+1} words are gathered in LDS and written whole.
// Both directions give identical distances and parents: the reference's
// sorted frontier makes parent[v] the smallest u with dist[u] = dist[v]-1
// and an edge u -> v, which is what either direction computes.
#include "lgmi_internal.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <thread>
#include <vector>

namespace lgmi {

namespace {

constexpr int kBlock = 1024;
constexpr int kWaves = kBlock / 64;
constexpr int kLaneBits = 8;                    // frontier bits per lane in a tile
constexpr int kTileVerts = 64 * kLaneBits;      // vertices per wavefront tile
constexpr int kTileWords = kTileVerts / 32;     // bitmap words per tile
constexpr int kUnroll = 2;          // top-down: adjacency entries per lane per round
constexpr int kGraphLevels = 18;    // level launches per captured graph (a multiple of 6: level mod 6 is static)
constexpr int kPartRows = kBlock / kPart;
constexpr int32_t kRangeFront = 2048;    // range mode: frontier vertices at most (LDS layout)
constexpr int kEmit = 8;                 // range mode: 1024-vertex slices emitted per round
constexpr int32_t kQueueFront = 4096;    // queue mode: frontier vertices at most (LDS layout)
constexpr int32_t kSoloFront = 4096;     // solo mode: frontier vertices at most (LDS layout)
constexpr int kSoloPer = kSoloFront / kBlock;   // solo mode: frontier entries per thread
constexpr int kSoloGroup = 8;            // solo mode: edges per lane with their loads in flight together

struct StepArgs {
    const int32_t* td_rp;     // traversal adjacency (top-down expands it)
    const int32_t* td_col;
    const int32_t* td_item_v; // heavy-row items of the traversal adjacency
    const int32_t* td_item_beg;
    const int32_t* td_item_end;
    int32_t td_items;
    const int32_t* td_heavy_id; // range table of the traversal adjacency (range_size 0: none)
    const int32_t* td_rtab;
    int32_t range_size;
    const int32_t* bu_rp;     // reverse adjacency (bottom-up scans it)
    const int32_t* bu_col;
    const uint32_t* bu_nz;    // bit v: row v of the reverse adjacency is non-empty
    const int32_t* bu_first;  // [nv]: first entry of reverse row v (-1: empty)
    int32_t nv;
    int32_t nwords;           // 2 * ceil(nv / 64)
    int32_t* dist;
    uint32_t* parent;         // 0xffffffff = unset (nullptr: no parents)
    uint32_t* bm0;
    uint32_t* bm1;
    uint32_t* bm2;
    uint32_t* vis;
    int32_t* qbuf;            // [2][grid][kQCap] new vertices per work-group, by level parity
    BfsShared* sh;
    unsigned long long* part; // [2][grid][kPart] by launch parity
    int32_t grid;
    LevelRec* log;
    int32_t max_log;
    const PolicyWeights* policy;
    int32_t* done_host;       // host-mapped copy of done_level (nullptr: none)
    // launch thresholds (lgmi_tuning)
    int32_t wave_row;         // top-down: rows this long or longer are expanded by a wavefront
    int32_t lane_rounds;      // bottom-up: per-lane rounds of four entries before whole-wavefront scans
    int32_t range_front;      // range mode: frontier vertices at most (<= kRangeFront)
    int32_t queue_front;      // queue mode: frontier vertices at most (<= kQueueFront)
    int64_t range_light;      // range mode: light-row frontier edges at most
    int64_t queue_min_edges;  // queue mode: frontier edges at least
    int64_t queue_max_edges;  // queue mode: frontier edges at most
    int64_t solo_edges;       // solo mode: frontier edges at most
    int32_t solo_front;       // solo mode: frontier vertices at most (<= kSoloFront; 0: off)
};

// done_level (the launch from which on launches exit at once), the level
// count of the ended search, and their host-mapped copies (which the host
// reads after the launches complete, instead of copying the state back).
__device__ inline void set_done(const StepArgs& a, int32_t v, int32_t levels) {
    a.sh->done_level = v;
    a.sh->done_levels = levels;
    if (a.done_host)   // one 64-bit store: the host never sees one half without the other
        __hip_atomic_store(reinterpret_cast<unsigned long long*>(a.done_host),
                           ((unsigned long long)(uint32_t)levels << 32) | (uint32_t)v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ inline uint32_t* bm_at(const StepArgs& a, int k) {
    return k == 0 ? a.bm0 : (k == 1 ? a.bm1 : a.bm2);
}

__device__ inline int hist_bin(int32_t deg) {
    // powers of four: 0, 1-3, 4-15, 16-63, 64-255, 256-1023, >= 1024
    if (deg <= 0) return 0;
    const int b = (31 - __clz(deg)) / 2 + 1;
    return b > 6 ? 6 : b;
}

__device__ inline float flog2p(float x) { return __log2f(1.0f + (x > 0.f ? x : 0.f)); }

// Exclusive prefix sum over the 64 lanes; total in `total`.
__device__ inline int32_t wave_excl_scan(int32_t x, int lane, int32_t& total) {
    int32_t incl = x;
    #pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int32_t y = __shfl_up(incl, d);
        if (lane >= d) incl += y;
    }
    total = __shfl(incl, 63);
    return incl - x;
}

// Sum over the 64 lanes with DPP row shifts and row broadcasts (VALU
// cross-lane operations; no LDS traffic).  Lane 63 collects the total,
// which readlane returns to every lane.
__device__ inline uint32_t wave_sum(uint32_t x) {
    // inclusive sums within rows of 16 lanes: row_shr:1, 2, 4, 8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);
    // row_bcast:15 adds a row's total into the next row (rows 1 and 3),
    // row_bcast:31 adds lane 31's total into rows 2 and 3
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);
    return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

// wave_sum for FP32 (the same DPP sequence, FP32 adds).
__device__ inline float wave_sum_f(float x) {
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x111, 0xf, 0xf, false));
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x112, 0xf, 0xf, false));
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x114, 0xf, 0xf, false));
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x118, 0xf, 0xf, false));
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x142, 0xa, 0xf, false));
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x143, 0xc, 0xf, false));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
}

// Largest lane lo with excl[lo] <= g (excl non-decreasing over lanes, excl[0] = 0).
__device__ inline int owner_lane(int32_t excl, int32_t g) {
    int lo = 0;
    #pragma unroll
    for (int step = 32; step >= 1; step >>= 1) {
        const int probe = lo + step;
        const int32_t pv = __shfl(excl, probe & 63);
        if (probe < 64 && pv <= g) lo = probe;
    }
    return lo;
}

// The set bits of a tile as a list: lane l holds the byte cw of vertices
// v0 + 8l .. v0 + 8l + 7, with exclusive popcount prefix excl, and writes
// those vertices in ascending order to entries excl .. of the wavefront's
// LDS list (kTileVerts entries); the lanes then read it 64 entries (one
// batch) at a time.  One LDS write per bit position and one read per batch
// take the place of a per-batch rank search over the lanes (a chain of
// dependent cross-lane reads).  Must be called by all 64 lanes.
__device__ __attribute__((always_inline)) inline void tile_list(int32_t* s_list, uint32_t cw, int32_t excl,
                                                                int32_t v0, int lane) {
    int32_t p = excl;
    const int32_t vb = v0 + lane * kLaneBits;
    #pragma unroll
    for (int j = 0; j < kLaneBits; j++) {
        if ((cw >> j) & 1u) {
            s_list[p] = vb + j;
            p++;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Features of the level state (LGMI_POLICY_FEATURES), in FP32 (the
// hardware log2).
__device__ __attribute__((always_inline)) inline void compute_features(const BfsCtrl* c, float* f) {
    const float nf = (float)c->n_front, mf = (float)c->m_front, mu = (float)c->m_unvisited;
    const float nu = (float)c->n_unvisited;
    const float lmf = flog2p(mf), lmu = flog2p(mu), lnf = flog2p(nf);
    f[0] = lnf;
    f[1] = lmf;
    f[2] = lmu;
    f[3] = flog2p(nu);
    f[4] = flog2p((float)c->level);
    f[5] = nf / (float)(c->nv > 0 ? c->nv : 1);
    f[6] = lmf - lmu;
    f[7] = lnf - flog2p((float)c->n_prev);
    f[8] = (float)c->prev_dir;
    uint32_t tot = 0;
    for (int b = 0; b < 7; b++) tot += c->front_hist[b];
    const float inv = tot ? 1.f / (float)tot : 0.f;
    for (int b = 0; b < 7; b++) f[9 + b] = (float)c->front_hist[b] * inv;
}

struct Counters {
    unsigned long long n = 0, out = 0, in = 0, scan = 0, cand = 0, light = 0;
    uint32_t hist[kHistBins] = {};   // new vertices per degree bin (wavefront-uniform)
};

// A new vertex v of F_{k+1}: its degrees feed the next level's state.
// din_known >= 0 is v's reverse-adjacency degree when the caller has it.
// Returns the histogram bin of v's traversal-direction degree (push_new
// records it).
__device__ inline int account_new(const StepArgs& a, int32_t v, int32_t din_known, Counters& k) {
    int32_t dout, din;
    if (a.td_rp == a.bu_rp && din_known >= 0) {
        dout = din = din_known;
    } else {
        dout = a.td_rp[v + 1] - a.td_rp[v];
        din = (a.td_rp == a.bu_rp) ? dout : (din_known >= 0 ? din_known : a.bu_rp[v + 1] - a.bu_rp[v]);
    }
    k.n++;
    k.out += (unsigned long long)dout;
    k.in += (unsigned long long)din;
    if (dout <= kHeavyRow) k.light += (unsigned long long)dout;
    return hist_bin(dout);
}

// Records the lanes' new vertices of N rounds (flag[q] set): their degree
// bins go to the wavefront's histogram counters (three ballots give every
// lane's bin; no LDS traffic), and their ids to the work-group's queue, the
// first kQCap of them, with one LDS atomic on the queue count *s_qn per call.
// Every lane of the wavefront calls it.
template <int N>
__device__ __attribute__((always_inline)) inline void push_new_n(const bool* flag, const int32_t* v, const int* bin,
                                                                 Counters& k, uint32_t* s_qn, int32_t* qout, int lane) {
    uint64_t m[N];
    uint32_t tot = 0;
    #pragma unroll
    for (int q = 0; q < N; q++) {
        m[q] = __ballot(flag[q]);
        tot += (uint32_t)__popcll(m[q]);
    }
    if (!tot) return;
    #pragma unroll
    for (int q = 0; q < N; q++) {
        const uint64_t b0 = __ballot(bin[q] & 1), b1 = __ballot(bin[q] & 2), b2 = __ballot(bin[q] & 4);
        #pragma unroll
        for (int b = 0; b < 7; b++)
            k.hist[b] += (uint32_t)__popcll(m[q] & ((b & 1) ? b0 : ~b0) & ((b & 2) ? b1 : ~b1) & ((b & 4) ? b2 : ~b2));
    }
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(s_qn, tot);
    base = __shfl(base, 0);
    const uint64_t lt = (1ull << lane) - 1ull;
    #pragma unroll
    for (int q = 0; q < N; q++) {
        const uint32_t pos = base + (uint32_t)__popcll(m[q] & lt);
        if (flag[q] && pos < (uint32_t)kQCap) qout[pos] = v[q];
        base += (uint32_t)__popcll(m[q]);
    }
}

__device__ __attribute__((always_inline)) inline void push_new(bool flag, int32_t v, int bin, Counters& k,
                                                               uint32_t* s_qn, int32_t* qout, int lane) {
    push_new_n<1>(&flag, &v, &bin, k, s_qn, qout, lane);
}

// Work-group barrier that orders LDS accesses only: the global stores a
// wavefront has in flight are not waited for (they complete before the
// kernel ends, and nothing in the work-group reads them back).
__device__ inline void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Top-down visit of N edges u[q] -> v[q] (v[q] < 0: none).  All loads of the
// N edges are issued before any result is used.  With parents, the atomicMin
// that finds a vertex's parent unset is its claim (one atomic per fresh edge
// that passes the plain-load filter, plus one atomicOr per new vertex);
// without parents the claim is an atomicOr on F_{k+1}.
template <bool PARENTS, int N>
__device__ inline void visit_n(const StepArgs& a, const int32_t* u, const int32_t* v, int32_t L,
                               const uint32_t* FL, uint32_t* Fn, Counters& k, uint32_t* s_qn, int32_t* qout,
                               int lane) {
    uint32_t seen[N], pre[N];
    #pragma unroll
    for (int q = 0; q < N; q++) {
        seen[q] = 0xffffffffu;
        pre[q] = 0;
        if (v[q] >= 0) {
            const int32_t w = v[q] >> 5;
            seen[q] = a.vis[w] | FL[w];
            pre[q] = PARENTS ? a.parent[v[q]] : Fn[w];
        }
    }
    bool claim[N];
    #pragma unroll
    for (int q = 0; q < N; q++) {
        const bool fresh = v[q] >= 0 && !((seen[q] >> (v[q] & 31)) & 1u);   // distance > L
        // parent only decreases, so a stale value <= u already rules u out
        claim[q] = fresh && (PARENTS ? pre[q] > (uint32_t)u[q] : !((pre[q] >> (v[q] & 31)) & 1u));
    }
    uint32_t old[N];
    #pragma unroll
    for (int q = 0; q < N; q++) {
        old[q] = 0;
        if (claim[q]) {
            if (PARENTS) old[q] = atomicMin(&a.parent[v[q]], (uint32_t)u[q]);
            else old[q] = atomicOr(&Fn[v[q] >> 5], 1u << (v[q] & 31));
        }
    }
    bool won[N];
    int bin[N];
    #pragma unroll
    for (int q = 0; q < N; q++) {
        won[q] = claim[q] && (PARENTS ? old[q] == 0xffffffffu : !((old[q] >> (v[q] & 31)) & 1u));
        bin[q] = 0;
        if (won[q]) {
            if (PARENTS) atomicOr(&Fn[v[q] >> 5], 1u << (v[q] & 31));
            __builtin_nontemporal_store(L + 1, &a.dist[v[q]]);
            bin[q] = account_new(a, v[q], -1, k);
        }
    }
    push_new_n<N>(won, v, bin, k, s_qn, qout, lane);
}

// Top-down expansion of row entries [b, e) of frontier vertex u by a whole
// wavefront, kUnroll entries per lane per round.
template <bool PARENTS>
__device__ inline void expand_row(const StepArgs& a, int32_t u, int32_t b, int32_t e, int32_t L,
                                  const uint32_t* FL, uint32_t* Fn, Counters& k, uint32_t* s_qn, int32_t* qout,
                                  int lane) {
    for (int32_t x0 = b; x0 < e; x0 += 64 * kUnroll) {
        int32_t uu[kUnroll], vv[kUnroll];
        #pragma unroll
        for (int q = 0; q < kUnroll; q++) {
            const int32_t x = x0 + q * 64 + lane;
            uu[q] = u;
            vv[q] = x < e ? a.td_col[x] : -1;
        }
        visit_n<PARENTS, kUnroll>(a, uu, vv, L, FL, Fn, k, s_qn, qout, lane);
    }
}

// Top-down step L -> L+1 of one wavefront (gw of nw).
template <bool PARENTS>
__device__ __attribute__((always_inline)) inline void td_step(const StepArgs& a, const BfsCtrl& c, int32_t L, const uint32_t* FL, uint32_t* Fn,
                        uint32_t* Fc, Counters& k, uint32_t* s_qn, int32_t* qout, int32_t* s_list, int gw, int nw,
                        int lane, uint32_t fl_first) {
    const int32_t ntiles = (a.nwords + kTileWords - 1) / kTileWords;
    const int sh8 = (lane & 3) * kLaneBits;
    // Heavy rows: 256-entry items; lane j of wavefront gw tests item
    // gw + (base + j) * nw, so consecutive items of one hub go to
    // consecutive wavefronts, and expands those of frontier vertices.
    // Skipped when no frontier vertex has 256 or more entries.
    if (c.front_hist[5] + c.front_hist[6] > 0) {
        for (int64_t base = 0; (int64_t)gw + base * nw < a.td_items; base += 64) {
            const int64_t it = (int64_t)gw + (base + lane) * nw;
            int32_t iu = -1, ib = 0, ie = 0;
            if (it < a.td_items) {
                iu = a.td_item_v[it];
                ib = a.td_item_beg[it];
                ie = a.td_item_end[it];
            }
            if (iu >= 0 && !((FL[iu >> 5] >> (iu & 31)) & 1u)) iu = -1;
            uint64_t m = __ballot(iu >= 0);
            while (m) {
                const int j = __ffsll((unsigned long long)m) - 1;
                m &= m - 1;
                const int32_t u = __shfl(iu, j), b = __shfl(ib, j), e = __shfl(ie, j);
                if (lane == 0) k.scan += (unsigned long long)(e - b);
                expand_row<PARENTS>(a, u, b, e, L, FL, Fn, k, s_qn, qout, lane);
            }
        }
    }
    // frontier tiles; the next tile's frontier word is loaded before the
    // current tile is expanded (the first tile's came with the prologue)
    int32_t t = gw;
    uint32_t fl_next = t < ntiles ? fl_first : 0u;
    while (t < ntiles) {
        const int32_t word0 = t * kTileWords;
        const int32_t w = word0 + (lane >> 2);
        const uint32_t fl = fl_next;
        t += nw;
        {
            const int32_t wn = t * kTileWords + (lane >> 2);
            fl_next = (t < ntiles && wn < a.nwords) ? FL[wn] : 0u;
        }
        uint32_t cw = 0;
        if (w < a.nwords) {
            cw = (fl >> sh8) & 0xffu;
            if ((lane & 3) == 0) {
                if (fl) a.vis[w] |= fl;   // fold F_k into vis
                Fc[w] = 0;                // clear F_{k+2}
            }
        }
        int32_t total;
        const int32_t excl = wave_excl_scan(__popc(cw), lane, total);
        if (total > 0) tile_list(s_list, cw, excl, word0 * 32, lane);
        for (int32_t base = 0; base < total; base += 64) {
            const int32_t u = base + lane < total ? s_list[base + lane] : -1;
            int32_t beg = 0, deg = 0;
            if (u >= 0) {
                beg = a.td_rp[u];
                deg = a.td_rp[u + 1] - beg;
                if (deg > kHeavyRow) deg = 0;     // expanded through its items
            }
            k.scan += (unsigned long long)deg;
            // wavefront-wide rows
            uint64_t m = __ballot(deg >= a.wave_row);
            while (m) {
                const int j = __ffsll((unsigned long long)m) - 1;
                m &= m - 1;
                const int32_t wu = __shfl(u, j), wb = __shfl(beg, j), wd = __shfl(deg, j);
                expand_row<PARENTS>(a, wu, wb, wb + wd, L, FL, Fn, k, s_qn, qout, lane);
            }
            if (deg >= a.wave_row) deg = 0;
            // short rows: scan-based gather over the batch, kUnroll entries per lane
            int32_t rtot;
            const int32_t rex = wave_excl_scan(deg, lane, rtot);
            for (int32_t off = 0; off < rtot; off += 64 * kUnroll) {
                int32_t uu[kUnroll], vv[kUnroll];
                #pragma unroll
                for (int q = 0; q < kUnroll; q++) {
                    const int32_t e = off + q * 64 + lane;
                    const int o = owner_lane(rex, e);
                    const int32_t ob = __shfl(beg, o), oe = __shfl(rex, o), ou = __shfl(u, o);
                    uu[q] = ou;
                    vv[q] = e < rtot ? a.td_col[ob + (e - oe)] : -1;
                }
                visit_n<PARENTS, kUnroll>(a, uu, vv, L, FL, Fn, k, s_qn, qout, lane);
            }
        }
    }
}

// A candidate of the parent scan waiting in a wavefront's LDS queue.
struct Pending {
    int32_t v, b, e, b0;   // vertex, next entry, end of row, start of row
};

// Records a found vertex v (parent par, tile offset off, row [b0, e)).
template <bool PARENTS>
__device__ inline int bu_found(const StepArgs& a, int32_t v, int32_t par, int32_t off, int32_t b0, int32_t e,
                               int32_t L, Counters& k, uint32_t* s_fn) {
    __builtin_nontemporal_store(L + 1, &a.dist[v]);
    if (PARENTS) __builtin_nontemporal_store((uint32_t)par, &a.parent[v]);
    atomicOr(&s_fn[off >> 5], 1u << (off & 31));
    return account_new(a, v, e - b0, k);
}

// Second stage of the parent scan: up to 64 queued candidates, one per lane,
// scanning four entries per round trip from entry b onwards, for at most
// a.lane_rounds rounds; rows still open after that are scanned by the whole
// wavefront, 64 entries per round trip.
template <bool PARENTS>
__device__ __attribute__((always_inline)) inline void bu_drain(const StepArgs& a, const Pending* q, int n, int32_t v0, int32_t L, const uint32_t* FL,
                         Counters& k, uint32_t* s_qn, int32_t* qout, uint32_t* s_fn, int lane) {
    int32_t v = -1, b = 0, e = 0, b0 = 0;
    // field by field: a whole-struct copy can leave a private copy that the
    // compiler places in LDS at a per-thread address, which needs the
    // dispatch packet (block dimensions) in every wave of every launch
    if (lane < n) { v = q[lane].v; b = q[lane].b; e = q[lane].e; b0 = q[lane].b0; }
    int32_t par = -1;
    int32_t x = b;
    if (v >= 0) {
        for (int r = 0; r < a.lane_rounds && x < e; r++, x += 4) {
            k.scan += (unsigned long long)min(4, e - x);
            int32_t uu[4];
            uint32_t bw[4];
            #pragma unroll
            for (int t = 0; t < 4; t++) uu[t] = (x + t < e) ? a.bu_col[x + t] : -1;
            #pragma unroll
            for (int t = 0; t < 4; t++) bw[t] = uu[t] >= 0 ? FL[uu[t] >> 5] : 0u;
            #pragma unroll
            for (int t = 0; t < 4; t++)
                if (par < 0 && uu[t] >= 0 && ((bw[t] >> (uu[t] & 31)) & 1u)) par = uu[t];
            if (par >= 0) break;
        }
    }
    const bool open = v >= 0 && par < 0 && x < e;
    uint64_t mb = __ballot(open);
    while (mb) {
        const int j = __ffsll((unsigned long long)mb) - 1;
        mb &= mb - 1;
        const int32_t jb = __shfl(x, j), je = __shfl(e, j);
        int32_t found_u = -1;
        for (int32_t off = jb; off < je; off += 64) {
            const int32_t y = off + lane;
            bool hit = false;
            int32_t uu = 0;
            if (y < je) { uu = a.bu_col[y]; hit = (FL[uu >> 5] >> (uu & 31)) & 1u; }
            const uint64_t hm = __ballot(hit);
            if (lane == j) k.scan += (unsigned long long)min(64, je - off);
            if (hm) { found_u = __shfl(uu, __ffsll((unsigned long long)hm) - 1); break; }
        }
        if (lane == j) par = found_u;
    }
    int bin = 0;
    if (par >= 0) bin = bu_found<PARENTS>(a, v, par, v - v0, b0, e, L, k, s_fn);
    push_new(par >= 0, v, bin, k, s_qn, qout, lane);
}

// The candidates of one tile (byte cw of lane l: vertices v0 + 8l ..
// v0 + 8l + 7; excl its exclusive popcount prefix, total > 0 their number):
// stage one, the found vertices recorded, and the open ones drained.
template <bool PARENTS>
__device__ __attribute__((always_inline)) inline Counters bu_tile(const StepArgs& a, int32_t L, const uint32_t* FL, uint32_t* s_qn,
                                         int32_t* qout, uint32_t* s_fn, Pending* s_q, int32_t* s_list, uint32_t cw,
                                         int32_t excl, int32_t total, int32_t word0, int lane) {
    Counters k;
    const int32_t v0 = word0 * 32;
    int qn = 0;   // queued candidates (uniform over the wavefront)
    // Stage one for all batches of the tile (kTileVerts / 64 at most) at
    // once: every batch's row bounds and first entry are loaded, then
    // every first entry's frontier word, so the tile pays two memory
    // round trips rather than two per batch.  The first entry comes from
    // the contiguous first-entry array (a tile's candidates read one run
    // of it), not from a scattered line of the adjacency.
    constexpr int kTB = kTileVerts / 64;
    const int nb = (total + 63) >> 6;   // uniform
    int32_t qv[kTB], qb[kTB], qe[kTB], qp[kTB];
    tile_list(s_list, cw, excl, v0, lane);
    #pragma unroll
    for (int q = 0; q < kTB; q++) qv[q] = (q < nb && q * 64 + lane < total) ? s_list[q * 64 + lane] : -1;
    #pragma unroll
    for (int q = 0; q < kTB; q++) {
        qb[q] = 0;
        qe[q] = 0;
        qp[q] = -1;
        if (qv[q] >= 0) { qb[q] = a.bu_rp[qv[q]]; qe[q] = a.bu_rp[qv[q] + 1]; qp[q] = a.bu_first[qv[q]]; }
    }
    uint32_t qw[kTB];
    #pragma unroll
    for (int q = 0; q < kTB; q++) qw[q] = qp[q] >= 0 ? FL[qp[q] >> 5] : 0u;
    #pragma unroll
    for (int q = 0; q < kTB; q++) qp[q] = (qp[q] >= 0 && ((qw[q] >> (qp[q] & 31)) & 1u)) ? qp[q] : -1;
    // record the batches in order; the registers of batch 0 are used and
    // the others move down one place per batch (constant indices only)
    for (int b = 0; b < nb; b++) {
        const int32_t v = qv[0], beg = qb[0], end = qe[0], par = qp[0];
        #pragma unroll
        for (int q = 0; q + 1 < kTB; q++) {
            qv[q] = qv[q + 1];
            qb[q] = qb[q + 1];
            qe[q] = qe[q + 1];
            qp[q] = qp[q + 1];
        }
        bool open = false;
        if (v >= 0 && end > beg) {
            k.scan += 1ull;
            open = par < 0 && end - beg > 1;
        }
        int bin = 0;
        if (par >= 0) bin = bu_found<PARENTS>(a, v, par, v - v0, beg, end, L, k, s_fn);
        push_new(par >= 0, v, bin, k, s_qn, qout, lane);
        const uint64_t om = __ballot(open);
        if (open) {
            const int pos = qn + (int)__popcll(om & ((1ull << lane) - 1ull));
            s_q[pos] = Pending{v, beg + 1, end, beg};
        }
        qn += (int)__popcll(om);
        if (qn >= 64) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            bu_drain<PARENTS>(a, s_q, 64, v0, L, FL, k, s_qn, qout, s_fn, lane);
            qn -= 64;
            int32_t kv = -1, kb = 0, ke = 0, kb0 = 0;   // field by field, as in bu_drain
            if (lane < qn) { kv = s_q[64 + lane].v; kb = s_q[64 + lane].b; ke = s_q[64 + lane].e; kb0 = s_q[64 + lane].b0; }
            __builtin_amdgcn_wave_barrier();
            if (lane < qn) s_q[lane] = Pending{kv, kb, ke, kb0};
        }
    }
    if (qn > 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        bu_drain<PARENTS>(a, s_q, qn, v0, L, FL, k, s_qn, qout, s_fn, lane);
    }
    return k;
}

// Bottom-up step for one wavefront.  Candidates are the unvisited vertices
// with a non-empty reverse row; each scans its sorted reverse row for the
// first neighbour in F_k, which is its minimum-id parent.  Stage one tests
// every candidate's first entry; candidates left open go to a per-wavefront
// LDS queue that bu_drain empties 64 at a time.
template <bool PARENTS>
__device__ __attribute__((always_inline)) inline void bu_step(const StepArgs& a, int32_t L, const uint32_t* FL, uint32_t* Fn, uint32_t* Fc,
                        Counters& k, uint32_t* s_qn, int32_t* qout, uint32_t* s_fn, Pending* s_q, int32_t* s_list,
                        int gw, int nw, int lane, uint32_t fl_first, uint32_t vis_first, uint32_t nz_first) {
    const int32_t ntiles = (a.nwords + kTileWords - 1) / kTileWords;
    const int sh8 = (lane & 3) * kLaneBits;
    // the next tile's words are loaded before the current tile is scanned
    int32_t t = gw;
    uint32_t fl_n = 0, vs_n = 0, nz_n = 0;
    auto load_tile = [&](int32_t tt) {
        const int32_t wn = tt * kTileWords + (lane >> 2);
        const bool ok = tt < ntiles && wn < a.nwords;
        fl_n = ok ? FL[wn] : 0u;
        vs_n = ok ? a.vis[wn] : 0u;
        nz_n = ok ? a.bu_nz[wn] : 0u;
    };
    // the first tile's words came with the prologue
    if (t < ntiles) {
        fl_n = fl_first;
        vs_n = vis_first;
        nz_n = nz_first;
    }
    while (t < ntiles) {
        const int32_t word0 = t * kTileWords;
        const int32_t w = word0 + (lane >> 2);
        uint32_t cw = 0;
        if (w < a.nwords) {
            const uint32_t visw = vs_n | fl_n;
            cw = ((~visw & nz_n) >> sh8) & 0xffu;   // unvisited with a non-empty reverse row
            if ((lane & 3) == 0) {
                if (fl_n) a.vis[w] = visw;          // fold F_k into vis
                Fc[w] = 0;                          // clear F_{k+2}
            }
        }
        t += nw;
        load_tile(t);
        if (lane < kTileWords) s_fn[lane] = 0;
        int32_t total;
        const int32_t excl = wave_excl_scan(__popc(cw), lane, total);
        k.cand += (lane == 0) ? (unsigned long long)total : 0ull;
        if (total > 0) {
            const Counters kt = bu_tile<PARENTS>(a, L, FL, s_qn, qout, s_fn, s_q, s_list, cw, excl, total, word0, lane);
            k.n += kt.n;
            k.out += kt.out;
            k.in += kt.in;
            k.scan += kt.scan;
            k.cand += kt.cand;
            k.light += kt.light;
            #pragma unroll
            for (int b = 0; b < kHistBins; b++) k.hist[b] += kt.hist[b];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < kTileWords && word0 + lane < a.nwords) Fn[word0 + lane] = s_fn[lane];
        __builtin_amdgcn_wave_barrier();
    }
}

// Top-down step in range mode, for a small frontier with heavy rows.  Work-
// group b owns the vertex ids [b * range_size, (b + 1) * range_size).  It
// reads the whole frontier list (the work-group queues of the previous
// level, s_qoff[] their offsets) and takes from every frontier row the
// segment in its range: heavy rows through the range table, light rows
// whole (entries outside the range are dropped).  The segments are laid out
// in LDS with a prefix sum, so that all wavefronts share the entries
// evenly; the minimum frontier parent of each unvisited vertex of the range
// is formed with LDS atomics, and distances, parents and the range's
// F_{k+1} words are written with plain stores.  No device-scope atomics.
template <bool PARENTS>
__device__ __attribute__((always_inline)) inline void range_step(const StepArgs& a, int32_t L, const uint32_t* FL, uint32_t* Fn, uint32_t* Fc,
                           Counters& k, uint32_t* s_qn, int32_t* qout, uint32_t* s_lds, const int32_t* s_qoff,
                           int32_t nf, int lpar, int tid, int lane, int wid) {
    const int32_t RS = a.range_size;              // a multiple of 64
    const int32_t R0 = (int32_t)blockIdx.x * RS;
    const int32_t nwr = RS / 32;
    const int32_t w0 = R0 / 32;
    uint32_t* s_par = s_lds;                      // [RS] minimum frontier parent, 0xffffffff = none
    uint32_t* s_seen = s_par + RS;                // [nwr] visited | F_k words of the range
    int32_t* s_su = reinterpret_cast<int32_t*>(s_seen + nwr);   // [kRangeFront] segment vertex
    int32_t* s_sb = s_su + kRangeFront;           // [kRangeFront] segment first entry
    int32_t* s_sx = s_sb + kRangeFront;           // [kRangeFront + 1] segment offsets (prefix)
    int32_t* s_ws = s_sx + kRangeFront + 1;       // [kWaves] wavefront totals
    // segments: thread t takes frontier entries 2t and 2t + 1.  Their
    // vertices, row bounds and heavy-row ids are loaded first (the three
    // loads of a vertex together), so that their round trips overlap the
    // sweep of the range's bitmap words below.
    const int32_t* qin = a.qbuf + (size_t)(lpar ^ 1) * a.grid * kQCap;
    int32_t fu[2] = {-1, -1};
    #pragma unroll
    for (int q = 0; q < 2; q++) {
        const int32_t j = 2 * tid + q;
        if (j < nf) {
            int lo = 0, hi = a.grid;                  // largest b with s_qoff[b] <= j
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_qoff[mid] <= j) lo = mid;
                else hi = mid;
            }
            fu[q] = qin[(size_t)lo * kQCap + (j - s_qoff[lo])];
        }
    }
    int32_t fb[2] = {0, 0}, fe[2] = {0, 0}, fh[2] = {-1, -1};
    #pragma unroll
    for (int q = 0; q < 2; q++) {
        if (fu[q] >= 0) {
            fb[q] = a.td_rp[fu[q]];
            fe[q] = a.td_rp[fu[q] + 1];
            fh[q] = a.td_heavy_id[fu[q]];
        }
    }
    for (int32_t i = tid; i < nwr; i += kBlock) {
        const int32_t w = w0 + i;
        uint32_t sw = 0xffffffffu;
        if (w < a.nwords) {
            const uint32_t fl = FL[w], vs = a.vis[w];
            sw = vs | fl;
            if (fl) a.vis[w] = sw;                // fold F_k into vis
            Fc[w] = 0;                            // clear F_{k+2}
        }
        s_seen[i] = sw;
    }
    for (int32_t i = tid; i < RS; i += kBlock) s_par[i] = 0xffffffffu;
    int32_t len[2] = {0, 0};
    #pragma unroll
    for (int q = 0; q < 2; q++) {
        if (fu[q] >= 0) {
            int32_t sb = fb[q], se = fe[q];
            if (se - sb > kHeavyRow) {
                const int32_t* t = a.td_rtab + (size_t)fh[q] * (a.grid + 1) + blockIdx.x;
                sb = t[0];
                se = t[1];
            }
            const int32_t j = 2 * tid + q;
            s_su[j] = fu[q];
            s_sb[j] = sb;
            len[q] = se - sb;
        }
    }
    int32_t wtot;
    const int32_t ex = wave_excl_scan(len[0] + len[1], lane, wtot);
    if (lane == 0) s_ws[wid] = wtot;
    lds_barrier();
    if (wid == 0) {
        const int32_t x = lane < kWaves ? s_ws[lane] : 0;
        int32_t t2;
        const int32_t e2 = wave_excl_scan(x, lane, t2);
        if (lane < kWaves) s_ws[lane] = e2;
        if (lane == 0) s_sx[nf] = t2;
    }
    lds_barrier();
    {
        const int32_t off = s_ws[wid] + ex;
        if (2 * tid < nf) s_sx[2 * tid] = off;
        if (2 * tid + 1 < nf) s_sx[2 * tid + 1] = off + len[0];
    }
    lds_barrier();
    // entries: lane l of a round takes 4 consecutive entries
    const int32_t T = s_sx[nf];
    if (lane == 0 && wid == 0) k.scan += (unsigned long long)T;
    for (int32_t e0 = wid * 256; e0 < T; e0 += kWaves * 256) {
        const int32_t e = e0 + lane * 4;
        int32_t uu[4], vv[4];
        int seg = 0;
        if (e < T) {
            int lo = 0, hi = nf;                      // largest segment with s_sx[seg] <= e
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_sx[mid] <= e) lo = mid;
                else hi = mid;
            }
            seg = lo;
        }
        #pragma unroll
        for (int q = 0; q < 4; q++) {
            const int32_t eq = e + q;
            uu[q] = 0;
            vv[q] = -1;
            if (eq < T) {
                while (s_sx[seg + 1] <= eq) seg++;
                uu[q] = s_su[seg];
                vv[q] = a.td_col[s_sb[seg] + (eq - s_sx[seg])];
            }
        }
        #pragma unroll
        for (int q = 0; q < 4; q++) {
            const int32_t loc = vv[q] - R0;
            if (vv[q] >= 0 && loc >= 0 && loc < RS && !((s_seen[loc >> 5] >> (loc & 31)) & 1u))
                atomicMin(&s_par[loc], (uint32_t)uu[q]);
        }
    }
    lds_barrier();
    // emit: eight 1024-vertex slices per round, so that the degree loads of
    // their new vertices are in flight together
    for (int32_t c0 = 0; c0 < RS; c0 += kEmit * kBlock) {
        int32_t v[kEmit], d0[kEmit], d1[kEmit];
        uint32_t par[kEmit];
        bool fresh[kEmit];
        #pragma unroll
        for (int q = 0; q < kEmit; q++) {
            const int32_t loc = c0 + q * kBlock + tid;
            par[q] = loc < RS ? s_par[loc] : 0xffffffffu;
            v[q] = R0 + loc;
            fresh[q] = par[q] != 0xffffffffu && v[q] < a.nv;
            d0[q] = d1[q] = 0;
            if (fresh[q]) { d0[q] = a.td_rp[v[q]]; d1[q] = a.td_rp[v[q] + 1]; }
        }
        int bin[kEmit];
        #pragma unroll
        for (int q = 0; q < kEmit; q++) {
            const int32_t base = c0 + q * kBlock + wid * 64;
            const uint64_t m = __ballot(fresh[q]);
            if (lane < 2 && base < RS) {
                const int32_t w = w0 + (base >> 5) + lane;
                if (w < a.nwords) Fn[w] = (uint32_t)(m >> (32 * lane));
            }
            bin[q] = 0;
            if (fresh[q]) {
                __builtin_nontemporal_store(L + 1, &a.dist[v[q]]);
                if (PARENTS) a.parent[v[q]] = par[q];
                const int32_t dout = d1[q] - d0[q];
                const int32_t din = (a.td_rp == a.bu_rp) ? dout : (a.bu_rp[v[q] + 1] - a.bu_rp[v[q]]);
                k.n++;
                k.out += (unsigned long long)dout;
                k.in += (unsigned long long)din;
                if (dout <= kHeavyRow) k.light += (unsigned long long)dout;
                bin[q] = hist_bin(dout);
            }
        }
        push_new_n<kEmit>(fresh, v, bin, k, s_qn, qout, lane);
    }
}

// Top-down step in queue mode, for a small frontier whose work-group queues
// are complete (or the single source of level 0).  Every work-group reads
// the whole frontier list, lays the frontier rows' entries out with a prefix
// sum in LDS and expands an equal share of them, work-group b the entries
// [b * T / G, (b + 1) * T / G), one entry per lane and round; the visit is
// the sweep's (plain-load filter, atomicMin claim).  No bitmap is swept: the
// frontier's vertices are folded into the visited bitmap one by one, and
// F_{k+2} is cleared with plain stores only.  One entry per lane keeps the
// inlined code small: with four, the level kernel was 2.5 KB larger and
// every level, of any variant, slower (a 4096 x 4096 grid 238 against
// 218 ms for two searches, RMAT-22 10.08 against 9.92 ms for 64).
template <bool PARENTS>
__device__ __attribute__((always_inline)) inline void queue_step(const StepArgs& a, int32_t L, const uint32_t* FL,
                                                                 uint32_t* Fn, uint32_t* Fc, Counters& k,
                                                                 uint32_t* s_qn, int32_t* qout, uint32_t* s_lds,
                                                                 const int32_t* s_qoff, int32_t nf, int32_t src0,
                                                                 int lpar, int tid, int lane, int wid) {
    constexpr int kPer = kQueueFront / kBlock;   // frontier entries per thread
    int32_t* s_su = reinterpret_cast<int32_t*>(s_lds);   // [kQueueFront] frontier vertex
    int32_t* s_sb = s_su + kQueueFront;                   // [kQueueFront] first entry of its row
    int32_t* s_sx = s_sb + kQueueFront;                   // [kQueueFront + 1] entry offsets (prefix)
    int32_t* s_ws = s_sx + kQueueFront + 1;               // [kWaves] wavefront totals
    const int32_t* qin = a.qbuf + (size_t)(lpar ^ 1) * a.grid * kQCap;
    // thread t holds frontier entries kPer * t .. kPer * t + kPer - 1
    int32_t fu[kPer];
    #pragma unroll
    for (int q = 0; q < kPer; q++) {
        const int32_t j = kPer * tid + q;
        fu[q] = -1;
        if (j < nf) {
            if (src0 >= 0) {
                fu[q] = src0;
            } else {
                int lo = 0, hi = a.grid;                  // largest b with s_qoff[b] <= j
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (s_qoff[mid] <= j) lo = mid;
                    else hi = mid;
                }
                fu[q] = qin[(size_t)lo * kQCap + (j - s_qoff[lo])];
            }
        }
    }
    int32_t fb[kPer], len[kPer];
    #pragma unroll
    for (int q = 0; q < kPer; q++) {
        fb[q] = 0;
        len[q] = 0;
        if (fu[q] >= 0) {
            fb[q] = a.td_rp[fu[q]];
            len[q] = a.td_rp[fu[q] + 1] - fb[q];
        }
    }
    // F_k into the visited bitmap (each frontier vertex by one work-group),
    // F_{k+2} cleared (stores only); neither is read by this step
    #pragma unroll
    for (int q = 0; q < kPer; q++) {
        const int32_t j = kPer * tid + q;
        if (fu[q] >= 0 && (j % a.grid) == (int32_t)blockIdx.x)
            atomicOr(&a.vis[fu[q] >> 5], 1u << (fu[q] & 31));
    }
    for (int64_t i = (int64_t)blockIdx.x * kBlock + tid; i < a.nwords / 2; i += (int64_t)a.grid * kBlock)
        reinterpret_cast<uint2*>(Fc)[i] = make_uint2(0u, 0u);
    // entry offsets
    int32_t tsum = 0;
    #pragma unroll
    for (int q = 0; q < kPer; q++) tsum += len[q];
    int32_t wtot;
    const int32_t ex = wave_excl_scan(tsum, lane, wtot);
    if (lane == 0) s_ws[wid] = wtot;
    lds_barrier();
    if (wid == 0) {
        const int32_t x = lane < kWaves ? s_ws[lane] : 0;
        int32_t t2;
        const int32_t e2 = wave_excl_scan(x, lane, t2);
        if (lane < kWaves) s_ws[lane] = e2;
        if (lane == 0) s_sx[nf] = t2;
    }
    lds_barrier();
    {
        int32_t off = s_ws[wid] + ex;
        #pragma unroll
        for (int q = 0; q < kPer; q++) {
            const int32_t j = kPer * tid + q;
            if (j < nf) {
                s_su[j] = fu[q];
                s_sb[j] = fb[q];
                s_sx[j] = off;
            }
            off += len[q];
        }
    }
    lds_barrier();
    const int64_t T = s_sx[nf];
    const int32_t e0 = (int32_t)(T * (int64_t)blockIdx.x / a.grid);
    const int32_t e1 = (int32_t)(T * (int64_t)(blockIdx.x + 1) / a.grid);
    for (int32_t base = e0 + wid * 64; base < e1; base += kWaves * 64) {
        const int32_t e = base + lane;
        int32_t uu[1] = {0}, vv[1] = {-1};
        if (e < e1) {
            int lo = 0, hi = nf;                          // largest segment with s_sx[seg] <= e
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_sx[mid] <= e) lo = mid;
                else hi = mid;
            }
            uu[0] = s_su[lo];
            vv[0] = a.td_col[s_sb[lo] + (e - s_sx[lo])];
            k.scan += 1ull;
        }
        visit_n<PARENTS, 1>(a, uu, vv, L, FL, Fn, k, s_qn, qout, lane);
    }
}

// Solo-mode visit of N edges u[q] -> v[q] (v[q] < 0: none) by work-group 0
// alone.  The visited bitmap holds every level up to L (each level's
// frontier is folded in at its start, and the reads are coherent with
// those atomics), so a vertex is fresh when its bit is clear.  With parents
// the atomicMin that finds the parent unset claims the vertex; without, an
// atomicCAS of its distance.  New vertices go to the LDS list s_next (one
// LDS atomic per wavefront and call); those past its capacity have their
// bits set in Fn at once and are counted in `ovf` (per thread, histogram
// included).  Every lane of the wavefront calls it.
template <bool PARENTS, int N>
__device__ __attribute__((always_inline)) inline void visit_solo(const StepArgs& a, const int32_t* u, const int32_t* v,
                                                                 int32_t L, uint32_t* Fn, int32_t* s_next,
                                                                 uint32_t* s_nn, Counters& ovf, int lane) {
    uint32_t seen[N], pre[N];
    #pragma unroll
    for (int q = 0; q < N; q++) {
        seen[q] = 0xffffffffu;
        pre[q] = 0;
        if (v[q] >= 0) {
            seen[q] = __hip_atomic_load(&a.vis[v[q] >> 5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (PARENTS) pre[q] = a.parent[v[q]];
        }
    }
    bool claim[N];
    #pragma unroll
    for (int q = 0; q < N; q++) {
        const bool fresh = v[q] >= 0 && !((seen[q] >> (v[q] & 31)) & 1u);
        // parent only decreases, so a stale value <= u already rules u out
        claim[q] = fresh && (PARENTS ? pre[q] > (uint32_t)u[q] : true);
    }
    uint32_t old[N];
    #pragma unroll
    for (int q = 0; q < N; q++) {
        old[q] = 0;
        if (claim[q]) {
            if (PARENTS) old[q] = atomicMin(&a.parent[v[q]], (uint32_t)u[q]);
            else old[q] = (uint32_t)atomicCAS(&a.dist[v[q]], -1, L + 1);
        }
    }
    bool won[N];
    uint64_t m[N];
    uint32_t tot = 0;
    #pragma unroll
    for (int q = 0; q < N; q++) {
        won[q] = claim[q] && old[q] == 0xffffffffu;
        if (PARENTS && won[q]) __builtin_nontemporal_store(L + 1, &a.dist[v[q]]);
        m[q] = __ballot(won[q]);
        tot += (uint32_t)__popcll(m[q]);
    }
    if (!tot) return;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(s_nn, tot);
    base = __shfl(base, 0);
    const uint64_t lt = (1ull << lane) - 1ull;
    #pragma unroll
    for (int q = 0; q < N; q++) {
        if (won[q]) {
            const uint32_t pos = base + (uint32_t)__popcll(m[q] & lt);
            if (pos < (uint32_t)kSoloFront) {
                s_next[pos] = v[q];
            } else {
                atomicOr(&Fn[v[q] >> 5], 1u << (v[q] & 31));
                const int bin = account_new(a, v[q], -1, ovf);
                #pragma unroll
                for (int b = 0; b < kHistBins; b++) ovf.hist[b] += bin == b ? 1u : 0u;
            }
        }
        base += (uint32_t)__popcll(m[q]);
    }
}

// Solo mode.  Work-group 0 alone runs top-down levels L, L+1, ... while the
// frontier has at most a.solo_front vertices and a.solo_edges edges;
// work-group barriers take the place of the launches between them.  The
// frontier is a list in LDS, expanded with the edges dealt out evenly over
// the threads (a contiguous run each, found with one binary search over the
// rows' edge offsets).  No frontier bitmap is written for the levels inside
// the loop: each level's frontier is folded into the visited bitmap at its
// start, and the visited bitmap alone tells fresh vertices.  The loop ends
// at the first level s whose frontier F_s is too large or empty:
//  * F_s too large (or past the list): F_s's bits go to Fn = F_{p+1}, the
//    frontier bitmap of launch p + 1, its list to this work-group's queue,
//    S_{s-1} to state[p & 1] and the counters of step s-1 to `k` for the
//    epilogue: what the next launch's prologue reads, as if launch p had
//    run level s-1 alone;
//  * F_s empty: the search has s levels; S_s and done_level = p + 1 are
//    stored, as the next launch's prologue would have stored them.
// Every level run here gets its record in the level log.
template <bool PARENTS>
__device__ __attribute__((always_inline)) inline void solo_loop(const StepArgs& a, BfsCtrl* s_cp, int32_t p, int lpar,
                                                                uint32_t* Fn, int32_t* qout, uint32_t* s_lds,
                                                                const int32_t* s_qoff, Counters& k, int tid, int lane,
                                                                int wid) {
    int32_t* s_fa = reinterpret_cast<int32_t*>(s_lds);    // [kSoloFront] frontier list
    int32_t* s_fb = s_fa + kSoloFront;                     // [kSoloFront] next list
    int32_t* s_sb = s_fb + kSoloFront;                     // [kSoloFront] first entry of each frontier row
    int32_t* s_sx = s_sb + kSoloFront;                     // [kSoloFront + 1] edge offsets (prefix)
    int32_t* s_ws = s_sx + kSoloFront + 1;                 // [kWaves] wavefront totals
    uint32_t* s_nn = reinterpret_cast<uint32_t*>(s_ws + kWaves);   // entries appended to the next list
    uint32_t* s_red = s_nn + 1;                            // [kWaves][12] counter partials
    uint32_t* s_tot = s_red + kWaves * 12;                 // [12] totals: n, out, in, light, hist[8]
    int32_t* s_go = reinterpret_cast<int32_t*>(s_tot + 12);   // 0 run the level, 1 hand over, 2 ended
    static_assert(4 * kSoloFront + 1 + kWaves + 1 + kWaves * 12 + 12 + 1 <=
                      kRangeMax + kRangeMax / 32 + 3 * kRangeFront + 1 + kWaves,
                  "solo-mode lists exceed the LDS area");
    static_assert(kSoloFront <= kQCap, "solo mode hands its frontier over through a work-group queue");
    static_assert(kSoloFront % kBlock == 0, "solo mode: whole frontier entries per thread");

    int32_t n = s_cp->n_front;
    const int32_t L0 = s_cp->level;
    {
        // F_L0 from the work-group queues of launch p-1
        const int32_t* qin = a.qbuf + (size_t)(lpar ^ 1) * a.grid * kQCap;
        for (int32_t j = tid; j < n; j += kBlock) {
            int lo = 0, hi = a.grid;                      // largest b with s_qoff[b] <= j
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_qoff[mid] <= j) lo = mid;
                else hi = mid;
            }
            s_fa[j] = qin[(size_t)lo * kQCap + (j - s_qoff[lo])];
        }
    }
    if (tid == 0) {
        s_cp->dir = 0;
        s_cp->variant = 3;
        s_cp->phase = p;
        *s_nn = 0;
    }
    __syncthreads();
    int32_t* cur = s_fa;
    int32_t* nxt = s_fb;
    int32_t sl = L0;              // the level being run
    int64_t t_prev = 0;           // edges expanded by step sl - 1
    bool over = false;            // F_sl has vertices beyond the list
    Counters ovf;                 // this thread's share of them
    for (;;) {
        // rows of F_sl, F_sl folded into the visited bitmap, F_sl's counters
        uint32_t cnt[12];
        cnt[0] = (uint32_t)ovf.n;
        cnt[1] = (uint32_t)ovf.out;
        cnt[2] = (uint32_t)ovf.in;
        cnt[3] = (uint32_t)ovf.light;
        #pragma unroll
        for (int b = 0; b < kHistBins; b++) cnt[4 + b] = ovf.hist[b];
        ovf = Counters();
        int32_t fb[kSoloPer], len[kSoloPer];
        #pragma unroll
        for (int q = 0; q < kSoloPer; q++) {
            const int32_t j = kSoloPer * tid + q;
            fb[q] = 0;
            len[q] = 0;
            if (j < n) {
                const int32_t u = cur[j];
                fb[q] = a.td_rp[u];
                len[q] = a.td_rp[u + 1] - fb[q];
                const int32_t din = a.td_rp == a.bu_rp ? len[q] : a.bu_rp[u + 1] - a.bu_rp[u];
                atomicOr(&a.vis[u >> 5], 1u << (u & 31));
                cnt[0] += 1u;
                cnt[1] += (uint32_t)len[q];
                cnt[2] += (uint32_t)din;
                if (len[q] <= kHeavyRow) cnt[3] += (uint32_t)len[q];
                const int bin = hist_bin(len[q]);
                #pragma unroll
                for (int b = 0; b < kHistBins; b++) cnt[4 + b] += bin == b ? 1u : 0u;
            }
        }
        int32_t tsum = 0;
        #pragma unroll
        for (int q = 0; q < kSoloPer; q++) tsum += len[q];
        int32_t wtot;
        const int32_t ex = wave_excl_scan(tsum, lane, wtot);
        #pragma unroll
        for (int i = 0; i < 12; i++) cnt[i] = wave_sum(cnt[i]);
        if (lane == 0) {
            s_ws[wid] = wtot;
            #pragma unroll
            for (int i = 0; i < 12; i++) s_red[wid * 12 + i] = cnt[i];
        }
        __syncthreads();   // also completes the visited-bitmap folds
        if (wid == 0) {
            const int32_t x = lane < kWaves ? s_ws[lane] : 0;
            int32_t t2;
            const int32_t e2 = wave_excl_scan(x, lane, t2);
            if (lane < kWaves) s_ws[lane] = e2;
            if (lane < 12) {
                uint32_t v = 0;
                for (int w = 0; w < kWaves; w++) v += s_red[w * 12 + lane];
                s_tot[lane] = v;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane == 0) {
                s_sx[n] = t2;
                int go = 0;
                if (sl > L0) {
                    // S_sl from S_{sl-1} and the counters of step sl-1
                    const uint64_t now = wall_clock64();
                    const BfsCtrl q = *s_cp;
                    BfsCtrl c = q;
                    const int32_t ntot = (int32_t)s_tot[0];
                    c.level = sl;
                    c.n_prev = q.n_front;
                    c.n_front = ntot;
                    c.prev_dir = 0;
                    c.m_front = (int64_t)s_tot[1];
                    c.m_unvisited = q.m_unvisited - (int64_t)s_tot[2];
                    c.n_unvisited = q.n_unvisited - ntot;
                    c.m_light = (int64_t)s_tot[3];
                    for (int b = 0; b < kHistBins; b++) c.front_hist[b] = s_tot[4 + b];
                    c.qovf = over ? 1 : 0;
                    c.t_main = now;
                    if (sl - 1 < a.max_log) {
                        LevelRec* r = &a.log[sl - 1];
                        r->dir = 0;
                        r->variant = 3;
                        r->n_front = q.n_front;
                        r->m_front = q.m_front;
                        r->m_unvisited = q.m_unvisited;
                        r->n_next = ntot;
                        r->scanned = t_prev;
                        r->cand = 0;
                        r->t_start = q.t_main;
                        r->t_end = now;
                        r->t_prev_end = q.t_main;
                        r->launch = p;
                    }
                    if (ntot == 0) {
                        go = 2;
                        a.sh->state[lpar] = c;
                        set_done(a, p + 1, sl);
                    } else if (over || ntot > a.solo_front || c.m_front > a.solo_edges ||
                               c.m_front > 8 * (int64_t)ntot || ntot > 2 * q.n_front + 64) {
                        go = 1;
                        a.sh->state[lpar] = q;
                    } else {
                        if (sl < a.max_log) {
                            float f[kFeat];
                            compute_features(&c, f);
                            for (int i = 0; i < kFeat; i++) a.log[sl].feat[i] = f[i];
                        }
                        *s_cp = c;
                    }
                }
                *s_go = go;
            }
        }
        __syncthreads();
        const int go = *s_go;
        if (go == 2) return;          // k stays zero
        if (go == 1) {
            // hand F_sl over to launch p + 1
            for (int32_t j = tid; j < n; j += kBlock) {
                const int32_t u = cur[j];
                atomicOr(&Fn[u >> 5], 1u << (u & 31));
                qout[j] = u;
            }
            if (wid == 0) {
                if (lane == 0) {
                    k.n = s_tot[0];
                    k.out = s_tot[1];
                    k.in = s_tot[2];
                    k.light = s_tot[3];
                    k.scan = (unsigned long long)t_prev;
                }
                #pragma unroll
                for (int b = 0; b < kHistBins; b++) k.hist[b] = s_tot[4 + b];
            }
            return;
        }
        if (tid == 0) *s_nn = 0;
        // the rows' first entries and edge offsets
        {
            int32_t off = s_ws[wid] + ex;
            #pragma unroll
            for (int q = 0; q < kSoloPer; q++) {
                const int32_t j = kSoloPer * tid + q;
                if (j < n) {
                    s_sb[j] = fb[q];
                    s_sx[j] = off;
                }
                off += len[q];
            }
        }
        __syncthreads();
        // step sl -> sl + 1: a contiguous run of edges per thread
        const int32_t T = s_sx[n];
        const int32_t per = (T + kBlock - 1) / kBlock;
        const int32_t e0 = min(T, tid * per), e1 = min(T, e0 + per);
        int lo = 0;
        if (e0 < e1) {
            int hi = n;                                   // largest row with s_sx[row] <= e0
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_sx[mid] <= e0) lo = mid;
                else hi = mid;
            }
        }
        for (int32_t base = 0; base < per; base += kSoloGroup) {
            int32_t uu[kSoloGroup], vv[kSoloGroup];
            #pragma unroll
            for (int q = 0; q < kSoloGroup; q++) {
                const int32_t e = e0 + base + q;
                uu[q] = 0;
                vv[q] = -1;
                if (e < e1) {
                    while (s_sx[lo + 1] <= e) lo++;
                    uu[q] = cur[lo];
                    vv[q] = a.td_col[s_sb[lo] + (e - s_sx[lo])];
                }
            }
            visit_solo<PARENTS, kSoloGroup>(a, uu, vv, sl, Fn, nxt, s_nn, ovf, lane);
        }
        __syncthreads();
        const uint32_t nn = *s_nn;
        over = nn > (uint32_t)kSoloFront;
        n = over ? kSoloFront : (int32_t)nn;
        t_prev = T;
        int32_t* t = cur;
        cur = nxt;
        nxt = t;
        sl++;
    }
}

// The learned policy's weights as lane j of wavefront 0 needs them (hidden
// unit j), loaded with the prologue's other loads.  The device copy has the
// feature standardisation folded into W1 and b1 (done on the host when the
// policy is set), so inference is 16 FMAs per lane and a butterfly sum.
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

// Beamer alpha/beta (GAP form): leave top-down when the frontier's edges
// exceed the unvisited edges / alpha; leave bottom-up when the frontier
// shrinks and holds at most nv / beta vertices.
__device__ inline int heuristic_direction(const BfsCtrl& c) {
    if (c.policy_mode == LGMI_POLICY_TOPDOWN) return 0;
    if (c.prev_dir == 0)
        return ((double)c.m_front > (double)c.m_unvisited / (double)c.alpha) ? 1 : 0;
    const bool shrinking = c.n_front < c.n_prev;
    const bool small = (double)c.n_front <= (double)c.nv / (double)c.beta;
    return (shrinking && small) ? 0 : 1;
}

// The network decides contested levels only: after a bottom-up step, or
// when the frontier's edges exceed 1/64 of the unvisited edges.  Below that
// a top-down step costs a small fraction of a bottom-up sweep and the
// alpha/beta rule decides (top-down for any alpha up to 64); skipping the
// network there saves its features and inference in the prologue.
__device__ inline bool policy_contested(const BfsCtrl& c) {
    return c.prev_dir == 1 || (double)c.m_front * 64.0 > (double)c.m_unvisited;
}

// MLP logits by wavefront 0 (lane j = hidden unit j); returns the
// direction in every lane.
__device__ inline int mlp_direction(const float* xs, const PolicyRegs& pr) {
    float acc = pr.b1;
    #pragma unroll
    for (int q = 0; q < kFeat / 4; q++) {
        const float4 r = pr.w1[q];
        acc += r.x * xs[4 * q] + r.y * xs[4 * q + 1] + r.z * xs[4 * q + 2] + r.w * xs[4 * q + 3];
    }
    const float hj = acc > 0.f ? acc : 0.f;
    const float a0 = wave_sum_f(pr.w2a * hj), a1 = wave_sum_f(pr.w2b * hj);
    return (a1 + pr.b2b) > (a0 + pr.b2a) ? 1 : 0;
}

// One level of the search.  karg >= 0: the level p itself; karg < 0
// (captured graphs): p is unknown to the host but p mod 6 = -1 - karg (so
// its parity and the bitmap rotation are static), and p is read from the
// state the launch before stored.  forced >= 0 forces the direction.
// Every device function the level kernel calls is inlined (always_inline
// on the large ones): a call that stays a call makes the kernel pass its
// arguments through per-lane scratch memory.
template <bool PARENTS>
__global__ __launch_bounds__(kBlock) void bfs_level(StepArgs a, int32_t karg, int32_t forced) {
    __shared__ BfsCtrl s_c;
    __shared__ unsigned long long s_red[kWaves][kPart];
    __shared__ unsigned long long s_sum[kPart];
    __shared__ float s_feat[kFeat];
    __shared__ int32_t s_p, s_dir, s_exit, s_dl, s_mode, s_src0, s_solo, s_learn;
    __shared__ uint32_t s_qn[1];                 // the work-group queue's count
    __shared__ uint32_t s_fn[kWaves][kTileWords];
    __shared__ int32_t s_qoff[kMaxGrid + 1];      // per-work-group new-vertex counts, then queue offsets
    // bottom-up: the wavefronts' pending queues; range mode: parents and
    // visited words of the work-group's range
    constexpr int kRangeWords = kRangeMax + kRangeMax / 32 + 3 * kRangeFront + 1 + kWaves;
    constexpr int kLdsWords = kRangeWords > (kWaves * 128 * 4) ? kRangeWords : (kWaves * 128 * 4);
    __shared__ uint32_t s_lds[kLdsWords];
    Pending* s_q = reinterpret_cast<Pending*>(s_lds);
    // per-wavefront tile lists of the sweep steps, after the queues
    static_assert(kLdsWords >= kWaves * 128 * 4 + kWaves * kTileVerts, "tile lists exceed the LDS area");
    static_assert(kLdsWords >= 3 * kQueueFront + 1 + kWaves, "queue-mode lists exceed the LDS area");
    static_assert(kQueueFront % kBlock == 0, "queue mode: whole frontier entries per thread");
    int32_t* s_lists = reinterpret_cast<int32_t*>(s_lds + kWaves * 128 * 4);

    // the wavefront index is uniform over a wavefront; readfirstlane lets the
    // compiler keep it, and the tile loops that depend on it, scalar
    const int tid = threadIdx.x, lane = tid & 63, wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint64_t t_now = wall_clock64();
    const int pmod6 = karg >= 0 ? (karg % 6) : (-1 - karg);   // p mod 6
    const int lpar = pmod6 & 1;                                 // parity of p
    const bool first = karg == 0;
    BfsShared* sh = a.sh;
    const int gw = blockIdx.x * kWaves + wid, nw = gridDim.x * kWaves;

    // The words of this wavefront's first tile of F_p, the visited bitmap and
    // the non-empty-row bitmap: the sweep steps read them first whatever the
    // direction, so their loads go out with the prologue's.
    uint32_t pf_fl = 0, pf_vis = 0, pf_nz = 0;
    {
        const uint32_t* FLp = bm_at(a, pmod6 % 3);
        const int32_t w = gw * kTileWords + (lane >> 2);
        if (w < a.nwords) {
            pf_fl = FLp[w];
            pf_vis = a.vis[w];
            pf_nz = a.bu_nz[w];
        }
    }

    // ---- prologue: the loads below are independent of each other ----
    unsigned long long acc = 0;
    if (!first) {
        const unsigned long long* P = a.part + (size_t)(lpar ^ 1) * a.grid * kPart;
        const int slot = tid % kPart, row = tid / kPart;
        unsigned long long x0 = 0, x1 = 0, x2 = 0, x3 = 0;
        int b = row;
        for (; b + 3 * kPartRows < a.grid; b += 4 * kPartRows) {
            const unsigned long long y0 = P[(size_t)b * kPart + slot];
            const unsigned long long y1 = P[(size_t)(b + kPartRows) * kPart + slot];
            const unsigned long long y2 = P[(size_t)(b + 2 * kPartRows) * kPart + slot];
            const unsigned long long y3 = P[(size_t)(b + 3 * kPartRows) * kPart + slot];
            if (slot == P_NEXT) {
                s_qoff[b] = (int32_t)y0;
                s_qoff[b + kPartRows] = (int32_t)y1;
                s_qoff[b + 2 * kPartRows] = (int32_t)y2;
                s_qoff[b + 3 * kPartRows] = (int32_t)y3;
            }
            x0 += y0; x1 += y1; x2 += y2; x3 += y3;
        }
        for (; b < a.grid; b += kPartRows) {
            const unsigned long long y0 = P[(size_t)b * kPart + slot];
            if (slot == P_NEXT) s_qoff[b] = (int32_t)y0;
            x0 += y0;
        }
        acc = (x0 + x1) + (x2 + x3);
    }
    PolicyRegs pr;
    const bool may_learn = a.policy != nullptr && forced < 0 && !first;
    if (may_learn && tid < kHidden) load_policy(a.policy, tid, pr);
    if (tid == 0) {
        s_c = sh->state[lpar ^ 1];   // the first launch reads S_0 from state[1]
        s_dl = sh->done_level;
        s_src0 = first ? sh->src0 : -1;
    }
    if (tid == 0) s_qn[0] = 0;
    // lanes l, l+16, l+32, l+48 of a wavefront hold the same slot
    acc += __shfl_xor(acc, 16);
    acc += __shfl_xor(acc, 32);
    if (lane < kPart) s_red[wid][lane] = acc;
    __syncthreads();
    // Wavefront 0 derives the level state (lane 0) and the direction (lane 0
    // for the heuristic, all 64 lanes for the MLP); wavefront-level ordering
    // suffices inside it, and one work-group barrier publishes the result.
    if (wid == 0) {
        if (lane < kPart) {
            unsigned long long s = 0;
            for (int q = 0; q < kWaves; q++) s += s_red[q][lane];
            s_sum[lane] = s;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane == 0) {
            const int32_t p = first ? 0 : (karg > 0 ? karg : s_c.phase + 1);
            int ex = s_dl <= p ? 1 : 0;
            if (!ex && !first) {
                // S_k from S_{k-1} and the counters of level k-1
                const BfsCtrl& q = s_c;
                BfsCtrl c = q;
                const int32_t k = q.level + 1;
                const int32_t n_next = (int32_t)s_sum[P_NEXT];
                c.level = k;
                c.n_prev = q.n_front;
                c.n_front = n_next;
                c.prev_dir = q.dir;
                c.m_front = (int64_t)s_sum[P_OUT];
                c.m_unvisited = q.m_unvisited - (int64_t)s_sum[P_IN];
                c.n_unvisited = q.n_unvisited - n_next;
                c.m_light = (int64_t)s_sum[P_LIGHT];
                c.qovf = s_sum[P_QOVF] != 0 ? 1 : 0;
                c.phase = p;
                c.t_main = t_now;
                for (int b = 0; b < kHistBins; b++) c.front_hist[b] = (uint32_t)s_sum[P_HIST + b];
                if (blockIdx.x == 0 && k - 1 < a.max_log) {
                    LevelRec* r = &a.log[k - 1];
                    r->dir = q.dir;
                    r->variant = q.variant;
                    r->n_front = q.n_front;
                    r->m_front = q.m_front;
                    r->m_unvisited = q.m_unvisited;
                    r->n_next = n_next;
                    r->scanned = (int64_t)s_sum[P_SCAN];
                    r->cand = (int64_t)s_sum[P_CAND];
                    r->t_start = q.t_main;
                    r->t_end = t_now;
                    r->t_prev_end = q.t_main;
                    r->launch = q.phase;
                }
                if (n_next == 0) {
                    ex = 1;
                    if (blockIdx.x == 0) {
                        sh->state[p & 1] = c;
                        set_done(a, p, k);
                    }
                } else if ((may_learn && policy_contested(c)) || (blockIdx.x == 0 && k < a.max_log)) {
                    // the features feed the network and work-group 0's level log only
                    compute_features(&c, s_feat);
                }
                s_c = c;
            }
            s_p = p;
            s_exit = ex;
            // solo mode: a run of small levels -- this frontier and the
            // last one small, this one of low-degree rows (at most 8
            // entries per vertex on average), its list complete, neither
            // growing nor shrinking fast: a chain, a grid -- whatever
            // direction the policy would pick.  A solo level costs more
            // than an ordinary one when the run ends after it (the next
            // frontier's rows are read before the hand-over), so a lone
            // small level (the last levels of a search shrink fast), a
            // frontier of hub rows that explodes next, and level 0 stay
            // with the other modes.
            const bool solo = !ex && !first && forced < 0 && a.solo_front > 0 && s_c.qovf == 0 &&
                              s_c.n_front <= a.solo_front && s_c.n_prev <= a.solo_front &&
                              s_c.m_front <= a.solo_edges && s_c.m_front <= 8 * (int64_t)s_c.n_front &&
                              s_c.n_front <= 2 * s_c.n_prev + 64 && 2 * s_c.n_front >= s_c.n_prev;
            s_solo = solo ? 1 : 0;
            const bool learn = !ex && may_learn && !solo && policy_contested(s_c);
            s_learn = learn ? 1 : 0;
            if (!ex && !learn)
                s_dir = solo ? 0 : (forced >= 0 ? forced : (first ? s_c.dir : heuristic_direction(s_c)));
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (s_learn) {
            const int d = mlp_direction(s_feat, pr);
            if (lane == 0) s_dir = d;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // queue mode: top-down over a small frontier with few edges whose
        // work-group queues are complete (level 0: a single source);
        // range mode: top-down over a small frontier that holds heavy rows
        // and few light-row edges, when the range table exists and every
        // work-group queue of the frontier is complete
        const BfsCtrl& c = s_c;
        const bool solo = !s_exit && s_solo;
        const bool top = !s_exit && !solo && s_dir == 0;
        const bool queue = top && (first ? s_src0 >= 0 : c.qovf == 0) && c.n_front <= a.queue_front &&
                           c.m_front >= a.queue_min_edges && c.m_front <= a.queue_max_edges;
        const bool range = top && !queue && !first && a.range_size > 0 && c.qovf == 0 &&
                           c.n_front <= a.range_front && c.m_light <= a.range_light &&
                           c.front_hist[5] + c.front_hist[6] > 0;
        if (lane == 0) s_mode = solo ? 3 : (queue ? 2 : (range ? 1 : 0));
        if ((range || queue || solo) && !first) {
            // queue offsets: exclusive scan of the per-work-group counts,
            // four consecutive counts per lane (one wavefront scan per 256)
            int32_t carry = 0;
            for (int32_t b0 = 0; b0 < a.grid; b0 += 256) {
                int32_t cnt[4], sum = 0;
                #pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int32_t b = b0 + lane * 4 + j;
                    cnt[j] = b < a.grid ? s_qoff[b] : 0;
                    sum += cnt[j];
                }
                int32_t tot;
                int32_t ex = carry + wave_excl_scan(sum, lane, tot);
                #pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int32_t b = b0 + lane * 4 + j;
                    if (b < a.grid) s_qoff[b] = ex;
                    ex += cnt[j];
                }
                carry += tot;
            }
            if (lane == 0) s_qoff[a.grid] = carry;
        }
    }
    __syncthreads();
    if (s_exit) return;
    const int32_t p = s_p;
    const int32_t L = s_c.level;
    const int dir = s_dir;
    const int vmode = s_mode;

    if (blockIdx.x == 0 && tid == 0) {
        BfsCtrl c = s_c;
        c.phase = p;
        c.dir = dir;
        c.variant = vmode;
        if (first) c.t_main = t_now;
        if (!first && L < a.max_log)
            for (int i = 0; i < kFeat; i++) a.log[L].feat[i] = s_feat[i];
        sh->state[p & 1] = c;
    }

    // launch p reads F from bitmap p mod 3, writes the next frontier to
    // p + 1 and clears p + 2 (a solo launch runs several levels, so the
    // rotation follows the launch, not the level)
    const uint32_t* FL = bm_at(a, pmod6 % 3);
    uint32_t* Fn = bm_at(a, (pmod6 + 1) % 3);
    uint32_t* Fc = bm_at(a, (pmod6 + 2) % 3);
    int32_t* qout = a.qbuf + ((size_t)lpar * a.grid + blockIdx.x) * kQCap;
    Counters k;
    if (vmode == 3) {
        for (int64_t i = (int64_t)blockIdx.x * kBlock + tid; i < a.nwords / 2; i += (int64_t)a.grid * kBlock)
            reinterpret_cast<uint2*>(Fc)[i] = make_uint2(0u, 0u);
        if (blockIdx.x == 0)
            solo_loop<PARENTS>(a, &s_c, p, lpar, Fn, qout, s_lds, s_qoff, k, tid, lane, wid);
    } else if (dir == 1) bu_step<PARENTS>(a, L, FL, Fn, Fc, k, s_qn, qout, s_fn[wid], s_q + wid * 128,
                                  s_lists + wid * kTileVerts, gw, nw, lane, pf_fl, pf_vis, pf_nz);
    else if (vmode == 2) queue_step<PARENTS>(a, L, FL, Fn, Fc, k, s_qn, qout, s_lds, s_qoff, s_c.n_front,
                                             s_src0, lpar, tid, lane, wid);
    else if (vmode == 1) range_step<PARENTS>(a, L, FL, Fn, Fc, k, s_qn, qout, s_lds, s_qoff, s_qoff[a.grid], lpar, tid, lane, wid);
    else td_step<PARENTS>(a, s_c, L, FL, Fn, Fc, k, s_qn, qout, s_lists + wid * kTileVerts, gw, nw, lane, pf_fl);

    // block reduction -> this block's counters (plain stores)
    // A work-group's counters of one step fit in 32 bits (at most nnz < 2^31
    // adjacency entries, nv vertices).
    uint32_t r[6] = {(uint32_t)k.n, (uint32_t)k.out, (uint32_t)k.in, (uint32_t)k.scan, (uint32_t)k.cand,
                     (uint32_t)k.light};
    #pragma unroll
    for (int i = 0; i < 6; i++) r[i] = wave_sum(r[i]);
    static_assert(6 + kHistBins <= kPart, "block reduction slots");
    if (lane == 0) {
        for (int i = 0; i < 6; i++) s_red[wid][i] = r[i];
        for (int b = 0; b < kHistBins; b++) s_red[wid][6 + b] = k.hist[b];
    }
    lds_barrier();
    unsigned long long* out = a.part + ((size_t)lpar * a.grid + blockIdx.x) * kPart;
    if (tid < 6) {
        unsigned long long s = 0;
        for (int q = 0; q < kWaves; q++) s += s_red[q][tid];
        out[tid < 5 ? tid : P_LIGHT] = s;
        if (tid == P_NEXT) out[P_QOVF] = s > (unsigned long long)kQCap ? 1ull : 0ull;
    } else if (tid >= 6 && tid < 6 + kHistBins) {   // the degree histogram
        unsigned long long s = 0;
        for (int q = 0; q < kWaves; q++) s += s_red[q][tid];
        out[P_HIST + tid - 6] = s;
    }
}

// S_0 of a search from nsrc sources (level-0 counters given).
__device__ void init_state(const StepArgs& a, int32_t nsrc, int32_t src0, unsigned long long m_out,
                           unsigned long long m_in, const uint32_t* hist, int64_t nnz_scan, int32_t forced,
                           int32_t mode, float alpha, float beta) {
    BfsCtrl c;
    c.level = 0;
    c.dir = forced >= 0 ? forced : 0;
    c.n_front = nsrc;
    c.n_prev = 0;
    c.prev_dir = 0;
    c.policy_mode = mode;
    c.m_front = (int64_t)m_out;
    c.m_unvisited = nnz_scan - (int64_t)m_in;
    c.n_unvisited = (int64_t)a.nv - nsrc;
    for (int b = 0; b < kHistBins; b++) c.front_hist[b] = hist[b];
    c.alpha = alpha;
    c.beta = beta;
    c.nv = a.nv;
    c.phase = -1;
    c.nnz_scan = nnz_scan;
    c.m_light = 0;
    c.qovf = 1;              // level 0 has no work-group queues
    c.variant = 0;
    c.t_main = wall_clock64();
    a.sh->state[0] = c;
    a.sh->state[1] = c;      // what the first launch reads
    a.sh->src0 = src0;
    set_done(a, nsrc == 0 ? 0 : INT_MAX, 0);
    if (a.max_log > 0) {
        float f[kFeat];
        compute_features(&c, f);
        for (int i = 0; i < kFeat; i++) a.log[0].feat[i] = f[i];
    }
}


// Clears the search state.  With src >= 0 it also seeds that single source
// and stores S_0 (no separate seed launch).
__global__ void bfs_init(StepArgs a, int32_t src, int64_t nnz_scan, int32_t forced, int32_t mode, float alpha,
                         float beta) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // dist / parent in 16-byte stores where both arrays allow it (the
    // handle's buffers always do; a caller's buffer may not); each element
    // is written by exactly one thread, the source's with its seed value
    const bool vec = ((uintptr_t)a.dist & 15) == 0 && ((uintptr_t)a.parent & 15) == 0;
    const int64_t n4 = vec ? a.nv / 4 : 0;
    for (int64_t i = t0; i < n4; i += stride) {
        int4 d = make_int4(-1, -1, -1, -1), p = d;
        if (src >= 0 && (src >> 2) == i) {
            const int c = src & 3;
            if (c == 0) { d.x = 0; p.x = src; }
            else if (c == 1) { d.y = 0; p.y = src; }
            else if (c == 2) { d.z = 0; p.z = src; }
            else { d.w = 0; p.w = src; }
        }
        // streaming stores: the whole arrays are written once, nothing
        // here reads them back
        typedef int v4i __attribute__((ext_vector_type(4)));
        __builtin_nontemporal_store(v4i{d.x, d.y, d.z, d.w}, reinterpret_cast<v4i*>(a.dist) + i);
        if (a.parent) __builtin_nontemporal_store(v4i{p.x, p.y, p.z, p.w}, reinterpret_cast<v4i*>(a.parent) + i);
    }
    for (int64_t i = 4 * n4 + t0; i < a.nv; i += stride) {
        a.dist[i] = i == src ? 0 : -1;
        if (a.parent) a.parent[i] = i == src ? (uint32_t)src : 0xffffffffu;
    }
    // bitmaps: nwords is even and every bitmap 8-byte aligned
    for (int64_t i = t0; i < a.nwords / 2; i += stride) {
        uint2 f = make_uint2(0u, 0u);
        if (src >= 0 && (src >> 6) == i) {
            if ((src >> 5) & 1) f.y = 1u << (src & 31);
            else f.x = 1u << (src & 31);
        }
        const uint2 z = make_uint2(0u, 0u);
        reinterpret_cast<uint2*>(a.bm0)[i] = f;
        reinterpret_cast<uint2*>(a.bm1)[i] = z;
        reinterpret_cast<uint2*>(a.bm2)[i] = z;
        reinterpret_cast<uint2*>(a.vis)[i] = z;
    }
    if (src >= 0 && blockIdx.x == 0 && threadIdx.x == 0) {
        const int32_t dout = a.td_rp[src + 1] - a.td_rp[src];
        const int32_t din = a.bu_rp[src + 1] - a.bu_rp[src];
        uint32_t hist[kHistBins] = {0, 0, 0, 0, 0, 0, 0, 0};
        hist[hist_bin(dout)] = 1;
        init_state(a, 1, src, (unsigned long long)dout, (unsigned long long)din, hist, nnz_scan, forced, mode, alpha,
                   beta);
    }
}

// Marks a list of distinct sources and stores S_0.  One block.
__global__ __launch_bounds__(256) void bfs_seed(StepArgs a, const int32_t* srcs, int32_t nsrc,
                                                int64_t nnz_scan, int32_t forced, int32_t mode, float alpha,
                                                float beta) {
    __shared__ unsigned long long s_out, s_in;
    __shared__ uint32_t s_hist[kHistBins];
    const int tid = threadIdx.x;
    if (tid == 0) { s_out = 0; s_in = 0; }
    if (tid < kHistBins) s_hist[tid] = 0;
    __syncthreads();
    unsigned long long mo = 0, mi = 0;
    for (int i = tid; i < nsrc; i += blockDim.x) {
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
    if (tid == 0) init_state(a, nsrc, -1, s_out, s_in, s_hist, nnz_scan, forced, mode, alpha, beta);
}

// Level 0 of a single-source search, as launch 0.  Every entry of the
// source's row is a new vertex at distance 1 whose parent is the source
// (rows hold no duplicates; a self-loop is skipped), so no vertex needs a
// claim: plain stores of distance and parent, the F_1 bits, and what launch
// 1's prologue reads -- each work-group's queue of new vertices and its
// per-block counters (work-group b of a.grid takes the b-th share of the
// row) and the state launch 0 stores.  Where an ordinary launch sweeps every
// bitmap tile to find the source and expands its row with one wavefront,
// this takes two dependent round trips.  bfs_init has stored S_0, seeded
// the source and cleared the bitmaps and the distances and parents.
constexpr int kFirstBlock = 256;
__global__ __launch_bounds__(kFirstBlock) void bfs_first(StepArgs a, int32_t src) {
    __shared__ uint32_t s_red[kFirstBlock / 64][6 + kHistBins];
    __shared__ uint32_t s_qn[1];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int32_t nb = a.grid;
    const int32_t b0 = a.td_rp[src], deg = a.td_rp[src + 1] - b0;
    const int32_t lo = b0 + (int32_t)((int64_t)deg * blockIdx.x / nb);
    const int32_t hi = b0 + (int32_t)((int64_t)deg * (blockIdx.x + 1) / nb);
    uint32_t* Fn = a.bm1;                                      // launch 0 writes bitmap 1
    int32_t* qout = a.qbuf + (size_t)blockIdx.x * kQCap;       // launch parity 0
    if (tid == 0) s_qn[0] = 0;
    __syncthreads();
    Counters k;
    for (int32_t base = lo; base < hi; base += kFirstBlock) {  // uniform over the work-group
        const int32_t e = base + tid;
        int32_t v = e < hi ? a.td_col[e] : -1;
        if (v == src) v = -1;
        k.scan += e < hi ? 1ull : 0ull;
        int bin = 0;
        if (v >= 0) {
            a.dist[v] = 1;
            if (a.parent) a.parent[v] = (uint32_t)src;
            atomicOr(&Fn[v >> 5], 1u << (v & 31));
            bin = account_new(a, v, -1, k);
        }
        push_new(v >= 0, v, bin, k, s_qn, qout, lane);
    }
    uint32_t r[6] = {(uint32_t)k.n, (uint32_t)k.out, (uint32_t)k.in, (uint32_t)k.scan, (uint32_t)k.cand,
                     (uint32_t)k.light};
    #pragma unroll
    for (int i = 0; i < 6; i++) r[i] = wave_sum(r[i]);
    if (lane == 0) {
        #pragma unroll
        for (int i = 0; i < 6; i++) s_red[wid][i] = r[i];
        #pragma unroll
        for (int b = 0; b < kHistBins; b++) s_red[wid][6 + b] = k.hist[b];
    }
    __syncthreads();
    unsigned long long* out = a.part + (size_t)blockIdx.x * kPart;   // launch parity 0
    if (tid < 6 + kHistBins) {
        unsigned long long sum = 0;
        for (int q = 0; q < kFirstBlock / 64; q++) sum += s_red[q][tid];
        if (tid < 6) out[tid < 5 ? tid : P_LIGHT] = sum;
        else out[P_HIST + tid - 6] = sum;
        if (tid == P_NEXT) out[P_QOVF] = sum > (unsigned long long)kQCap ? 1ull : 0ull;
    }
    if (blockIdx.x == 0 && tid == 0) {
        atomicOr(&a.vis[src >> 5], 1u << (src & 31));          // F_0 folded into the visited bitmap
        BfsCtrl* c = &a.sh->state[0];    // field by field: no private copy of the struct
        c->phase = 0;
        c->dir = 0;
        c->variant = 2;
        c->t_main = wall_clock64();
    }
}

// Undoes one step: every vertex of F_{k+1} (`Fn`) returns to unvisited and
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
    dev_free(w->bm[0]); dev_free(w->bm[1]); dev_free(w->bm[2]); dev_free(w->vis); dev_free(w->qbuf);
    dev_free(w->sh); dev_free(w->log); dev_free(w->policy);
    if (w->h_sh) (void)hipHostFree(w->h_sh);
    if (w->h_done) (void)hipHostFree(w->h_done);
    if (w->ev0) (void)hipEventDestroy(w->ev0);
    if (w->ev1) (void)hipEventDestroy(w->ev1);
    for (auto e : w->step_ev) (void)hipEventDestroy(e);
    for (auto& x : w->gexec)
        if (x) (void)hipGraphExecDestroy(x);
    if (w->own_stream) (void)hipStreamDestroy(w->own_stream);
    delete w;
}

// Marks the wall vertices visited before the first level, so that no level
// claims them (has_path's exclude_vertices, src/traversals/bfs.jl:150-157).
__global__ void bfs_walls(StepArgs a, const int32_t* walls, int32_t n) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int32_t x = walls[i];
        if (x >= 0 && x < a.nv) atomicOr(&a.vis[x >> 5], 1u << (x & 31));
    }
}

// A search workspace on stream s.  With per_call false it holds no
// distance or parent array and no source list (the concurrent single-source
// searches write the caller's arrays).
static BfsWork* bfs_work_new(lgmi_graph* g, hipStream_t s, bool per_call) {
    BfsWork* w = new BfsWork();
    try {
        w->nv = g->nv;
        w->nwords = 2 * ((g->nv + 63) / 64);
        w->grid = std::min(std::max(1, g->cu_count), kMaxGrid);
        if (per_call) {
            w->dist = dev_alloc_n<int32_t>(g->nv);
            w->parent = dev_alloc_n<int32_t>(g->nv);
            w->srcbuf = dev_alloc_n<int32_t>((size_t)g->nv + 1);
        }
        w->part = dev_alloc_n<unsigned long long>((size_t)2 * w->grid * kPart);
        for (int k = 0; k < 3; k++) w->bm[k] = dev_alloc_n<uint32_t>(std::max(w->nwords, 2));
        w->vis = dev_alloc_n<uint32_t>(std::max(w->nwords, 2));
        w->qbuf = dev_alloc_n<int32_t>((size_t)2 * w->grid * kQCap);
        w->sh = dev_alloc_n<BfsShared>(1);
        w->max_log = 16384;   // covers a 4096 x 4096 grid (up to 8190 levels)
        w->log = dev_alloc_n<LevelRec>(w->max_log);
        w->policy = dev_alloc_n<PolicyWeights>(1);
        LGMI_HIP(hipHostMalloc((void**)&w->h_sh, sizeof(BfsShared), hipHostMallocDefault));
        LGMI_HIP(hipHostMalloc((void**)&w->h_done, 2 * sizeof(int32_t), hipHostMallocMapped | hipHostMallocCoherent));
        w->h_done[0] = 0;
        w->h_done[1] = 0;
        LGMI_HIP(hipHostGetDevicePointer((void**)&w->d_done, w->h_done, 0));
        LGMI_HIP(hipEventCreate(&w->ev0));
        LGMI_HIP(hipEventCreate(&w->ev1));
        LGMI_HIP(hipMemsetAsync(w->sh, 0, sizeof(BfsShared), s));
        LGMI_HIP(hipMemsetAsync(w->log, 0, sizeof(LevelRec) * w->max_log, s));
        LGMI_HIP(hipMemsetAsync(w->part, 0, sizeof(unsigned long long) * 2 * (size_t)w->grid * kPart, s));
    } catch (...) {
        bfs_free(w);
        throw;
    }
    return w;
}

void bfs_prepare(lgmi_graph* g) {
    if (g->bfs) return;
    g->bfs = bfs_work_new(g, g->stream, true);
}

// The adjacency half of the step arguments (traversal and reverse CSR with
// their derived structures) for a launch grid of `grid` work-groups.
static StepArgs graph_args(lgmi_graph* g, int dir_in, int32_t grid) {
    StepArgs a{};
    const DevCSR& td = dir_in ? g->in : g->out;
    const DevCSR& bu = dir_in ? g->out : g->in;
    a.td_rp = td.row_ptr; a.td_col = td.col;
    a.td_item_v = td.item_v; a.td_item_beg = td.item_beg; a.td_item_end = td.item_end;
    a.td_items = td.n_items;
    a.td_heavy_id = td.heavy_id;
    a.td_rtab = td.rtab;
    a.range_size = (td.range_size > 0 && td.range_grid == grid) ? td.range_size : 0;
    a.bu_rp = bu.row_ptr; a.bu_col = bu.col; a.bu_nz = bu.nz_bm; a.bu_first = bu.first;
    a.nv = g->nv;
    a.nwords = 2 * ((g->nv + 63) / 64);
    a.grid = grid;
    const lgmi_tuning t = tuning_get();
    a.wave_row = t.wave_row;
    a.lane_rounds = t.lane_rounds;
    a.range_front = std::min(t.range_front, kRangeFront);
    a.queue_front = std::min(t.queue_front, kQueueFront);
    a.range_light = t.range_light;
    a.queue_min_edges = t.queue_min_edges;
    a.queue_max_edges = t.queue_max_edges;
    a.solo_front = std::min(t.solo_front, kSoloFront);
    a.solo_edges = t.solo_edges;
    return a;
}

static StepArgs make_args(lgmi_graph* g, BfsWork* w, int dir_in, bool parents) {
    StepArgs a = graph_args(g, dir_in, w->grid);
    a.dist = w->dist;
    a.parent = parents ? reinterpret_cast<uint32_t*>(w->parent) : nullptr;
    a.bm0 = w->bm[0]; a.bm1 = w->bm[1]; a.bm2 = w->bm[2];
    a.vis = w->vis;
    a.qbuf = w->qbuf;
    a.sh = w->sh;
    a.part = w->part;
    a.grid = w->grid;
    a.log = w->log;
    a.max_log = w->max_log;
    a.policy = w->policy;
    return a;
}

static void launch_level(const StepArgs& a, bool parents, int32_t karg, int32_t forced, hipStream_t s) {
    if (parents) bfs_level<true><<<a.grid, kBlock, 0, s>>>(a, karg, forced);
    else bfs_level<false><<<a.grid, kBlock, 0, s>>>(a, karg, forced);
}

// The learned policy's weights into workspace w, on stream s, when they
// changed since its last search.
static void policy_upload(BfsWork* w, hipStream_t s) {
    std::lock_guard<std::mutex> lk(policy_mutex());
    PolicyHost& p = policy_host();
    if (p.have_mlp && w->policy_version != p.version) {
        LGMI_HIP(hipMemcpyAsync(w->policy, &p.w, sizeof(PolicyWeights), hipMemcpyHostToDevice, s));
        w->policy_version = p.version;
    }
}

// One search on workspace w and stream s.  out_dist / out_parent, when
// out_dist is not NULL, replace the workspace's own result arrays (the
// concurrent searches; captured level sequences, whose arguments hold the
// arrays, are then not used).
static void bfs_search(lgmi_graph* g, BfsWork* w, hipStream_t s, int32_t* out_dist, int32_t* out_parent,
                       const int32_t* h_srcs, int32_t nsrc, int32_t single_src, int dir_in, int want_parents,
                       int forced_dir, lgmi_bfs_stats* stats, const int32_t* h_walls, int32_t nwalls) {
    if (nsrc > g->nv) throw Error(LGMI_E_ARG, "source list longer than nv (de-duplicate first)");
    if (nwalls > 0 && single_src < 0) throw Error(LGMI_E_ARG, "walls need a single source");
    int mode;
    float alpha, beta;
    bool have_mlp;
    {
        std::lock_guard<std::mutex> lk(policy_mutex());
        PolicyHost& p = policy_host();
        mode = p.mode; alpha = p.alpha; beta = p.beta; have_mlp = p.have_mlp;
    }
    if (mode == LGMI_POLICY_LEARNED && have_mlp) policy_upload(w, s);
    if (mode == LGMI_POLICY_LEARNED && !have_mlp) mode = LGMI_POLICY_HEURISTIC;

    StepArgs a = make_args(g, w, dir_in, want_parents != 0);
    if (out_dist) {
        a.dist = out_dist;
        a.parent = want_parents ? reinterpret_cast<uint32_t*>(out_parent) : nullptr;
    }
    if (mode != LGMI_POLICY_LEARNED) a.policy = nullptr;
    a.done_host = w->d_done;
    const int64_t nnz_scan = g->nnz;  // the reverse CSR has as many entries as the forward one
    const int init_grid = std::max(1, std::min<int>(w->grid * 8, (g->nv + 255) / 256));
    static const bool graphs_env = [] {
        const char* e = std::getenv("LGMI_BFS_GRAPHS");
        return !(e && e[0] == '0');
    }();
    const bool graphs_on = graphs_env && out_dist == nullptr;
    const bool parents = want_parents != 0;
    const int flags = (parents ? 1 : 0) | (dir_in ? 2 : 0) | (a.policy ? 4 : 0) | ((forced_dir + 1) << 3);
    // Launches go out in batches; a launch after the last level exits at
    // once.  The first batch is sized by the last search with a margin of
    // one level: an exiting launch costs less than a second round trip to
    // the host when a search is one level deeper than the last.
    int batch = std::max(4, std::min(w->last_launches_used + 2, 64));
    int launched = 0;
    // Completion.  The launch that ends the search stores its index (and the
    // level count) into host-mapped memory.  Without statistics or per-step
    // timing the host polls that word and returns as soon as the search has
    // ended: the batch's spare launches, which exit at once, then finish
    // behind the caller's next work on the same stream, and every later read
    // of the results is ordered after them on that stream.  With either, the
    // host waits for the whole batch.  The word is reset here, before
    // anything is enqueued; the kernels of the last search wrote it before
    // that search returned.
    const bool poll = stats == nullptr && !w->time_steps;
    __atomic_store_n(reinterpret_cast<unsigned long long*>(w->h_done), (unsigned long long)INT_MAX, __ATOMIC_RELEASE);
    auto read_done = [&](int32_t& lev) {
        const unsigned long long x = __atomic_load_n(reinterpret_cast<unsigned long long*>(w->h_done), __ATOMIC_ACQUIRE);
        lev = (int32_t)(x >> 32);
        return (int32_t)(uint32_t)x;
    };
    int32_t levels = 0;
    LGMI_HIP(hipEventRecord(w->ev0, s));
    if (single_src < 0 && nsrc > 0)
        LGMI_HIP(hipMemcpyAsync(w->srcbuf, h_srcs, sizeof(int32_t) * (size_t)nsrc, hipMemcpyHostToDevice, s));
    bfs_init<<<init_grid, 256, 0, s>>>(a, single_src, nnz_scan, forced_dir, mode, alpha, beta);
    if (single_src < 0)
        bfs_seed<<<1, 256, 0, s>>>(a, w->srcbuf, nsrc, nnz_scan, forced_dir, mode, alpha, beta);
    if (nwalls > 0) {
        LGMI_HIP(hipMemcpyAsync(w->srcbuf, h_walls, sizeof(int32_t) * (size_t)nwalls, hipMemcpyHostToDevice, s));
        bfs_walls<<<std::max(1, std::min(w->grid, (nwalls + 255) / 256)), 256, 0, s>>>(a, w->srcbuf, nwalls);
    }
    LGMI_HIP(hipGetLastError());
    auto event_at = [&](size_t i) {
        while (w->step_ev.size() <= i) {
            hipEvent_t e;
            LGMI_HIP(hipEventCreate(&e));
            w->step_ev.push_back(e);
        }
        LGMI_HIP(hipEventRecord(w->step_ev[i], s));
    };
    // Level 0 of a single source is its own small kernel (launch 0) when the
    // policy decides the directions; a forced direction (the direction
    // probe's labels, the tests that force one) and has_path's excluded
    // vertices keep the ordinary launch.
    if (single_src >= 0 && forced_dir < 0 && nwalls == 0) {
        if (w->time_steps) event_at(0);
        bfs_first<<<a.grid, kFirstBlock, 0, s>>>(a, single_src);
        if (w->time_steps) event_at(1);
        launched = 1;
    }
    int32_t done = 0;
    for (bool first_batch = true;; first_batch = false) {
        if (!first_batch) {
            LGMI_HIP(hipEventRecord(w->ev1, s));
            if (poll) {
                // spin for up to 0.5 ms (a whole RMAT-22 search takes about
                // 0.16 ms), then wait for the batch without spinning
                const auto t0 = std::chrono::steady_clock::now();
                for (unsigned it = 0;; it++) {
                    done = read_done(levels);
                    if (done < launched) break;
                    if ((it & 255) == 255) {   // the batch may end without ending the search
                        if (std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(500)) {
                            LGMI_HIP(hipEventSynchronize(w->ev1));
                            done = read_done(levels);
                            break;
                        }
                        const hipError_t q = hipEventQuery(w->ev1);
                        if (q == hipSuccess) {
                            done = read_done(levels);
                            break;
                        }
                        if (q != hipErrorNotReady) LGMI_HIP(q);
                    }
                    __builtin_ia32_pause();
                }
            } else {
                LGMI_HIP(hipEventSynchronize(w->ev1));
                done = read_done(levels);
            }
            if (done < launched) break;
            if (launched > g->nv + 8) throw Error(LGMI_E_HIP, "BFS did not terminate");
            batch = std::min(batch * 2, 1024);
        }
        if (graphs_on && !w->time_steps && launched >= kGraphLevels && launched % 6 == 0) {
            // deep search: replay captured level sequences (each starts on a
            // level that is a multiple of 6)
            if (!w->gexec[flags]) {
                hipGraph_t gr = nullptr;
                LGMI_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                for (int i = 0; i < kGraphLevels; i++) launch_level(a, parents, -1 - (i % 6), forced_dir, s);
                LGMI_HIP(hipStreamEndCapture(s, &gr));
                hipGraphExec_t ex = nullptr;
                const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                LGMI_HIP(e);
                w->gexec[flags] = ex;
            }
            const int ng = std::max(1, batch / kGraphLevels);
            for (int i = 0; i < ng; i++) LGMI_HIP(hipGraphLaunch(w->gexec[flags], s));
            launched += ng * kGraphLevels;
        } else {
            // individual launches; a deep search stops at the next multiple
            // of 6 from which graph replays can start
            // the first batch counts the level-0 kernel among its launches
            int n = first_batch ? std::max(1, batch - launched) : batch;
            if (graphs_on && !w->time_steps && launched + n > kGraphLevels)
                n = std::max(1, (std::max(launched, kGraphLevels) + 5) / 6 * 6 - launched);
            for (int i = 0; i < n; i++) {
                const size_t li = (size_t)(launched + i);
                if (w->time_steps) event_at(2 * li);
                launch_level(a, parents, launched + i, forced_dir, s);
                if (w->time_steps) event_at(2 * li + 1);
            }
            launched += n;
        }
        LGMI_HIP(hipGetLastError());
    }
    // done: the launch from which on launches exit; levels: the level count
    // (equal unless solo launches ran several levels each)
    w->last_levels = levels;
    w->last_launches_used = done;
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
        stats->levels = levels;
        stats->device_ms = ms;
        stats->launches = launched;
        stats->steps_event_ms = w->last_steps_event_ms;
        const int nrec = std::min(levels, w->max_log);
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

void bfs_run(lgmi_graph* g, const int32_t* h_srcs, int32_t nsrc, int32_t single_src, int dir_in,
             int want_parents, int forced_dir, lgmi_bfs_stats* stats, const int32_t* h_walls, int32_t nwalls) {
    bfs_prepare(g);
    bfs_search(g, g->bfs, g->stream, nullptr, nullptr, h_srcs, nsrc, single_src, dir_in, want_parents, forced_dir,
               stats, h_walls, nwalls);
}

// Independent single-source searches from roots[0..nroots), up to
// kManyStreams at a time, each on a workspace and stream of its own driven by
// a host thread of its own; a thread takes the next root when its search
// ends.  Row i of dist / parent (device arrays of nroots x nv) receives the
// results of roots[i].  A level launch of one search fills the gaps that
// another search's launch boundaries and small levels leave.  RMAT-22,
// 64 roots (GTEPS, medians of four runs): 2 at a time 468, 4 475, 8 485,
// 12 481, 16 476, against 403 for one search per call.
constexpr int kManyStreams = 8;

void bfs_run_many(lgmi_graph* g, const int32_t* roots, int32_t nroots, int dir_in, int32_t* dist, int32_t* parent) {
    if (nroots <= 0) return;
    const int K = std::min<int>(kManyStreams, nroots);
    while ((int)g->bfs_pool.size() < K) {
        hipStream_t st = nullptr;
        LGMI_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        BfsWork* w = nullptr;
        try {
            w = bfs_work_new(g, st, false);
        } catch (...) {
            (void)hipStreamDestroy(st);
            throw;
        }
        w->own_stream = st;
        g->bfs_pool.push_back(w);
    }
    // the searches start after the handle's queued work (the CSR upload and
    // any earlier search) is complete
    hipEvent_t ready = nullptr;
    LGMI_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    try {
        LGMI_HIP(hipEventRecord(ready, g->stream));
        for (int k = 0; k < K; k++) LGMI_HIP(hipStreamWaitEvent(g->bfs_pool[(size_t)k]->own_stream, ready, 0));
    } catch (...) {
        (void)hipEventDestroy(ready);
        throw;
    }
    int dev = 0;
    LGMI_HIP(hipGetDevice(&dev));
    const int forced = g->bfs ? g->bfs_forced : -1;
    const int want_parents = parent != nullptr;
    const size_t nv = (size_t)g->nv;
    std::atomic<int32_t> next{0};
    std::atomic<bool> failed{false};
    std::vector<std::exception_ptr> err((size_t)K);
    auto worker = [&](int k) {
        BfsWork* w = g->bfs_pool[(size_t)k];
        try {
            LGMI_HIP(hipSetDevice(dev));
            for (;;) {
                const int32_t i = next.fetch_add(1);
                if (i >= nroots || failed.load()) break;
                bfs_search(g, w, w->own_stream, dist + (size_t)i * nv, want_parents ? parent + (size_t)i * nv : nullptr,
                           nullptr, 1, roots[i], dir_in, want_parents, forced, nullptr, nullptr, 0);
            }
        } catch (...) {
            err[(size_t)k] = std::current_exception();
            failed.store(true);
        }
        (void)hipStreamSynchronize(w->own_stream);
    };
    std::vector<std::thread> th;
    th.reserve((size_t)K - 1);
    try {
        for (int k = 1; k < K; k++) th.emplace_back(worker, k);
    } catch (...) {   // no thread for the rest: the started ones finish their searches
        failed.store(true);
        for (auto& t : th) t.join();
        (void)hipEventDestroy(ready);
        throw;
    }
    worker(0);
    for (auto& t : th) t.join();
    (void)hipEventDestroy(ready);
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
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

// ----------------------------------------------------------------------
// BFS without device allocation: the search state lives in one fixed
// device array (`scratch`), the results go to caller-owned device buffers.
// ----------------------------------------------------------------------
namespace {

constexpr size_t na_align(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr int kNaMaxGrid = 256;                                   // work-groups at most
constexpr size_t kNaWords = 2 * (((size_t)LGMI_NOALLOC_MAX_NV + 63) / 64);   // bitmap words at the largest nv
constexpr size_t kNaPartOff = 0;                                  // [2][grid][kPart] u64
constexpr size_t kNaBmOff = na_align(kNaPartOff + sizeof(unsigned long long) * 2 * kNaMaxGrid * kPart);
constexpr size_t kNaQOff = na_align(kNaBmOff + sizeof(uint32_t) * 4 * kNaWords);   // bm0, bm1, bm2, vis
constexpr size_t kNaShOff = na_align(kNaQOff + sizeof(int32_t) * 2 * kNaMaxGrid * kQCap);
constexpr size_t kNaPolOff = na_align(kNaShOff + sizeof(BfsShared));
constexpr size_t kNaBytes = na_align(kNaPolOff + sizeof(PolicyWeights));
static_assert(kNaBytes <= (size_t)LGMI_NOALLOC_SCRATCH_BYTES, "no-alloc scratch layout exceeds the stated size");

// Each GPU holds its own copy until the process ends; host code reaches it
// only through the address hipGetSymbolAddress returns for that device.
static __device__ __align__(256) unsigned char scratch[LGMI_NOALLOC_SCRATCH_BYTES];

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu[kNaMaxDevices];         // one no-alloc search per device at a time
unsigned char* g_na_ptr[kNaMaxDevices] = {};

// The device address of `scratch` on device dev, looked up once per device
// with that device current (the caller holds g_na_mu[dev]).
unsigned char* na_scratch(int dev) {
    if (!g_na_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            throw Error(LGMI_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        }
        g_na_ptr[dev] = static_cast<unsigned char*>(p);
    }
    return g_na_ptr[dev];
}

}  // namespace

void bfs_run_noalloc(lgmi_graph* g, int32_t src, int dir_in, int32_t* dist, int32_t* parent) {
    if (g->nv > LGMI_NOALLOC_MAX_NV)
        throw Error(LGMI_E_TOO_LARGE, "lgmi_bfs_run_noalloc: nv = " + std::to_string(g->nv) +
                                          " exceeds LGMI_NOALLOC_MAX_NV = " + std::to_string(LGMI_NOALLOC_MAX_NV));
    int dev = 0;
    LGMI_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNaMaxDevices) throw Error(LGMI_E_ARG, "device index beyond the no-alloc table");
    std::lock_guard<std::mutex> lk(g_na_mu[dev]);
    unsigned char* base = na_scratch(dev);
    const int32_t grid = std::min(std::max(1, g->cu_count), kNaMaxGrid);
    int mode;
    float alpha, beta;
    bool have_mlp;
    PolicyWeights pw;
    {
        std::lock_guard<std::mutex> pl(policy_mutex());
        PolicyHost& p = policy_host();
        mode = p.mode; alpha = p.alpha; beta = p.beta; have_mlp = p.have_mlp;
        if (have_mlp) pw = p.w;
    }
    if (mode == LGMI_POLICY_LEARNED && !have_mlp) mode = LGMI_POLICY_HEURISTIC;
    StepArgs a = graph_args(g, dir_in, grid);
    a.dist = dist;
    a.parent = reinterpret_cast<uint32_t*>(parent);
    a.part = reinterpret_cast<unsigned long long*>(base + kNaPartOff);
    uint32_t* bm = reinterpret_cast<uint32_t*>(base + kNaBmOff);
    a.bm0 = bm;
    a.bm1 = bm + kNaWords;
    a.bm2 = bm + 2 * kNaWords;
    a.vis = bm + 3 * kNaWords;
    a.qbuf = reinterpret_cast<int32_t*>(base + kNaQOff);
    a.sh = reinterpret_cast<BfsShared*>(base + kNaShOff);
    a.log = nullptr;
    a.max_log = 0;
    PolicyWeights* dpol = reinterpret_cast<PolicyWeights*>(base + kNaPolOff);
    a.policy = mode == LGMI_POLICY_LEARNED ? dpol : nullptr;
    hipStream_t s = g->stream;
    if (a.policy) LGMI_HIP(hipMemcpyAsync(dpol, &pw, sizeof(PolicyWeights), hipMemcpyHostToDevice, s));
    const int forced = g->bfs ? g->bfs_forced : -1;
    const bool parents = parent != nullptr;
    // bfs_init writes every word a search reads before writing it: the
    // bitmaps, dist / parent and the level state (the first level launch
    // reads no per-block counters)
    const int init_grid = std::max(1, std::min<int>(grid * 8, (g->nv + 255) / 256));
    bfs_init<<<init_grid, 256, 0, s>>>(a, src, g->nnz, forced, mode, alpha, beta);
    LGMI_HIP(hipGetLastError());
    int32_t launched = 0, batch = 8;
    for (;;) {
        for (int i = 0; i < batch; i++) launch_level(a, parents, launched + i, forced, s);
        launched += batch;
        LGMI_HIP(hipGetLastError());
        int32_t done = 0;
        LGMI_HIP(hipMemcpyAsync(&done, &a.sh->done_level, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        LGMI_HIP(hipStreamSynchronize(s));
        if (done < launched) break;
        if (launched > g->nv + 8) throw Error(LGMI_E_HIP, "BFS did not terminate");
        batch = std::min(batch * 2, 1024);
    }
}

}  // namespace lgmi

extern "C" int lgmi_bfs_run_noalloc(lgmi_graph_t g, int32_t src, int dir_in, int32_t* dist, int32_t* parent) {
    using namespace lgmi;
    try {
        if (!g || !dist) throw Error(LGMI_E_ARG, "bad arguments");
        std::lock_guard<std::mutex> lk(g->mu);
        if (src < 0 || src >= g->nv) throw Error(LGMI_E_BOUNDS, "source vertex " + std::to_string(src) + " out of range");
        bfs_run_noalloc(g, src, dir_in ? 1 : 0, dist, parent);
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
        LGMI_HIP(hipStreamSynchronize(g->stream));   // the search may have returned before its last launches ended
        std::vector<LevelRec> recs((size_t)n);
        LGMI_HIP(hipMemcpy(recs.data(), w->log, sizeof(LevelRec) * (size_t)n, hipMemcpyDeviceToHost));
        int dev = 0, khz = 0;
        LGMI_HIP(hipGetDevice(&dev));
        LGMI_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        const double us_per_tick = khz > 0 ? 1000.0 / (double)khz : 0.01;
        // levels per launch: a launch's event time is shared equally among them
        std::vector<int> per_launch(w->last_launch_ms.size(), 0);
        for (int i = 0; i < n; i++) {
            const int li = recs[(size_t)i].launch;
            if (li >= 0 && (size_t)li < per_launch.size()) per_launch[(size_t)li]++;
        }
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
            o.variant = r.variant;
            o.scanned = r.scanned;
            o.cand = r.cand;
            o.kernel_us = (float)((double)(r.t_end - r.t_start) * us_per_tick);
            o.step_us = (float)((double)(r.t_end - r.t_prev_end) * us_per_tick);
            const int li = r.launch;
            o.event_ms = (li >= 0 && (size_t)li < per_launch.size() && per_launch[(size_t)li] > 0)
                             ? w->last_launch_ms[(size_t)li] / (float)per_launch[(size_t)li]
                             : 0.f;
            o.launch = li;
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
        StepArgs a = make_args(g, w, dir_in, true);
        a.policy = nullptr;
        a.solo_front = 0;   // one level per launch: launch index = level
        hipStream_t s = g->stream;
        const int init_grid = std::max(1, std::min<int>(w->grid * 8, (g->nv + 255) / 256));
        bfs_init<<<init_grid, 256, 0, s>>>(a, src, g->nnz, -1, LGMI_POLICY_HEURISTIC, 15.0f, 18.0f);
        LGMI_HIP(hipGetLastError());
        hipEvent_t ev[2];
        for (auto& e : ev) LGMI_HIP(hipEventCreate(&e));
        auto sync_sh = [&]() {
            LGMI_HIP(hipMemcpyAsync(w->h_sh, w->sh, sizeof(BfsShared), hipMemcpyDeviceToHost, s));
            LGMI_HIP(hipStreamSynchronize(s));
        };
        int32_t launched = 0;   // level launches so far: the next one is level `launched`
        auto finished = [&]() { return w->h_sh->done_level < launched; };
        int nout = 0;
        try {
            sync_sh();
            for (int32_t li = 0; li < nlevels; li++) {
                const int32_t target = levels[li];
                while (!finished() && launched < target) {
                    const int n = std::min(target - launched, 256);
                    for (int i = 0; i < n; i++) launch_level(a, true, launched + i, -1, s);
                    launched += n;
                    LGMI_HIP(hipGetLastError());
                    sync_sh();
                }
                if (finished()) break;
                // Level `target` is next.  A launch of level L reads only what
                // level L-1 wrote, so it can run again on the same inputs after
                // the new vertices are undone.  Each direction runs kReps
                // times, in alternating order; the minimum is kept, so one
                // slow launch (a cold start after the host synchronised) does
                // not decide the label.  The launch finds the frontier empty
                // when the search ends at this level.
                const int32_t L = target;
                uint32_t* Fn = w->bm[(L + 1) % 3];
                constexpr int kReps = 3;
                float best[2] = {1e30f, 1e30f};
                const int first = ((L + src) & 1) ? 1 : 0;
                bool ended = false;
                for (int rep = 0; rep < kReps && !ended; rep++) {
                    for (int q = 0; q < 2; q++) {
                        const int dir = (first + q + rep) & 1;
                        LGMI_HIP(hipEventRecord(ev[0], s));
                        launch_level(a, true, L, dir, s);
                        LGMI_HIP(hipEventRecord(ev[1], s));
                        LGMI_HIP(hipGetLastError());
                        sync_sh();
                        if (w->h_sh->done_level <= L) { ended = true; break; }
                        float t = 0.f;
                        LGMI_HIP(hipEventElapsedTime(&t, ev[0], ev[1]));
                        best[dir] = std::min(best[dir], t);
                        if (!(rep == kReps - 1 && q == 1)) {
                            bfs_undo<<<init_grid, 256, 0, s>>>(a, Fn);
                            LGMI_HIP(hipGetLastError());
                        }
                    }
                }
                launched = L + 1;
                if (ended) break;
                LevelRec rec;
                LGMI_HIP(hipMemcpy(&rec, w->log + std::min(L, w->max_log - 1), sizeof(LevelRec), hipMemcpyDeviceToHost));
                lgmi_probe_rec& o = out[nout++];
                std::memset(&o, 0, sizeof(o));
                for (int q = 0; q < kFeat; q++) o.feat[q] = rec.feat[q];
                o.td_us = 1000.f * best[0];
                o.bu_us = 1000.f * best[1];
                o.level = L;
                o.n_front = w->h_sh->state[L & 1].n_front;
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
