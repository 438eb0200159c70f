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
// lanes are the 64 sources; sig[v][0..63] holds sigma by lane.
// The gathers read a neighbour's values for the lanes of one level only, so
// those are kept a second time, rank-packed: stage[L & 1][v][r] holds the
// value of the r-th lane (in lane order) of F_L[v].  The lanes one hit needs
// then sit in the first popcount(F_L[v]) entries of a 512-byte row instead
// of spread over the whole row (about 1.6 instead of 2.5 lines per hit on
// RMAT graphs with hub-ordered sources).  Two planes suffice: a step reads
// the level before (forward) or after (backward) and writes its own.
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
#include <atomic>
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

// Barrier and result words of the persistent launch.  arrive and closed are
// cleared before every launch, bar_count is zero between launches; bar_gen
// (the last barrier generation of the last complete launch) only grows.
// abort is set when a barrier wait times out (the blocks of a launch were
// not all resident), overflow when a slot needs more levels than F holds;
// either makes the host redo the computation level by level.
// Each word that blocks update concurrently has a 128-byte line of its own:
// measured on gfx950, stores into a line that other work-groups update with
// atomics can lose those updates.
struct BcPersistCtl {
    alignas(128) uint32_t arrive;   // blocks that have started (reset before each launch)
    alignas(128) uint32_t closed;   // participants + 1 once fixed, 0 before
    alignas(128) uint32_t bar_count;
    alignas(128) uint32_t bar_gen;
    alignas(128) int32_t abort;
    alignas(128) int32_t overflow;  // the rest: written by block 0 only, read by the host
    int32_t levels;
    int32_t barriers;
    uint64_t t_fwd;
    uint64_t t_bwd;
    uint64_t t_end;
};

struct BcWork {
    int32_t nv = 0;
    int32_t nslots = 0;
    int32_t cap_levels = 0;
    uint64_t* seen = nullptr;    // [slot][nv]
    uint64_t* F = nullptr;       // [slot][cap][nv]
    double* sig = nullptr;       // [slot][nv][64]
    double* stage = nullptr;     // [slot][2][nv][64] rank-packed values of one level
    double* bc_acc = nullptr;    // [slot][nv]
    double* bc_out = nullptr;    // [nv]
    double* endp = nullptr;      // [nv] endpoints term
    uint64_t* active = nullptr;  // [slot]
    int32_t* lane_src = nullptr; // [slot][64]
    BcCtrl* ctrl = nullptr;      // [slot]
    BcCtrl* h_ctrl = nullptr;    // pinned mirror
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool has_sig = false;         // sig and bc_acc allocated (not needed by distances only)
    uint64_t* st = nullptr;       // distances only: [slot][chunks + heavy rows][3][64]
    uint64_t* st_out = nullptr;   // [slot][3][64]
    size_t st_words = 0;
    double* hpart = nullptr;      // [slot][seg_stride][64]
    uint64_t* hgot = nullptr;     // [slot][seg_stride]
    int32_t seg_stride = 0;
    // persistent launch
    BcPersistCtl* pctl = nullptr;    // barrier and result words
    BcPersistCtl* h_pctl = nullptr;  // pinned mirror
    uint64_t* found = nullptr;              // [2][grid][words] per-block slot masks
    uint32_t* rel = nullptr;                // [grid][32] per-block barrier flags
    hipGraphExec_t pexec = nullptr;         // captured launch, replayed while its key matches
    std::vector<unsigned char> pkey;
};

// Per-graph host row pointers and the heavy-row segment lists (device).
struct BcGraph {
    std::vector<int32_t> rp_out, rp_in;
    int32_t* dev = nullptr;           // all segment arrays, one allocation
    int32_t nseg_in = 0, nh_in = 0, nseg_out = 0, nh_out = 0;
    int32_t* seg_in[6] = {};          // seg_v, seg_b, seg_e, seg_h, h_first, h_nseg
    int32_t* seg_out[6] = {};
    int32_t seg_len = 1024;           // segment length the arrays were cut with
};

void bc_graph_free(BcGraph* b) {
    if (!b) return;
    dev_free(b->dev);
    delete b;
}

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

// Rows longer than the segment length (lgmi_tuning.bc_seg_len, fixed for a
// graph when its segments are built) are cut into segments of that many
// entries, each a work item of its own; after all segments of a level step,
// one wavefront per row adds the segments' partial sums in segment order.

struct BcSegs {
    const int32_t* seg_v;    // [nseg] the row's vertex
    const int32_t* seg_b;    // [nseg] first entry
    const int32_t* seg_e;    // [nseg] end entry
    const int32_t* seg_h;    // [nseg] heavy-row index
    const int32_t* h_first;  // [nh] first segment of the row
    const int32_t* h_nseg;   // [nh] segments of the row
    int32_t nseg;
    int32_t nh;
};

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
    double* stage;           // [slot][2][nv][64]: plane L & 1 holds level L rank-packed
    double* bc_acc;
    const uint64_t* active;
    BcCtrl* ctrl;
    BcSegs hin;              // heavy in-rows (forward)
    BcSegs hout;             // heavy out-rows (backward)
    double* hpart;           // [slot][seg_stride][64] per-segment lane sums
    uint64_t* hgot;          // [slot][seg_stride] per-segment new-lane masks
    int32_t seg_stride;
    int32_t heavy;           // rows longer than this are heavy (the segment length)
    int32_t fmod;            // levels F keeps: cap (all), or 2 (distances only)
    int32_t fstride;         // slots interleaved in F: F[level][v][slot]
    uint64_t* st;            // distances only: [slot][chunk or heavy row][3][64]
                             // per-lane sums of distances, counts, last level
};

// Three computations share these kernels: MODE 0 is betweenness (FP64 sigma
// and delta), MODE 1 is stress centrality (src/centrality/stress.jl:33-77),
// whose forward pass counts each vertex's predecessors per lane and whose
// backward pass sums Int64 path counts (wrapping like Julia's Int), and
// MODE 2 is the forward pass alone, which records per source lane the sum
// of the distances, the number of vertices reached and the last level
// (closeness, radiality and eccentricity, src/centrality/closeness.jl,
// src/centrality/radiality.jl, src/distance.jl).  MODE 2 keeps two levels
// of F (fmod = 2) and no sigma.
template <int MODE> struct BwdAcc { using T = double; };
template <> struct BwdAcc<1> { using T = uint64_t; };

// Position of this lane among the set bits of the wave-uniform mask m below
// it (the rank-packed index of the lane's entry).
__device__ inline int lane_rank(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ inline uint64_t readlane_u64(uint64_t x, int t) {
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(x >> 32), t) << 32) |
           (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, t);
}

// Up to eight of the hits in `hits` (lanes whose adjacency entry uk[t] has
// a non-empty lane mask mk[t] = fk[t] & wanted lanes, fk[t] being the
// neighbour's lanes on the level read): lane i adds, for every hit whose
// mask has bit i, the neighbour's value for lane i -- entry rank(i in fk[t])
// of its rank-packed row in `stage` -- or 1 when LOAD is false, one after
// another in ascending hit order (the order of the row), with the eight
// loads in flight together.  Removes the hits it took and ORs their masks
// into `got`.  Every lane calls it with the same `hits`.
template <typename T, bool LOAD>
__device__ inline void gather_hits(uint64_t& hits, int32_t uk, uint64_t mk, uint64_t fk, const T* stage, int lane,
                                   T& acc, uint64_t& got) {
    constexpr int kH = 8;
    int32_t uu[kH];
    uint64_t mm[kH], ff[kH];
    #pragma unroll
    for (int q = 0; q < kH; q++) {
        uu[q] = 0;
        mm[q] = 0;
        ff[q] = 0;
        if (hits) {
            // t is the same in every lane, so the reads are scalar
            const int t = __builtin_amdgcn_readfirstlane(__ffsll((unsigned long long)hits) - 1);
            hits &= hits - 1;
            uu[q] = __builtin_amdgcn_readlane(uk, t);
            mm[q] = readlane_u64(mk, t);
            if constexpr (LOAD) ff[q] = readlane_u64(fk, t);
        }
    }
    T x[kH];
    #pragma unroll
    for (int q = 0; q < kH; q++) {
        if constexpr (LOAD) x[q] = ((mm[q] >> lane) & 1) ? stage[(size_t)uu[q] * 64 + lane_rank(ff[q])] : T(0);
        else x[q] = ((mm[q] >> lane) & 1) ? T(1) : T(0);
    }
    #pragma unroll
    for (int q = 0; q < kH; q++) {
        acc += x[q];
        got |= mm[q];
    }
}

// As gather_hits over a pass that holds the entries of several rows, entry
// t belonging to the row of chunk lane rk[t] (rows are contiguous and in
// order).  The running sum acc and mask got belong to row `cur`; when a hit
// of another row comes, flush(cur, acc, got) finishes that row and both
// restart from zero, so every row adds its values in the order of its
// entries, as one row at a time does.
template <typename T, bool LOAD, typename Flush>
__device__ inline void gather_rows(uint64_t& hits, int32_t uk, uint64_t mk, uint64_t fk, int32_t rk, const T* stage,
                                   int lane, T& acc, uint64_t& got, int32_t& cur, Flush& flush) {
    constexpr int kH = 8;
    int32_t uu[kH], rr[kH];
    uint64_t mm[kH], ff[kH];
    #pragma unroll
    for (int q = 0; q < kH; q++) {
        uu[q] = 0;
        rr[q] = -1;
        mm[q] = 0;
        ff[q] = 0;
        if (hits) {
            const int t = __builtin_amdgcn_readfirstlane(__ffsll((unsigned long long)hits) - 1);
            hits &= hits - 1;
            uu[q] = __builtin_amdgcn_readlane(uk, t);
            rr[q] = __builtin_amdgcn_readlane(rk, t);
            mm[q] = readlane_u64(mk, t);
            if constexpr (LOAD) ff[q] = readlane_u64(fk, t);
        }
    }
    T x[kH];
    #pragma unroll
    for (int q = 0; q < kH; q++) {
        if constexpr (LOAD) x[q] = ((mm[q] >> lane) & 1) ? stage[(size_t)uu[q] * 64 + lane_rank(ff[q])] : T(0);
        else x[q] = ((mm[q] >> lane) & 1) ? T(1) : T(0);
    }
    #pragma unroll
    for (int q = 0; q < kH; q++) {
        if (rr[q] >= 0) {
            if (rr[q] != cur) {
                if (cur >= 0) flush(cur, acc, got);
                cur = rr[q];
                acc = T(0);
                got = 0;
            }
            acc += x[q];
            got |= mm[q];
        }
    }
}

// One 64-entry pass over the packed short rows of a chunk (row j's entries
// end at entry E_j of the packed list, start at S_j = E_j - len_j): the
// whole rows that fit from entry `base` on, each lane's entry, its row and
// that row's lane mask `key`, and the lane's column entry, whose load is
// issued here so that it runs while the previous pass gathers.  Every lane
// calls it.  pend < 0: no row starts at base (not reached, as the next row
// starts at base and holds at most 64 entries).
struct PackedPass {
    uint64_t rows;   // the rows of the pass (chunk lanes)
    uint64_t key;    // this lane's row's lane mask
    int32_t pend;    // where the pass ends in the packed list
    int32_t rk;      // this lane's row
    int32_t uk;      // this lane's column entry (lanes past the pass: 0)
    bool on;         // this lane holds an entry
};

__device__ inline PackedPass packed_pass(int32_t base, int32_t E, int32_t lenj, int32_t bj, uint64_t keyj,
                                         const int32_t* col, int lane) {
    PackedPass ps;
    ps.rows = __ballot(lenj > 0 && E > base && E <= base + 64);
    ps.pend = -1;
    ps.key = 0;
    ps.rk = 0;
    ps.uk = 0;
    ps.on = false;
    if (!ps.rows) return ps;
    ps.pend = __builtin_amdgcn_readlane(E, 63 - __builtin_clzll(ps.rows));
    // lanes past the pass repeat its last entry: every lane takes part in the shuffles
    const int32_t kk = min(base + lane, ps.pend - 1);
    ps.rk = first_lane_above(E, kk);
    const int32_t idx = __shfl(bj, ps.rk) + (kk - __shfl(E - lenj, ps.rk));
    ps.key = __shfl(keyj, ps.rk);
    ps.on = base + lane < ps.pend;
    if (ps.on) ps.uk = col[idx];
    return ps;
}

// For source i: its out-neighbour of largest out-degree (the smallest id
// among ties), -1 for an empty row.  One wavefront per source.
__global__ void bc_source_hubs(const int32_t* rp, const int32_t* col, const int32_t* srcs, int32_t n, int32_t* hub) {
    const int lane = threadIdx.x & 63;
    const int32_t i = (int32_t)blockIdx.x * (int32_t)(blockDim.x / 64) + (int32_t)(threadIdx.x >> 6);
    if (i >= n) return;   // uniform over the wavefront
    const int32_t s = srcs[i];
    const int32_t b = rp[s], e = rp[s + 1];
    long long best = -1;   // (degree << 32) | (0x7fffffff - id)
    for (int32_t k = b + lane; k < e; k += 64) {
        const int32_t u = col[k];
        const long long key = ((long long)(rp[u + 1] - rp[u]) << 32) | (long long)(0x7fffffff - u);
        best = key > best ? key : best;
    }
    for (int o = 32; o >= 1; o >>= 1) {
        const long long x = __shfl_xor(best, o);
        best = x > best ? x : best;
    }
    if (lane == 0) hub[i] = best < 0 ? -1 : 0x7fffffff - (int32_t)(best & 0xffffffffll);
}

__global__ void bc_clear_acc(double* acc, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        acc[i] = 0.0;
}

// Clears seen and F_0 of every slot; one grid over nv, blockIdx.y = slot.
__global__ void bc_init(BcArgs a) {
    const int slot = blockIdx.y;
    uint64_t* seen = a.seen + (size_t)slot * a.nv;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < a.nv; v += (int64_t)gridDim.x * blockDim.x) {
        seen[v] = 0;
        a.F[(size_t)v * a.fstride + slot] = 0;   // level 0 of F[level][v][slot]
    }
}

// Marks the sources of every slot (one block per slot, lane = source lane).
__global__ void bc_seed(BcArgs a, const int32_t* lane_src) {
    const int slot = blockIdx.x;
    const int lane = threadIdx.x;
    const int32_t s = lane_src[slot * 64 + lane];
    uint64_t* seen = a.seen + (size_t)slot * a.nv;
    double* sig = a.sig + (size_t)slot * a.nv * 64;
    if (s >= 0) {
        atomicOr((unsigned long long*)&a.F[(size_t)s * a.fstride + slot], 1ull << lane);
        atomicOr((unsigned long long*)&seen[s], 1ull << lane);
        if (a.sig) sig[(size_t)s * 64 + lane] = 1.0;
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

// Forward step L -> L+1 of one slot over the 64 vertices of chunk ch (one
// wavefront): every vertex with an unseen lane ORs its in-neighbours'
// frontier masks and adds their sigma per lane.  Returns, in every lane,
// whether any vertex of the chunk got a new lane.
// F[level][v][slot]: the masks of one vertex for all slots share a line,
// which the items of one chunk for the slots of a level read together.
__device__ inline size_t fidx(const BcArgs& a, int slot, int32_t L, int32_t v) {
    const int32_t lev = a.fmod == 2 ? (L & 1) : L;
    return ((size_t)lev * a.nv + v) * a.fstride + slot;
}

// Plane of the rank-packed values of level L of one slot ([nv][64]).
template <typename T = double>
__device__ inline T* stage_plane(const BcArgs& a, int slot, int32_t L) {
    return reinterpret_cast<T*>(a.stage) + ((size_t)slot * 2 + (size_t)(L & 1)) * (size_t)a.nv * 64;
}

__device__ inline uint64_t wave_or_u64(uint64_t m) {
    for (int o = 32; o >= 1; o >>= 1) m |= __shfl_xor(m, o);
    return m;
}

// Distances only: adds level L + 1 for the lanes of `got` to one (slot,
// chunk or heavy row) entry of the per-lane statistics.
__device__ inline void st_add(const BcArgs& a, int slot, int32_t idx, int32_t L, uint64_t got, int lane) {
    uint64_t* e = a.st + ((size_t)slot * (a.nchunks + a.hin.nh) + idx) * 192;
    if ((got >> lane) & 1) {
        e[lane] += (uint64_t)(L + 1);
        e[64 + lane] += 1;
        e[128 + lane] = (uint64_t)(L + 1);
    }
}

template <int MODE>
__device__ bool fwd_chunk(const BcArgs& a, int slot, int32_t ch, int32_t L, int lane) {
    const int32_t nv = a.nv;
    const uint64_t act = a.active[slot];
    uint64_t* seen = a.seen + (size_t)slot * nv;
    double* sig = MODE == 2 ? nullptr : a.sig + (size_t)slot * nv * 64;
    // sigma on levels 0 and 1 is 1 in every lane (one path from the source),
    // so those steps count the hits; later ones read the level's plane
    const bool ld = MODE == 0 && L >= 2;
    const double* rd = ld ? stage_plane(a, slot, L) : nullptr;
    double* wr = MODE == 0 && L >= 1 ? stage_plane(a, slot, L + 1) : nullptr;
    const int32_t vj = ch * 64 + lane;
    uint64_t sj = 0;
    int32_t bj = 0, ej = 0;
    bool cand = false, heavy = false;
    uint64_t st_d = 0, st_c = 0, st_e = 0;   // distances only
    if (vj < nv) {
        sj = seen[vj];
        bj = a.in_rp[vj];
        ej = a.in_rp[vj + 1];
        heavy = ej - bj > a.heavy;
        cand = (~sj & act) && ej > bj && !heavy;
    }
    uint64_t newF = 0;
    // a finished row j of the chunk: its sums and mask
    auto finish = [&](int32_t j, double acc, uint64_t got) {
        if constexpr (MODE == 2) {
            if ((got >> lane) & 1) {
                st_d += (uint64_t)(L + 1);
                st_c += 1;
                st_e = (uint64_t)(L + 1);
            }
        } else {
            if ((got >> lane) & 1) {
                const int32_t v = ch * 64 + j;
                __builtin_nontemporal_store(acc, &sig[(size_t)v * 64 + lane]);
                if (wr) wr[(size_t)v * 64 + lane_rank(got)] = acc;
            }
        }
        if (lane == j) newF = got;
    };
    // rows of more than 64 entries: one row at a time, 64 entries a pass
    uint64_t cm = __ballot(cand && ej - bj > 64);
    while (cm) {
        const int j = __ffsll((unsigned long long)cm) - 1;
        cm &= cm - 1;
        const uint64_t unseen = ~__shfl(sj, j) & act;
        const int32_t b = __shfl(bj, j), e = __shfl(ej, j);
        double acc = 0.0;
        uint64_t got = 0;
        // each 64-entry piece's column loads go out before the piece
        // before it gathers (not for the distances, which may end a row
        // early and keep the registers for a seventh wavefront per SIMD)
        int32_t un = MODE != 2 && b + lane < e ? a.in_col[b + lane] : 0;
        for (int32_t off = b; off < e; off += 64) {
            const int32_t k = off + lane;
            const int32_t uk = MODE != 2 ? un : (k < e ? a.in_col[k] : 0);
            if (MODE != 2 && off + 64 < e) un = off + 64 + lane < e ? a.in_col[off + 64 + lane] : 0;
            uint64_t mk = 0, fk = 0;
            if (k < e) {
                fk = a.F[fidx(a, slot, L, uk)];
                mk = fk & unseen;
            }
            if constexpr (MODE == 2) {
                got |= wave_or_u64(mk);
                if (got == unseen) break;    // every unseen lane found
            } else {
                uint64_t hits = __ballot(mk != 0);
                if (ld) while (hits) gather_hits<double, true>(hits, uk, mk, fk, rd, lane, acc, got);
                else while (hits) gather_hits<double, false>(hits, uk, mk, fk, rd, lane, acc, got);
            }
        }
        finish(j, acc, got);
    }
    // rows of at most 64 entries, packed: their entries laid end to end in
    // lane order, each 64-entry pass holding whole rows (row j ends at
    // entry E_j of the packed list)
    const bool shortr = cand && ej - bj <= 64;
    const int32_t lenj = shortr ? ej - bj : 0;
    const int32_t E = wave_incl_sum(lenj, lane);
    const int32_t total = __builtin_amdgcn_readlane(E, 63);
    const uint64_t unseen_j = ~sj & act;
    // each pass's column loads go out before the previous pass gathers
    PackedPass pc = packed_pass(0, E, lenj, bj, unseen_j, a.in_col, lane);
    while (pc.pend > 0) {
        PackedPass pn;
        pn.pend = -1;
        if (pc.pend < total) pn = packed_pass(pc.pend, E, lenj, bj, unseen_j, a.in_col, lane);
        uint64_t fk = 0, mk = 0;
        if (pc.on) {
            fk = a.F[fidx(a, slot, L, pc.uk)];
            mk = fk & pc.key;
        }
        uint64_t hits = __ballot(mk != 0);
        double acc = 0.0;
        uint64_t got = 0;
        int32_t cur = -1;
        if (MODE != 0 || !ld) while (hits) gather_rows<double, false>(hits, pc.uk, mk, fk, pc.rk, rd, lane, acc, got, cur, finish);
        else while (hits) gather_rows<double, true>(hits, pc.uk, mk, fk, pc.rk, rd, lane, acc, got, cur, finish);
        if (cur >= 0) finish(cur, acc, got);
        pc = pn;
    }
    if (vj < nv && !heavy) {
        a.F[fidx(a, slot, L + 1, vj)] = newF;
        if (newF) seen[vj] = sj | newF;
    }
    const bool any = __ballot(newF != 0) != 0;
    if constexpr (MODE == 2) {
        if (any) {
            uint64_t* e = a.st + ((size_t)slot * (a.nchunks + a.hin.nh) + ch) * 192;
            if (st_c) {
                e[lane] += st_d;
                e[64 + lane] += st_c;
                e[128 + lane] = st_e;
            }
        }
    }
    return any;
}

// Forward step of one slot over heavy-row segment sg (one wavefront): the
// lane sums of sigma over the segment's in-neighbours on the frontier, and
// the lanes they reach, go to the segment's partial.  fwd_combine finishes
// the row once every segment is done.  Returns, in every lane, whether the
// segment reached a new lane.
template <int MODE>
__device__ bool fwd_seg(const BcArgs& a, int slot, int32_t sg, int32_t L, int lane) {
    const int32_t nv = a.nv;
    const int32_t v = a.hin.seg_v[sg];
    const uint64_t act = a.active[slot];
    const bool ld = MODE == 0 && L >= 2;   // as in fwd_chunk
    const double* rd = ld ? stage_plane(a, slot, L) : nullptr;
    const uint64_t unseen = ~a.seen[(size_t)slot * nv + v] & act;
    double acc = 0.0;
    uint64_t got = 0;
    if (unseen) {
        const int32_t b = a.hin.seg_b[sg], e = a.hin.seg_e[sg];
        // each 64-entry piece's column loads go out before the piece
        // before it gathers (not for the distances, which may end a row
        // early and keep the registers for a seventh wavefront per SIMD)
        int32_t un = MODE != 2 && b + lane < e ? a.in_col[b + lane] : 0;
        for (int32_t off = b; off < e; off += 64) {
            const int32_t k = off + lane;
            const int32_t uk = MODE != 2 ? un : (k < e ? a.in_col[k] : 0);
            if (MODE != 2 && off + 64 < e) un = off + 64 + lane < e ? a.in_col[off + 64 + lane] : 0;
            uint64_t mk = 0, fk = 0;
            if (k < e) {
                fk = a.F[fidx(a, slot, L, uk)];
                mk = fk & unseen;
            }
            if constexpr (MODE == 2) {
                got |= wave_or_u64(mk);
                if (got == unseen) break;
            } else {
                uint64_t hits = __ballot(mk != 0);
                if (ld) while (hits) gather_hits<double, true>(hits, uk, mk, fk, rd, lane, acc, got);
                else while (hits) gather_hits<double, false>(hits, uk, mk, fk, rd, lane, acc, got);
            }
        }
    }
    if constexpr (MODE != 2) a.hpart[((size_t)slot * a.seg_stride + sg) * 64 + lane] = acc;
    if (lane == 0) a.hgot[(size_t)slot * a.seg_stride + sg] = got;
    return got != 0;
}

// Finishes heavy in-row h of one slot after its segments: the partials in
// segment order (sigma is integer-valued, so the sum equals the row-order
// one exactly) give the vertex's sigma; its F_{L+1} and seen follow.
template <int MODE>
__device__ void fwd_combine(const BcArgs& a, int slot, int32_t h, int32_t L, int lane) {
    const int32_t nv = a.nv;
    const int32_t f0 = a.hin.h_first[h], nsg = a.hin.h_nseg[h];
    const int32_t v = a.hin.seg_v[f0];
    uint64_t* Fn = a.F + fidx(a, slot, L + 1, v);
    const uint64_t sv = a.seen[(size_t)slot * nv + v];
    if (!(~sv & a.active[slot])) {   // every lane seen: the segments did nothing
        if (lane == 0) *Fn = 0;
        return;
    }
    const uint64_t* pgot = a.hgot + (size_t)slot * a.seg_stride + f0;
    uint64_t gm = 0;
    if constexpr (MODE == 2) {
        for (int32_t q = 0; q < nsg; q++) gm |= pgot[q];
        if (gm) st_add(a, slot, a.nchunks + h, L, gm, lane);
    } else {
        const double* part = a.hpart + ((size_t)slot * a.seg_stride + f0) * 64;
        double tot = 0.0;
        for (int32_t q = 0; q < nsg; q++) {
            tot += part[(size_t)q * 64 + lane];
            gm |= pgot[q];
        }
        if ((gm >> lane) & 1) {
            __builtin_nontemporal_store(tot, &a.sig[((size_t)slot * nv + v) * 64 + lane]);
            if (MODE == 0 && L >= 1) stage_plane(a, slot, L + 1)[(size_t)v * 64 + lane_rank(gm)] = tot;
        }
    }
    if (lane == 0) {
        *Fn = gm;
        if (gm) a.seen[(size_t)slot * nv + v] = sv | gm;
    }
}

// Forward level step, part 1: every light vertex of every chunk and every
// heavy-row segment; sets the slot's found flag when anything new is reached.
template <int MODE>
__global__ __launch_bounds__(kBlock) void bc_forward(BcArgs a) {
    const int slot = blockIdx.y;
    BcCtrl* c = &a.ctrl[slot];
    __shared__ int32_t s_L, s_done, s_found;
    const int tid = threadIdx.x, lane = tid & 63, wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    if (tid == 0) {
        s_done = c->done;
        s_L = c->level;
        s_found = 0;
        if (!s_done && a.fmod != 2 && s_L + 1 >= a.cap) {
            c->overflow = 1;
            s_done = 1;
        }
    }
    __syncthreads();
    if (s_done) return;
    const int32_t L = s_L;
    bool any = false;
    const int32_t nseg = a.hin.nseg;
    for (int32_t i = blockIdx.x * kWaves + wid; i < nseg + a.nchunks; i += gridDim.x * kWaves)
        any |= i < nseg ? fwd_seg<MODE>(a, slot, i, L, lane) : fwd_chunk<MODE>(a, slot, i - nseg, L, lane);
    if (any && lane == 0) s_found = 1;
    __syncthreads();
    if (tid == 0 && s_found) atomicOr(&c->found, 1);
}

// Forward level step, part 2: finishes the heavy rows; the last block of
// the launch advances the slot's level, or marks the slot done when the
// step found nothing (maxlevel = L).
template <int MODE>
__global__ __launch_bounds__(kBlock) void bc_forward_fin(BcArgs a) {
    const int slot = blockIdx.y;
    BcCtrl* c = &a.ctrl[slot];
    __shared__ int32_t s_L, s_done;
    __shared__ uint32_t s_ticket;
    const int tid = threadIdx.x, lane = tid & 63, wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    if (tid == 0) {
        s_done = c->done;
        s_L = c->level;
        if (!s_done && a.fmod != 2 && s_L + 1 >= a.cap) s_done = 1;
    }
    __syncthreads();
    if (s_done) return;
    const int32_t L = s_L;
    for (int32_t h = blockIdx.x * kWaves + wid; h < a.hin.nh; h += gridDim.x * kWaves) fwd_combine<MODE>(a, slot, h, L, lane);
    __syncthreads();
    if (tid == 0) {
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

// Backward step of one slot at level L (>= 1) over chunk ch (one
// wavefront): delta of every vertex of the chunk with lanes on level L from
// its successors' coefficients, the vertex's coefficient (1 + delta) / sigma
// into its sigma slot, and the lane sum of delta into the slot's partial
// betweenness.
// One lane of a vertex on level L once its successor sum acc is complete
// (sv: the lane's entry of sig): betweenness writes the coefficient
// (1 + delta) / sigma to *outp (the lane's entry of the level's plane) and
// returns delta = sigma * acc; stress (_stress_accumulate_basic!,
// stress.jl:55-77) returns acc times the predecessor count the forward pass
// left in sig, and writes acc + 1 for the vertex's predecessors.
template <int MODE>
__device__ inline typename BwdAcc<MODE>::T bwd_finish_lane(double sv, typename BwdAcc<MODE>::T* outp,
                                                           typename BwdAcc<MODE>::T acc) {
    if constexpr (MODE == 0) {
        const double d = sv * acc;
        *outp = (1.0 + d) / sv;
        return d;
    } else {
        *outp = acc + 1;
        return acc * (uint64_t)sv;
    }
}

template <typename T>
__device__ inline T lane_sum(T x) {
    for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

template <int MODE>
__device__ void bwd_chunk(const BcArgs& a, int slot, int32_t ch, int32_t L, int lane) {
    using T = typename BwdAcc<MODE>::T;
    const int32_t nv = a.nv;
    const double* sig = a.sig + (size_t)slot * nv * 64;
    const T* rd = stage_plane<T>(a, slot, L + 1);
    T* wr = stage_plane<T>(a, slot, L);
    T* acc_bc = reinterpret_cast<T*>(a.bc_acc) + (size_t)slot * nv;
    const int32_t vj = ch * 64 + lane;
    uint64_t Aj = 0;
    int32_t bj = 0, ej = 0;
    if (vj < nv) {
        Aj = a.F[fidx(a, slot, L, vj)];
        if (Aj) {
            bj = a.out_rp[vj];
            ej = a.out_rp[vj + 1];
            if (ej - bj > a.heavy) Aj = 0;   // a heavy row: its segments do it
        }
    }
    T bcj = 0;
    // row j of the chunk done: its successor sum acc gives the lanes'
    // coefficients and the vertex's contribution
    auto finish = [&](int32_t j, T acc) {
        const uint64_t A = readlane_u64(Aj, j);
        const int32_t v = ch * 64 + j;
        T contrib = 0;
        if ((A >> lane) & 1)
            contrib = bwd_finish_lane<MODE>(__builtin_nontemporal_load(&sig[(size_t)v * 64 + lane]), &wr[(size_t)v * 64 + lane_rank(A)], acc);
        contrib = lane_sum(contrib);
        if (lane == j) bcj = contrib;
    };
    const uint64_t am = __ballot(Aj != 0);
    // rows of more than 64 entries: one row at a time, 64 entries a pass
    uint64_t lm = am & __ballot(ej - bj > 64);
    while (lm) {
        const int j = __ffsll((unsigned long long)lm) - 1;
        lm &= lm - 1;
        const uint64_t A = __shfl(Aj, j);
        const int32_t b = __shfl(bj, j), e = __shfl(ej, j);
        T acc = 0;
        int32_t wn = b + lane < e ? a.out_col[b + lane] : 0;   // as in fwd_chunk
        for (int32_t off = b; off < e; off += 64) {
            const int32_t kk = off + lane;
            const int32_t wk = wn;
            if (off + 64 < e) wn = off + 64 + lane < e ? a.out_col[off + 64 + lane] : 0;
            uint64_t mk = 0, fk = 0;
            if (kk < e) {
                fk = a.F[fidx(a, slot, L + 1, wk)];
                mk = fk & A;
            }
            uint64_t hits = __ballot(mk != 0);
            uint64_t unused = 0;
            while (hits) gather_hits<T, true>(hits, wk, mk, fk, rd, lane, acc, unused);
        }
        finish(j, acc);
    }
    // rows without entries have no successors
    uint64_t zm = am & __ballot(ej == bj);
    while (zm) {
        const int j = __ffsll((unsigned long long)zm) - 1;
        zm &= zm - 1;
        finish(j, T(0));
    }
    // rows of 1 to 64 entries, packed as in fwd_chunk; a row of a pass with
    // no hit is finished with a zero sum, in row order with the others
    const int32_t lenj = (Aj != 0 && ej - bj <= 64) ? ej - bj : 0;
    const int32_t E = wave_incl_sum(lenj, lane);
    const int32_t total = __builtin_amdgcn_readlane(E, 63);
    uint64_t pending = 0;   // rows of the current pass not finished yet
    auto flush = [&](int32_t j, T acc, uint64_t) {
        uint64_t below = pending & ((1ull << j) - 1);
        while (below) {
            const int r = __ffsll((unsigned long long)below) - 1;
            below &= below - 1;
            finish(r, T(0));
        }
        finish(j, acc);
        pending &= j == 63 ? 0ull : ~((2ull << j) - 1);
    };
    PackedPass pc = packed_pass(0, E, lenj, bj, Aj, a.out_col, lane);
    while (pc.pend > 0) {
        PackedPass pn;
        pn.pend = -1;
        if (pc.pend < total) pn = packed_pass(pc.pend, E, lenj, bj, Aj, a.out_col, lane);
        pending = pc.rows;
        uint64_t fk = 0, mk = 0;
        if (pc.on) {
            fk = a.F[fidx(a, slot, L + 1, pc.uk)];
            mk = fk & pc.key;
        }
        uint64_t hits = __ballot(mk != 0);
        T acc = 0;
        uint64_t got = 0;
        int32_t cur = -1;
        while (hits) gather_rows<T, true>(hits, pc.uk, mk, fk, pc.rk, rd, lane, acc, got, cur, flush);
        if (cur >= 0) flush(cur, acc, got);
        while (pending) {
            const int r = __ffsll((unsigned long long)pending) - 1;
            pending &= pending - 1;
            finish(r, T(0));
        }
        pc = pn;
    }
    if (Aj) acc_bc[vj] += bcj;
}

// Backward step of one slot at level L over heavy-row segment sg: the lane
// sums of the successors' coefficients over the segment go to the segment's
// partial (nothing when the row's vertex has no lane on level L).
template <int MODE>
__device__ void bwd_seg(const BcArgs& a, int slot, int32_t sg, int32_t L, int lane) {
    using T = typename BwdAcc<MODE>::T;
    const int32_t nv = a.nv;
    const int32_t v = a.hout.seg_v[sg];
    const uint64_t A = a.F[fidx(a, slot, L, v)];
    if (!A) return;
    (void)nv;
    const T* rd = stage_plane<T>(a, slot, L + 1);
    const int32_t b = a.hout.seg_b[sg], e = a.hout.seg_e[sg];
    T acc = 0;
    int32_t wn = b + lane < e ? a.out_col[b + lane] : 0;   // as in fwd_chunk
    for (int32_t off = b; off < e; off += 64) {
        const int32_t kk = off + lane;
        const int32_t wk = wn;
        if (off + 64 < e) wn = off + 64 + lane < e ? a.out_col[off + 64 + lane] : 0;
        uint64_t mk = 0, fk = 0;
        if (kk < e) {
            fk = a.F[fidx(a, slot, L + 1, wk)];
            mk = fk & A;
        }
        uint64_t hits = __ballot(mk != 0);
        uint64_t unused = 0;
        while (hits) gather_hits<T, true>(hits, wk, mk, fk, rd, lane, acc, unused);
    }
    reinterpret_cast<T*>(a.hpart)[((size_t)slot * a.seg_stride + sg) * 64 + lane] = acc;
}

// Finishes heavy out-row h of one slot at level L after its segments: the
// partials in segment order give delta, then as bwd_chunk.
template <int MODE>
__device__ void bwd_combine(const BcArgs& a, int slot, int32_t h, int32_t L, int lane) {
    using T = typename BwdAcc<MODE>::T;
    const int32_t nv = a.nv;
    const int32_t f0 = a.hout.h_first[h], nsg = a.hout.h_nseg[h];
    const int32_t v = a.hout.seg_v[f0];
    const uint64_t A = a.F[fidx(a, slot, L, v)];
    if (!A) return;
    const T* part = reinterpret_cast<const T*>(a.hpart) + ((size_t)slot * a.seg_stride + f0) * 64;
    T tot = 0;
    for (int32_t q = 0; q < nsg; q++) tot += part[(size_t)q * 64 + lane];
    const double* sig = a.sig + (size_t)slot * nv * 64;
    T contrib = 0;
    if ((A >> lane) & 1)
        contrib = bwd_finish_lane<MODE>(__builtin_nontemporal_load(&sig[(size_t)v * 64 + lane]), &stage_plane<T>(a, slot, L)[(size_t)v * 64 + lane_rank(A)],
                                        tot);
    contrib = lane_sum(contrib);
    if (lane == 0) reinterpret_cast<T*>(a.bc_acc)[(size_t)slot * nv + v] += contrib;
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void bc_backward(BcArgs a, int32_t k) {
    const int slot = blockIdx.y;
    const int32_t L = a.ctrl[slot].maxlevel - k;
    if (L < 1) return;
    const int lane = threadIdx.x & 63, wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int32_t nseg = a.hout.nseg;
    for (int32_t i = blockIdx.x * kWaves + wid; i < nseg + a.nchunks; i += gridDim.x * kWaves) {
        if (i < nseg) bwd_seg<MODE>(a, slot, i, L, lane);
        else bwd_chunk<MODE>(a, slot, i - nseg, L, lane);
    }
}

// Backward level step, part 2: finishes the heavy rows of level maxlevel - k.
template <int MODE>
__global__ __launch_bounds__(kBlock) void bc_backward_fin(BcArgs a, int32_t k) {
    const int slot = blockIdx.y;
    const int32_t L = a.ctrl[slot].maxlevel - k;
    if (L < 1) return;
    const int lane = threadIdx.x & 63, wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int32_t h = blockIdx.x * kWaves + wid; h < a.hout.nh; h += gridDim.x * kWaves) bwd_combine<MODE>(a, slot, h, L, lane);
}

// Distances only: one block adds up the statistics entries of one slot
// (sums of distances and counts, the largest last level) into
// out[slot][3][64].  Every thread of the block calls it.
__device__ void st_reduce(const BcArgs& a, int slot, uint64_t* out) {
    __shared__ uint64_t s_part[kWaves][3][64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int32_t n = a.nchunks + a.hin.nh;
    const uint64_t* base = a.st + (size_t)slot * n * 192;
    uint64_t d = 0, c = 0, e = 0;
    for (int32_t i = wid; i < n; i += kWaves) {
        d += base[(size_t)i * 192 + lane];
        c += base[(size_t)i * 192 + 64 + lane];
        e = max(e, base[(size_t)i * 192 + 128 + lane]);
    }
    s_part[wid][0][lane] = d;
    s_part[wid][1][lane] = c;
    s_part[wid][2][lane] = e;
    __syncthreads();
    if (wid == 0) {
        for (int w = 1; w < kWaves; w++) {
            d += s_part[w][0][lane];
            c += s_part[w][1][lane];
            e = max(e, s_part[w][2][lane]);
        }
        out[(size_t)slot * 192 + lane] = d;
        out[(size_t)slot * 192 + 64 + lane] = c;
        out[(size_t)slot * 192 + 128 + lane] = e;
    }
    __syncthreads();
}

__global__ __launch_bounds__(kBlock) void bc_st_reduce(BcArgs a, uint64_t* out) { st_reduce(a, blockIdx.x, out); }

__global__ void bc_st_clear(BcArgs a, int32_t ns) {
    const int64_t n = (int64_t)ns * (a.nchunks + a.hin.nh) * 192;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a.st[i] = 0;
}

constexpr int kMaxSlotsP = 1024;        // slots of one persistent launch
constexpr int kMaxWordsP = kMaxSlotsP / 64;
constexpr int kMaxGridP = 4096;

constexpr int kPush0 = 16;   // level-0 push: wavefronts per source row

struct PersistArgs {
    BcPersistCtl* ctl;
    uint64_t* found;
    uint32_t* rel;        // [grid][32] per-block barrier flags
    const int32_t* lane_src;
    const double* endp;   // endpoints term or null
    double* out;
    uint64_t* st_out;     // distances only: [slot][3][64]
    double scale;
    uint64_t timeout;     // wall-clock ticks a barrier may wait
    uint64_t gather;      // wall-clock ticks without a new block that end the start-up
    int32_t ns;
    int32_t words;        // (ns + 63) / 64
    int32_t do_scale;
    int32_t pad;
};

// Barrier over the nblocks participating blocks of a launch (all resident):
// release fence and an arrival count; the last block to arrive writes the
// new generation into every block's own flag (one 128-byte line each, so
// the waiting blocks poll different addresses and their polls do not pile
// up on one memory channel beside the blocks still working); acquire
// fence.  A wait longer than `timeout` ticks, or an abort seen on the way,
// returns false in every thread of the block, which then leaves the kernel.
__device__ bool grid_sync(BcPersistCtl* c, uint32_t* rel, uint32_t vb, uint32_t nblocks, uint32_t& gen,
                          uint64_t timeout, int* s_ok) {
    __shared__ int s_last;
    __syncthreads();
    gen++;   // in every thread: the generation this barrier releases
    if (threadIdx.x == 0) {
        int ok = __hip_atomic_load(&c->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
        int last = 0;
        if (ok) {
            // the arrival releases this block's writes; only the last
            // arrival acquires the others' before it releases everyone
            const uint32_t prev = __hip_atomic_fetch_add(&c->bar_count, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            if (prev == nblocks - 1) {
                atomicExch(&c->bar_count, 0u);
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                last = 1;
            }
        }
        *s_ok = ok;
        s_last = last;
    }
    __syncthreads();
    if (s_last) {
        // one release fence in the first wavefront, then its lanes store
        // every block's flag: each flag store follows the fence in its own
        // lane (a release store per flag wrote the L2 back once per flag,
        // and kept this block late for the next barrier)
        if (threadIdx.x < 64) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            for (uint32_t b = threadIdx.x; b < nblocks; b += 64)
                __hip_atomic_store(&rel[(size_t)b * 32], gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    } else if (threadIdx.x == 0 && *s_ok) {
        const uint64_t t0 = wall_clock64();
        int nap = 0;
        // relaxed polls: an acquiring load would invalidate the L2 each time;
        // flags only grow, so a flag at or past gen has released this barrier
        while ((int32_t)(__hip_atomic_load(&rel[(size_t)vb * 32], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - gen) < 0) {
            if ((nap & 15) == 15 && __hip_atomic_load(&c->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                *s_ok = 0;
                break;
            }
            if (wall_clock64() - t0 > timeout) {
                atomicExch(&c->abort, 1);
                *s_ok = 0;
                break;
            }
            if (nap < 8) __builtin_amdgcn_s_sleep(4);
            else __builtin_amdgcn_s_sleep(32);
            nap++;
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // one L2 invalidate
    }
    __syncthreads();
    return *s_ok != 0;
}

__device__ inline uint64_t wave_or64(uint64_t m) {
    for (int o = 32; o >= 1; o >>= 1) m |= __shfl_xor(m, o);
    return m;
}

// The whole computation of up to kMaxSlotsP lane groups in one launch:
// clear, seed, forward levels until no slot finds a vertex, backward levels
// deepest first, then the reduction over slots with the endpoints term and
// the scale.  The chunk work is the same as bc_forward / bc_backward's,
// and every slot's sums are formed in the same order, so the result is
// bit-identical to the level-per-launch path.  Slot s finishes at the first
// level L whose step finds no new vertex for it (maxlevel = L); every block
// learns that from the per-block found masks after the barrier, so all
// blocks run the same number of barriers.  With heavy rows every level step
// has a second barrier, after the rows' partial sums are combined.
template <int MODE>
__global__ __launch_bounds__(kBlock) void bc_persist(BcArgs a, PersistArgs p) {
    __shared__ uint64_t s_found[kMaxWordsP];
    __shared__ uint64_t s_live[kMaxWordsP];
    __shared__ int32_t s_maxl[kMaxSlotsP];
    __shared__ int s_ok;
    __shared__ uint32_t s_vb, s_np;
    BcPersistCtl* c = p.ctl;
    const int tid = threadIdx.x, lane = tid & 63, wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    // Participants: a block may only wait for blocks that are resident.  The
    // blocks take tickets as they start; the first waits until no ticket has
    // been taken for kGatherTicks (or all are taken) and fixes the number of
    // participants, rounded down so that ns divides their wavefront count.
    // Blocks whose ticket is beyond it leave at once.  Which block does which
    // item does not change any result.
    if (tid == 0) {
        const uint32_t t = atomicAdd(&c->arrive, 1u);
        uint32_t np = 0;
        if (t == 0) {
            uint32_t last = 1;
            uint64_t t_last = wall_clock64();
            for (;;) {
                const uint32_t now = __hip_atomic_load(&c->arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint64_t tn = wall_clock64();
                if (now != last) {
                    last = now;
                    t_last = tn;
                }
                if (now >= gridDim.x || tn - t_last > p.gather) break;
                __builtin_amdgcn_s_sleep(4);
            }
            np = last;
            while (np > 0 && (np * kWaves) % (uint32_t)p.ns != 0) np--;
            __hip_atomic_store(&c->closed, np + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const uint64_t t0 = wall_clock64();
            for (;;) {
                const uint32_t cl = __hip_atomic_load(&c->closed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (cl) {
                    np = cl - 1;
                    break;
                }
                if (wall_clock64() - t0 > p.timeout) {
                    atomicExch(&c->abort, 1);
                    break;
                }
                __builtin_amdgcn_s_sleep(16);
            }
        }
        s_vb = t;
        s_np = np;
    }
    __syncthreads();
    const uint32_t vb = s_vb, G = s_np;   // this block's index among the G participants
    if (vb >= G) {
        if (G == 0 && vb == 0) atomicExch(&c->abort, 1);
        return;
    }
    // Work items: the heavy-row segments, then the 64-vertex chunks, each
    // once per slot, slot-minor.  Wavefront gw takes positions i = gw + r*nw
    // (r = 0, 1, ...), and position i is item (i / ns, slot (i + r) % ns):
    // as ns divides nw, every group of ns positions of one item is dealt in
    // one round r, so this is a bijection, and a wavefront's slot changes
    // from round to round, which spreads a slot with much work over all
    // wavefronts.  The host keeps the item counts below 2^31.
    const uint32_t gw = vb * kWaves + wid, nw = G * kWaves;
    const uint32_t nitems_f = (uint32_t)p.ns * (uint32_t)(a.hin.nseg + a.nchunks);
    const uint32_t nitems_b = (uint32_t)p.ns * (uint32_t)(a.hout.nseg + a.nchunks);
    const uint32_t nrows_f = (uint32_t)p.ns * (uint32_t)a.hin.nh, nrows_b = (uint32_t)p.ns * (uint32_t)a.hout.nh;
    const int64_t nv = a.nv;
    const int ns = p.ns, W = p.words;
    // barrier generations continue from the last launch's, so that no flag
    // left by it can release a barrier of this one
    uint32_t gen = __hip_atomic_load(&c->bar_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (vb == 0 && tid == 0) {
        c->overflow = 0;
        c->levels = 0;
        c->barriers = 0;
    }
    // clear seen, F_0 (and F_1, which the level-0 push below fills with
    // atomics) and the partial betweenness of every slot
    for (int64_t i = (int64_t)vb * kBlock + tid; i < (int64_t)ns * nv; i += (int64_t)G * kBlock) {
        const int64_t sl = i / nv;
        a.seen[i] = 0;
        a.F[fidx(a, (int)sl, 0, (int32_t)(i - sl * nv))] = 0;
        if constexpr (MODE != 2) {
            a.F[fidx(a, (int)sl, 1, (int32_t)(i - sl * nv))] = 0;
            a.bc_acc[i] = 0.0;
        }
    }
    if constexpr (MODE == 2) {
        const int64_t nst = (int64_t)ns * (a.nchunks + a.hin.nh) * 192;
        for (int64_t i = (int64_t)vb * kBlock + tid; i < nst; i += (int64_t)G * kBlock) a.st[i] = 0;
    }
    for (int w = tid; w < W; w += kBlock) s_live[w] = 0;
    __syncthreads();
    for (int sl = tid; sl < ns; sl += kBlock) {
        s_maxl[sl] = 0;
        if (a.active[sl]) atomicOr((unsigned long long*)&s_live[sl >> 6], 1ull << (sl & 63));
    }
    int nbar = 0;
    if (!grid_sync(c, p.rel, vb, G, gen, p.timeout, &s_ok)) return;
    nbar++;
    // seed: lane l of slot s starts at lane_src[s][l]
    for (int sl = (int)vb; sl < ns; sl += (int)G) {
        if (tid < 64) {
            const int32_t src = p.lane_src[sl * 64 + tid];
            if (src >= 0) {
                atomicOr((unsigned long long*)&a.seen[(size_t)sl * nv + src], 1ull << tid);
                atomicOr((unsigned long long*)&a.F[fidx(a, sl, 0, src)], 1ull << tid);
                if constexpr (MODE != 2) a.sig[((size_t)sl * nv + src) * 64 + tid] = 1.0;
            }
        }
    }
    if (!grid_sync(c, p.rel, vb, G, gen, p.timeout, &s_ok)) return;
    nbar++;
    if (vb == 0 && tid == 0) c->t_fwd = wall_clock64();
    int32_t L = 0;
    for (;;) {
        bool live = false;
        for (int w = 0; w < W; w++) live |= s_live[w] != 0;
        if (!live) break;
        if (a.fmod != 2 && L + 1 >= a.cap) {
            if (vb == 0 && tid == 0) c->overflow = 1;
            return;
        }
        for (int w = tid; w < W; w += kBlock) s_found[w] = 0;
        __syncthreads();
        const bool push = MODE != 2 && L == 0;
        if (push) {
            // Level 0 -> 1 by a push from the sources rather than a pull over
            // every vertex: lane i's level 1 is out(s_i) minus s_i itself, with
            // sigma 1 (a simple graph has one edge s_i -> v) -- for stress, one
            // predecessor -- which is what the pull computes.  Item (slot,
            // lane, part) takes every kPush0-th 64-entry round of the source's
            // row, so that a hub source is spread over kPush0 wavefronts.
            for (uint32_t i = gw; i < (uint32_t)ns * 64u * (uint32_t)kPush0; i += nw) {
                const int sl = (int)(i / (64u * (uint32_t)kPush0));
                const int ln = (int)((i / (uint32_t)kPush0) & 63u);
                const int part = (int)(i % (uint32_t)kPush0);
                if (!((s_live[sl >> 6] >> (sl & 63)) & 1)) continue;
                const int32_t src = p.lane_src[sl * 64 + ln];
                if (src < 0) continue;
                const int32_t b = a.out_rp[src], e = a.out_rp[src + 1];
                bool got = false;
                for (int32_t k = b + part * 64 + lane; k < e; k += kPush0 * 64) {
                    const int32_t v = a.out_col[k];
                    if (v == src) continue;
                    atomicOr((unsigned long long*)&a.F[fidx(a, sl, 1, v)], 1ull << ln);
                    atomicOr((unsigned long long*)&a.seen[(size_t)sl * nv + v], 1ull << ln);
                    a.sig[((size_t)sl * nv + v) * 64 + ln] = 1.0;
                    got = true;
                }
                if (__ballot(got) && lane == 0) atomicOr((unsigned long long*)&s_found[sl >> 6], 1ull << (sl & 63));
            }
        } else {
            for (uint32_t i = gw, r = 0; i < nitems_f; i += nw, r++) {
                const int sl = (int)((i + r) % (uint32_t)ns);
                const int32_t q = (int32_t)(i / (uint32_t)ns);
                if (!((s_live[sl >> 6] >> (sl & 63)) & 1)) continue;
                const bool got = q < a.hin.nseg ? fwd_seg<MODE>(a, sl, q, L, lane) : fwd_chunk<MODE>(a, sl, q - a.hin.nseg, L, lane);
                if (got && lane == 0) atomicOr((unsigned long long*)&s_found[sl >> 6], 1ull << (sl & 63));
            }
        }
        __syncthreads();
        uint64_t* fo = p.found + ((size_t)(L & 1) * G + vb) * W;
        for (int w = tid; w < W; w += kBlock) fo[w] = s_found[w];
        if (!grid_sync(c, p.rel, vb, G, gen, p.timeout, &s_ok)) return;
        nbar++;
        // heavy rows: one wavefront per (row, live slot) adds the partials
        // (none after the level-0 push, which wrote every row directly)
        for (uint32_t i = gw, r = 0; !push && i < nrows_f; i += nw, r++) {
            const int sl = (int)((i + r) % (uint32_t)ns);
            if ((s_live[sl >> 6] >> (sl & 63)) & 1) fwd_combine<MODE>(a, sl, (int32_t)(i / (uint32_t)ns), L, lane);
        }
        for (int w = tid; w < W; w += kBlock) s_found[w] = 0;
        __syncthreads();
        for (int w = 0; w < W; w++) {
            uint64_t m = 0;
            for (int b = tid; b < (int)G; b += kBlock)
                m |= __hip_atomic_load(&p.found[((size_t)(L & 1) * G + b) * W + w], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            m = wave_or64(m);
            if (lane == 0 && m) atomicOr((unsigned long long*)&s_found[w], m);
        }
        __syncthreads();
        for (int sl = tid; sl < ns; sl += kBlock)
            if (((s_live[sl >> 6] >> (sl & 63)) & 1) && !((s_found[sl >> 6] >> (sl & 63)) & 1)) s_maxl[sl] = L;
        __syncthreads();
        for (int w = tid; w < W; w += kBlock) s_live[w] &= s_found[w];
        if (nrows_f > 0) {
            if (!grid_sync(c, p.rel, vb, G, gen, p.timeout, &s_ok)) return;
            nbar++;
        } else {
            __syncthreads();
        }
        L++;
    }
    int32_t kmax = 0;
    for (int sl = 0; sl < ns; sl++) kmax = max(kmax, s_maxl[sl]);
    if (vb == 0) {
        if (tid == 0) {
            c->t_bwd = wall_clock64();
            c->levels = L;
        }
        for (int sl = tid; sl < ns; sl += kBlock) {
            a.ctrl[sl].maxlevel = s_maxl[sl];
            a.ctrl[sl].done = 1;
        }
    }
    if constexpr (MODE == 2) {
        if (vb == 0 && tid == 0) {
            c->t_end = wall_clock64();
            c->barriers = nbar;
            c->bar_gen = gen;
        }
        for (int sl = (int)vb; sl < ns; sl += (int)G) st_reduce(a, sl, p.st_out);
        return;
    }
    for (int32_t k = 0; k < kmax; k++) {
        for (uint32_t i = gw, r = 0; i < nitems_b; i += nw, r++) {
            const int sl = (int)((i + r) % (uint32_t)ns);
            const int32_t q = (int32_t)(i / (uint32_t)ns);
            const int32_t Ls = s_maxl[sl] - k;
            if (Ls < 1) continue;
            if (q < a.hout.nseg) bwd_seg<MODE>(a, sl, q, Ls, lane);
            else bwd_chunk<MODE>(a, sl, q - a.hout.nseg, Ls, lane);
        }
        if (!grid_sync(c, p.rel, vb, G, gen, p.timeout, &s_ok)) return;
        nbar++;
        if (nrows_b > 0) {
            for (uint32_t i = gw, r = 0; i < nrows_b; i += nw, r++) {
                const int sl = (int)((i + r) % (uint32_t)ns);
                const int32_t Ls = s_maxl[sl] - k;
                if (Ls >= 1) bwd_combine<MODE>(a, sl, (int32_t)(i / (uint32_t)ns), Ls, lane);
            }
            if (!grid_sync(c, p.rel, vb, G, gen, p.timeout, &s_ok)) return;
            nbar++;
        }
    }
    if (vb == 0 && tid == 0) {
        c->t_end = wall_clock64();
        c->barriers = nbar;
        c->bar_gen = gen;
    }
    for (int64_t v = (int64_t)vb * kBlock + tid; v < nv; v += (int64_t)G * kBlock) {
        if constexpr (MODE == 0) {
            double sum = 0.0;
            for (int k = 0; k < ns; k++) sum += a.bc_acc[(size_t)k * nv + v];
            if (p.endp) sum += p.endp[v];
            if (p.do_scale) sum *= p.scale;
            p.out[v] = sum;
        } else {
            uint64_t sum = 0;
            for (int k = 0; k < ns; k++) sum += reinterpret_cast<const uint64_t*>(a.bc_acc)[(size_t)k * nv + v];
            reinterpret_cast<uint64_t*>(p.out)[v] = sum;
        }
    }
}

// After a launch that stopped early: its blocks may have left flags up to
// some generations past bar_gen, so the next launch starts far beyond them.
__global__ void bc_ctl_reset(BcPersistCtl* c) {
    c->bar_count = 0;
    c->abort = 0;
    c->overflow = 0;
    c->bar_gen += 1u << 20;
}

// Runs before every persistent launch (in the same captured graph).
__global__ void bc_persist_prologue(BcPersistCtl* c) {
    c->arrive = 0;
    c->closed = 0;
}

// bc[v] = (sum over slots of bc_acc, in slot order, + endpoints term) * scale
template <int MODE>
__global__ void bc_finish(const double* acc, int32_t nslots, int32_t nv, const double* endp, int do_scale,
                          double scale, double* out) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x) {
        if constexpr (MODE == 0) {
            double s = 0.0;
            for (int k = 0; k < nslots; k++) s += acc[(size_t)k * nv + v];
            if (endp) s += endp[v];
            if (do_scale) s *= scale;
            out[v] = s;
        } else {
            uint64_t s = 0;
            for (int k = 0; k < nslots; k++) s += reinterpret_cast<const uint64_t*>(acc)[(size_t)k * nv + v];
            reinterpret_cast<uint64_t*>(out)[v] = s;
        }
    }
}

}  // namespace

void bc_free(BcWork* w) {
    if (!w) return;
    dev_free(w->seen); dev_free(w->F); dev_free(w->sig); dev_free(w->stage); dev_free(w->bc_acc); dev_free(w->bc_out);
    dev_free(w->endp); dev_free(w->active); dev_free(w->lane_src); dev_free(w->ctrl);
    if (w->h_ctrl) (void)hipHostFree(w->h_ctrl);
    for (auto& e : w->ev) if (e) (void)hipEventDestroy(e);
    if (w->pexec) (void)hipGraphExecDestroy(w->pexec);
    dev_free(w->pctl); dev_free(w->found); dev_free(w->rel);
    dev_free(w->hpart); dev_free(w->hgot); dev_free(w->st); dev_free(w->st_out);
    if (w->h_pctl) (void)hipHostFree(w->h_pctl);
    delete w;
}

static size_t free_device_bytes() {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) { (void)hipGetLastError(); return 0; }
    return fr;
}

// (Re)allocates the per-slot state for `nslots` slots and `cap` levels.
static void bc_alloc(lgmi_graph* g, int32_t nslots, int32_t cap, int32_t seg_stride, bool need_sig) {
    BcWork* w = g->bc;
    if (w && w->nslots >= nslots && w->cap_levels >= cap && w->nv == g->nv && w->seg_stride >= seg_stride &&
        (w->has_sig || !need_sig))
        return;
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
        if (need_sig) {
            w->sig = dev_alloc_n<double>(nv * nslots * 64);
            w->stage = dev_alloc_n<double>(nv * nslots * 2 * 64);
            w->bc_acc = dev_alloc_n<double>(nv * nslots);
            w->has_sig = true;
        }
        w->bc_out = dev_alloc_n<double>(nv);
        w->endp = dev_alloc_n<double>(nv);
        w->active = dev_alloc_n<uint64_t>(nslots);
        w->lane_src = dev_alloc_n<int32_t>((size_t)nslots * 64);
        w->ctrl = dev_alloc_n<BcCtrl>(nslots);
        LGMI_HIP(hipHostMalloc((void**)&w->h_ctrl, sizeof(BcCtrl) * nslots, hipHostMallocDefault));
        for (auto& e : w->ev) LGMI_HIP(hipEventCreate(&e));
        w->seg_stride = seg_stride;
        if (need_sig) w->hpart = dev_alloc_n<double>((size_t)std::max(seg_stride, 1) * nslots * 64);
        w->hgot = dev_alloc_n<uint64_t>((size_t)std::max(seg_stride, 1) * nslots);
        w->pctl = dev_alloc_n<BcPersistCtl>(1);
        LGMI_HIP(hipHostMalloc((void**)&w->h_pctl, sizeof(BcPersistCtl), hipHostMallocDefault));
        w->found = dev_alloc_n<uint64_t>((size_t)2 * kMaxGridP * kMaxWordsP);
        w->rel = dev_alloc_n<uint32_t>((size_t)kMaxGridP * 32);
        LGMI_HIP(hipMemsetAsync(w->pctl, 0, sizeof(BcPersistCtl), g->stream));
        LGMI_HIP(hipMemsetAsync(w->rel, 0, sizeof(uint32_t) * kMaxGridP * 32, g->stream));
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
    // F[level][v][slot]: the first `keep` levels are a prefix
    LGMI_HIP(hipMemcpyAsync(F2, w->F, sizeof(uint64_t) * nv * (size_t)w->nslots * (size_t)keep, hipMemcpyDeviceToDevice,
                            g->stream));
    LGMI_HIP(hipStreamSynchronize(g->stream));
    dev_free(w->F);
    w->F = F2;
    w->cap_levels = new_cap;
}

// Host row pointers and heavy-row segments of g, built on first use.
static BcGraph* bc_graph_get(lgmi_graph* g) {
    if (g->bcg) return g->bcg;
    const int32_t nv = g->nv;
    BcGraph* b = new BcGraph();
    const int32_t seg_len = tuning_get().bc_seg_len;
    b->seg_len = seg_len;
    try {
        b->rp_out.resize((size_t)nv + 1);
        LGMI_HIP(hipMemcpy(b->rp_out.data(), g->out.row_ptr, sizeof(int32_t) * ((size_t)nv + 1), hipMemcpyDeviceToHost));
        if (g->directed) {
            b->rp_in.resize((size_t)nv + 1);
            LGMI_HIP(hipMemcpy(b->rp_in.data(), g->in.row_ptr, sizeof(int32_t) * ((size_t)nv + 1), hipMemcpyDeviceToHost));
        }
        // seg_v, seg_b, seg_e, seg_h, h_first, h_nseg of one CSR
        auto build = [&](const std::vector<int32_t>& rp, std::vector<int32_t> (&arr)[6]) {
            for (int32_t v = 0; v < nv; v++) {
                const int32_t bb = rp[v], e = rp[v + 1];
                if (e - bb <= seg_len) continue;
                const int32_t h = (int32_t)arr[4].size();
                arr[4].push_back((int32_t)arr[0].size());
                for (int32_t x = bb; x < e; x += seg_len) {
                    arr[0].push_back(v);
                    arr[1].push_back(x);
                    arr[2].push_back(std::min(e, x + seg_len));
                    arr[3].push_back(h);
                }
                arr[5].push_back((int32_t)arr[0].size() - arr[4].back());
            }
        };
        std::vector<int32_t> ain[6], aout[6];
        build(g->directed ? b->rp_in : b->rp_out, ain);
        if (g->directed) build(b->rp_out, aout);
        size_t total = 0;
        for (auto& x : ain) total += x.size();
        for (auto& x : aout) total += x.size();
        b->dev = dev_alloc_n<int32_t>(std::max<size_t>(total, 1));
        std::vector<int32_t> packed;
        packed.reserve(total);
        int32_t* base = b->dev;
        for (int k = 0; k < 6; k++) {
            b->seg_in[k] = base + packed.size();
            packed.insert(packed.end(), ain[k].begin(), ain[k].end());
        }
        for (int k = 0; k < 6; k++) {
            b->seg_out[k] = base + packed.size();
            packed.insert(packed.end(), aout[k].begin(), aout[k].end());
        }
        b->nseg_in = (int32_t)ain[0].size();
        b->nh_in = (int32_t)ain[4].size();
        if (g->directed) {
            b->nseg_out = (int32_t)aout[0].size();
            b->nh_out = (int32_t)aout[4].size();
        } else {
            for (int k = 0; k < 6; k++) b->seg_out[k] = b->seg_in[k];
            b->nseg_out = b->nseg_in;
            b->nh_out = b->nh_in;
        }
        if (total > 0)
            LGMI_HIP(hipMemcpy(b->dev, packed.data(), sizeof(int32_t) * total, hipMemcpyHostToDevice));
    } catch (...) {
        bc_graph_free(b);
        throw;
    }
    g->bcg = b;
    return b;
}

static BcSegs make_segs(int32_t* const (&arr)[6], int32_t nseg, int32_t nh) {
    BcSegs r{};
    r.seg_v = arr[0]; r.seg_b = arr[1]; r.seg_e = arr[2]; r.seg_h = arr[3]; r.h_first = arr[4]; r.h_nseg = arr[5];
    r.nseg = nseg;
    r.nh = nh;
    return r;
}

// 0: one persistent launch when the slots fit (level-per-launch otherwise
// and as the fallback), 1: level-per-launch only.
static std::atomic<int> g_launch_mode{0};

void bc_set_launch(int mode) { g_launch_mode = mode; }

// One launch for all groups (ngroups <= slots), captured in a hipGraph (a
// plain launch when capture fails).  The launch is an ordinary one: the
// kernel itself keeps only the blocks that start together (tickets), so it
// needs no co-residency promise from a cooperative launch, and the runtime
// keeps no cooperative-launch state alive until the process ends.
// Returns false when the launch could not finish (a barrier timed out or a
// slot ran past the level capacity); the caller then runs the level-per-launch
// path, which overwrites everything this one wrote.
static bool bc_persistent(lgmi_graph* g, BcArgs a, int32_t ns, const double* endp, int do_scale, double scale,
                          int mode, int32_t* launch_kind, float* fwd_ms, float* bwd_ms, int32_t* max_levels,
                          int32_t* participants, int32_t* blocks) {
    const void* kfn = mode == 2 ? reinterpret_cast<const void*>(bc_persist<2>)
                      : mode == 1 ? reinterpret_cast<const void*>(bc_persist<1>)
                                  : reinterpret_cast<const void*>(bc_persist<0>);
    BcWork* w = g->bc;
    hipStream_t st = g->stream;
    int dev = 0;
    LGMI_HIP(hipGetDevice(&dev));
    int khz = 0, per_cu = 0;
    LGMI_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    if (khz <= 0) return false;
    LGMI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, kBlock, 0));
    const int cu = std::max(1, g->cu_count);
    // the blocks that the occupancy calculation says fit; the kernel itself
    // keeps only those that start together
    const int grid = std::min(std::min(per_cu, 8) * cu, kMaxGridP);
    if (per_cu < 1 || grid * kWaves < ns) return false;
    PersistArgs p{};
    p.ctl = w->pctl;
    p.found = w->found;
    p.rel = w->rel;
    p.st_out = w->st_out;
    p.lane_src = w->lane_src;
    p.endp = endp;
    p.out = w->bc_out;
    p.scale = scale;
    p.timeout = (uint64_t)khz * 1000ull * 5ull;  // 5 s
    p.gather = (uint64_t)khz / 50ull + 1;          // 20 us
    p.ns = ns;
    p.words = (ns + 63) / 64;
    p.do_scale = do_scale;
    void* args[] = {&a, &p};
    const dim3 gd(grid), bd(kBlock);
    std::vector<unsigned char> key(sizeof(a) + sizeof(p) + 2 * sizeof(int));
    std::memcpy(key.data(), &a, sizeof(a));
    std::memcpy(key.data() + sizeof(a), &p, sizeof(p));
    std::memcpy(key.data() + sizeof(a) + sizeof(p), &grid, sizeof(int));
    std::memcpy(key.data() + sizeof(a) + sizeof(p) + sizeof(int), &mode, sizeof(int));
    if (w->pexec && w->pkey != key) {
        (void)hipGraphExecDestroy(w->pexec);
        w->pexec = nullptr;
    }
    if (!w->pexec) {
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            bc_persist_prologue<<<1, 1, 0, st>>>(w->pctl);
            hipError_t el = hipGetLastError();
            if (el == hipSuccess)
                el = hipLaunchKernel(kfn, gd, bd, args, 0, st);
            e = hipStreamEndCapture(st, &graph);
            if (el != hipSuccess) e = el;
        }
        if (e == hipSuccess && graph) e = hipGraphInstantiate(&w->pexec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            if (w->pexec) (void)hipGraphExecDestroy(w->pexec);
            w->pexec = nullptr;
        } else {
            w->pkey = key;
        }
    }
    LGMI_HIP(hipEventRecord(w->ev[1], st));
    if (w->pexec) {
        LGMI_HIP(hipGraphLaunch(w->pexec, st));
        *launch_kind = 2;
    } else {
        bc_persist_prologue<<<1, 1, 0, st>>>(w->pctl);
        LGMI_HIP(hipGetLastError());
        LGMI_HIP(hipLaunchKernel(kfn, gd, bd, args, 0, st));
        *launch_kind = 1;
    }
    LGMI_HIP(hipEventRecord(w->ev[2], st));
    LGMI_HIP(hipMemcpyAsync(w->h_pctl, w->pctl, sizeof(BcPersistCtl), hipMemcpyDeviceToHost, st));
    LGMI_HIP(hipMemcpyAsync(w->h_ctrl, w->ctrl, sizeof(BcCtrl) * ns, hipMemcpyDeviceToHost, st));
    LGMI_HIP(hipStreamSynchronize(st));
    const BcPersistCtl c = *w->h_pctl;
    if (c.abort || c.overflow) {
        bc_ctl_reset<<<1, 1, 0, st>>>(w->pctl);
        LGMI_HIP(hipGetLastError());
        return false;
    }
    const double tick_ms = 1.0 / (double)khz;
    *participants = c.closed > 0 ? (int32_t)c.closed - 1 : 0;
    *blocks = grid;
    *fwd_ms = (float)((double)(c.t_bwd - c.t_fwd) * tick_ms);
    *bwd_ms = (float)((double)(c.t_end - c.t_bwd) * tick_ms);
    int32_t lm = 0;
    for (int s = 0; s < ns; s++) lm = std::max(lm, w->h_ctrl[s].maxlevel);
    *max_levels = lm;
    return true;
}

// Lane assignment of the sources.  A hit of the gathers reads the 64-lane
// row of one vertex for the lanes whose searches place it on the level
// before; the fewer 128-byte quarters of the row those lanes span, the
// fewer lines it costs.  Sources that share their highest-degree neighbour
// have nearly the same distances to everything else (offset by one at most),
// so the sources are ordered by that neighbour (highest degree first), then
// by their own degree, then by position: sources with alike searches end up
// in the same lane group and the same quarter (and, for the distance
// statistics, a vertex's unseen lanes tend to be found by the same
// neighbour, which ends its scan early).  Betweenness and stress sum over
// the sources, so only the floating-point summation order changes; the
// distance statistics go back to their sources' positions (pos).  The order
// is a function of the graph and the source list alone.
static void order_sources(lgmi_graph* g, std::vector<int32_t>& act, std::vector<int32_t>& pos) {
    const int32_t n = (int32_t)act.size();
    int32_t* d = dev_alloc_n<int32_t>((size_t)n * 2);
    std::vector<int32_t> hub((size_t)n);
    try {
        LGMI_HIP(hipMemcpyAsync(d, act.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, g->stream));
        bc_source_hubs<<<(unsigned)((n + 3) / 4), 256, 0, g->stream>>>(g->out.row_ptr, g->out.col, d, n, d + n);
        LGMI_HIP(hipGetLastError());
        LGMI_HIP(hipMemcpyAsync(hub.data(), d + n, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, g->stream));
        LGMI_HIP(hipStreamSynchronize(g->stream));
    } catch (...) {
        dev_free(d);
        throw;
    }
    dev_free(d);
    const std::vector<int32_t>& rp = g->bcg->rp_out;
    auto odeg = [&](int32_t v) { return v < 0 ? -1 : rp[(size_t)v + 1] - rp[(size_t)v]; };
    std::vector<int32_t> idx((size_t)n);
    for (int32_t i = 0; i < n; i++) idx[(size_t)i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t x, int32_t y) {
        const int32_t hx = hub[(size_t)x], hy = hub[(size_t)y];
        const int32_t dx = odeg(hx), dy = odeg(hy);
        if (dx != dy) return dx > dy;
        if (hx != hy) return hx < hy;
        const int32_t sx = odeg(act[(size_t)x]), sy = odeg(act[(size_t)y]);
        return sx > sy;
    });
    std::vector<int32_t> out((size_t)n);
    std::vector<int32_t> pos2((size_t)n);
    for (int32_t i = 0; i < n; i++) {
        out[(size_t)i] = act[(size_t)idx[(size_t)i]];
        pos2[(size_t)i] = pos[(size_t)idx[(size_t)i]];
    }
    act.swap(out);
    pos.swap(pos2);
}

void bc_run(lgmi_graph* g, const int32_t* srcs, int32_t nsrc, int64_t k_scale, int skip_degree0, int normalize,
            int endpoints, double* bc, lgmi_bc_stats* stats, int mode, uint64_t* dist_stats) {
    const int32_t nv = g->nv;
    hipStream_t st = g->stream;
    // degree(g, s) > 0 filter (betweenness.jl:58) on host copies of the row pointers
    BcGraph* bg = bc_graph_get(g);
    const std::vector<int32_t>& rp_out = bg->rp_out;
    const std::vector<int32_t>& rp_in = bg->rp_in;
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
    // pos[i]: the position among the active sources that act[i] had before
    // the ordering (the distance statistics are returned in that order)
    std::vector<int32_t> pos((size_t)nact);
    for (int64_t i = 0; i < nact; i++) pos[(size_t)i] = (int32_t)i;
    if (nact > 16) order_sources(g, act, pos);

    // slots: as many groups at once as fit a memory budget
    const int32_t cap0 = mode == 2 ? 2 : 32;
    const int32_t seg_stride = std::max(bg->nseg_in, bg->nseg_out);
    const int32_t nchunks = (nv + 63) / 64;
    const size_t per_slot = mode == 2
        ? (size_t)std::max(nv, 1) * (8 + 8 * (size_t)cap0 + 8) + (size_t)seg_stride * 8 +
              ((size_t)nchunks + bg->nh_in) * 192 * 8
        : (size_t)std::max(nv, 1) * (8 + 8 * (size_t)cap0 + 3 * 512 + 8) + (size_t)seg_stride * 520;
    // half the free memory: 16 slots of RMAT-22 (7.6 GB each with the
    // stage planes) then fit one persistent launch
    const size_t free_b = free_device_bytes();
    size_t budget = free_b / 2;
    if (budget == 0) budget = (size_t)4 << 30;
    // a graph whose single lane group does not fit is refused as too large
    // (the Julia shim then runs the reference method)
    if (free_b > 0 && per_slot > free_b - free_b / 8)
        throw Error(LGMI_E_TOO_LARGE, "the state of one 64-source lane group (" + std::to_string(per_slot) +
                                          " bytes) exceeds the free device memory");
    int32_t nslots = (int32_t)std::max<size_t>(1, std::min<size_t>((size_t)std::max(ngroups, 1), budget / per_slot));
    nslots = std::min(nslots, kMaxSlotsP);
    bc_alloc(g, nslots, cap0, seg_stride, mode != 2);
    BcWork* w = g->bc;
    nslots = std::min(w->nslots, std::max(ngroups, 1));
    if (mode == 2) {
        const size_t words = (size_t)nslots * ((size_t)nchunks + bg->nh_in) * 192;
        if (w->st_words < words) {
            dev_free(w->st);
            dev_free(w->st_out);
            w->st = nullptr;
            w->st_out = nullptr;
            w->st_words = 0;
            w->st = dev_alloc_n<uint64_t>(words);
            w->st_out = dev_alloc_n<uint64_t>((size_t)w->nslots * 192);
            w->st_words = words;
        }
    }

    BcArgs a{};
    a.in_rp = g->in.row_ptr; a.in_col = g->in.col;
    a.out_rp = g->out.row_ptr; a.out_col = g->out.col;
    a.nv = nv;
    a.nchunks = (nv + 63) / 64;
    a.seen = w->seen; a.F = w->F; a.sig = w->sig; a.stage = w->stage; a.bc_acc = w->bc_acc;
    a.active = w->active; a.ctrl = w->ctrl; a.cap = w->cap_levels;
    a.hin = make_segs(bg->seg_in, bg->nseg_in, bg->nh_in);
    a.hout = make_segs(bg->seg_out, bg->nseg_out, bg->nh_out);
    a.hpart = w->hpart; a.hgot = w->hgot;
    a.seg_stride = w->seg_stride;
    a.heavy = bg->seg_len;
    a.fmod = mode == 2 ? 2 : w->cap_levels;
    a.fstride = w->nslots;
    a.st = w->st;
    if (mode == 2) { a.sig = nullptr; a.stage = nullptr; a.bc_acc = nullptr; }

    const int cu = std::max(1, g->cu_count);
    const int target_blocks = cu * 8;
    const int bx = std::max(1, std::min((seg_stride + a.nchunks + kWaves - 1) / kWaves, std::max(1, target_blocks / nslots)));
    const int ibx = std::max(1, std::min((nv + 255) / 256, std::max(1, target_blocks / nslots)));
    const int hbx_in = std::max(1, std::min((a.hin.nh + kWaves - 1) / kWaves, std::max(1, target_blocks / nslots)));
    const int hbx_out = std::max(1, std::min((a.hout.nh + kWaves - 1) / kWaves, std::max(1, target_blocks / nslots)));

    // endpoints: every processed source s adds length(S)-1 = nv-1 to bc[s]
    // and 1 to every other vertex (betweenness.jl:122,130)
    const double* endp = nullptr;
    if (mode == 0 && endpoints && nact > 0) {
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
    if (mode != 0) do_scale = 0;   // stress: plain Int64 sums

    LGMI_HIP(hipEventRecord(w->ev[0], st));
    std::vector<int32_t> lane_src((size_t)nslots * 64);
    std::vector<uint64_t> active((size_t)nslots);
    auto fill_groups = [&](int32_t g0, int32_t ns) {
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
    };

    // distances only: slot s, lane l of groups g0.. is source (g0 + s) * 64 + l;
    // dist_stats[i] = {sum of distances, vertices reached (the source
    // included), largest distance}
    std::vector<uint64_t> st_host(mode == 2 ? (size_t)nslots * 192 : 0);
    auto scatter_stats = [&](int32_t g0, int32_t ns) {
        for (int32_t s = 0; s < ns; s++)
            for (int l = 0; l < 64; l++) {
                const int64_t idx = (int64_t)(g0 + s) * 64 + l;
                if (idx >= nact) continue;
                const int64_t o = pos[(size_t)idx];
                dist_stats[o * 3 + 0] = st_host[(size_t)s * 192 + l];
                dist_stats[o * 3 + 1] = st_host[(size_t)s * 192 + 64 + l] + 1;
                dist_stats[o * 3 + 2] = st_host[(size_t)s * 192 + 128 + l];
            }
    };
    float fwd_ms = 0.f, bwd_ms = 0.f;
    int32_t max_levels = 0;
    int32_t launch_kind = 0;
    int32_t participants = 0, blocks = 0;
    const int64_t max_items = (int64_t)ngroups * ((int64_t)seg_stride + a.nchunks);
    if (g_launch_mode == 0 && ngroups >= 1 && ngroups <= nslots && nv > 0 && max_items < ((int64_t)1 << 31)) {
        fill_groups(0, ngroups);
        if (!bc_persistent(g, a, ngroups, endp, do_scale, scale, mode, &launch_kind, &fwd_ms, &bwd_ms, &max_levels,
                           &participants, &blocks)) {
            launch_kind = 0;
            participants = blocks = 0;
            fwd_ms = bwd_ms = 0.f;
            max_levels = 0;
        } else if (mode == 2) {
            LGMI_HIP(hipMemcpyAsync(st_host.data(), w->st_out, sizeof(uint64_t) * 192 * (size_t)ngroups,
                                    hipMemcpyDeviceToHost, st));
            LGMI_HIP(hipStreamSynchronize(st));
            scatter_stats(0, ngroups);
        }
    }
    if (launch_kind == 0) {
        if (mode != 2)
            bc_clear_acc<<<std::max(1, std::min(target_blocks, (int)(((size_t)nv * nslots + 255) / 256))), 256, 0, st>>>(
                w->bc_acc, (int64_t)nv * nslots);
        for (int32_t g0 = 0; g0 < ngroups; g0 += nslots) {
            const int32_t ns = std::min(nslots, ngroups - g0);
            fill_groups(g0, ns);
            a.F = w->F; a.cap = w->cap_levels;
            if (mode != 2) a.fmod = w->cap_levels;
            bc_init<<<dim3(ibx, ns), 256, 0, st>>>(a);
            if (mode == 2) bc_st_clear<<<target_blocks, 256, 0, st>>>(a, ns);
            bc_seed<<<ns, 64, 0, st>>>(a, w->lane_src);
            LGMI_HIP(hipGetLastError());
            LGMI_HIP(hipEventRecord(w->ev[1], st));
            // forward: level steps in batches until every slot is done
            int32_t lcur = 0;
            int batch = 8;
            for (;;) {
                if (mode != 2 && lcur + batch + 2 >= w->cap_levels) {
                    int32_t nc = w->cap_levels;
                    while (lcur + batch + 2 >= nc) nc *= 2;
                    // a finished slot also needs its empty F_{maxlevel+1}
                    bc_grow_levels(g, nc, std::min(lcur + 2, w->cap_levels));
                    a.F = w->F; a.cap = w->cap_levels; a.fmod = w->cap_levels;
                }
                for (int i = 0; i < batch; i++) {
                    if (mode == 2) {
                        bc_forward<2><<<dim3(bx, ns), kBlock, 0, st>>>(a);
                        bc_forward_fin<2><<<dim3(hbx_in, ns), kBlock, 0, st>>>(a);
                    } else if (mode == 1) {
                        bc_forward<1><<<dim3(bx, ns), kBlock, 0, st>>>(a);
                        bc_forward_fin<1><<<dim3(hbx_in, ns), kBlock, 0, st>>>(a);
                    } else {
                        bc_forward<0><<<dim3(bx, ns), kBlock, 0, st>>>(a);
                        bc_forward_fin<0><<<dim3(hbx_in, ns), kBlock, 0, st>>>(a);
                    }
                }
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
            if (mode == 2) {
                bc_st_reduce<<<ns, kBlock, 0, st>>>(a, w->st_out);
                LGMI_HIP(hipGetLastError());
                LGMI_HIP(hipMemcpyAsync(st_host.data(), w->st_out, sizeof(uint64_t) * 192 * (size_t)ns,
                                        hipMemcpyDeviceToHost, st));
                LGMI_HIP(hipStreamSynchronize(st));
                scatter_stats(g0, ns);
            }
            for (int32_t k = 0; mode != 2 && k < lmax; k++) {
                if (mode) bc_backward<1><<<dim3(bx, ns), kBlock, 0, st>>>(a, k);
                else bc_backward<0><<<dim3(bx, ns), kBlock, 0, st>>>(a, k);
                if (a.hout.nh > 0) {
                    if (mode) bc_backward_fin<1><<<dim3(hbx_out, ns), kBlock, 0, st>>>(a, k);
                    else bc_backward_fin<0><<<dim3(hbx_out, ns), kBlock, 0, st>>>(a, k);
                }
            }
            LGMI_HIP(hipGetLastError());
            LGMI_HIP(hipEventRecord(w->ev[3], st));
            LGMI_HIP(hipEventSynchronize(w->ev[3]));
            float t1 = 0.f, t2 = 0.f;
            LGMI_HIP(hipEventElapsedTime(&t1, w->ev[1], w->ev[2]));
            LGMI_HIP(hipEventElapsedTime(&t2, w->ev[2], w->ev[3]));
            fwd_ms += t1;
            bwd_ms += t2;
        }
        if (nv > 0 && mode != 2) {
            const int fb = std::max(1, std::min(target_blocks, (nv + 255) / 256));
            if (mode) bc_finish<1><<<fb, 256, 0, st>>>(w->bc_acc, nslots, nv, endp, do_scale, scale, w->bc_out);
            else bc_finish<0><<<fb, 256, 0, st>>>(w->bc_acc, nslots, nv, endp, do_scale, scale, w->bc_out);
            LGMI_HIP(hipGetLastError());
        }
    }
    if (nv > 0 && mode != 2) LGMI_HIP(hipMemcpyAsync(bc, w->bc_out, sizeof(double) * (size_t)nv, hipMemcpyDeviceToHost, st));
    LGMI_HIP(hipEventRecord(w->ev[3], st));
    LGMI_HIP(hipEventSynchronize(w->ev[3]));
    if (stats) {
        float tot = 0.f;
        LGMI_HIP(hipEventElapsedTime(&tot, w->ev[0], w->ev[3]));
        stats->groups = ngroups;
        stats->max_levels = max_levels;
        stats->launch = launch_kind;
        stats->participants = participants;
        stats->blocks = blocks;
        stats->pad = 0;
        stats->device_ms = tot;
        stats->forward_ms = fwd_ms;
        stats->backward_ms = bwd_ms;
    }
}

}  // namespace lgmi
