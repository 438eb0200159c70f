// lgmi_internal.h -- shared definitions for the liblgmi.so sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lgmi.h"

namespace lgmi {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define LGMI_HIP(expr)                                                                        \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess)                                                                 \
            throw ::lgmi::Error(_e == hipErrorOutOfMemory ? LGMI_E_ALLOC : LGMI_E_HIP,        \
                                std::string(#expr) + ": " + hipGetErrorString(_e));           \
    } while (0)

// Device allocation with an optional injected failure (LGMI_FAIL_ALLOC=1)
// so that error paths can be exercised from the host.
void* dev_alloc(size_t bytes);
void dev_free(void* p);
void ensure_device();  // throws LGMI_E_NODEV when no GPU is visible
int device_cu_count();

template <class T>
T* dev_alloc_n(size_t n) {
    return static_cast<T*>(dev_alloc(sizeof(T) * (n ? n : 1)));
}

// Wavefront helpers for packing short adjacency rows end to end over the 64
// lanes.  Every lane of the wavefront calls them.
// Inclusive prefix sum of x over the lanes.
__device__ inline int32_t wave_incl_sum(int32_t x, int lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    return x;
}

// The first lane j whose value e (nondecreasing over the lanes) exceeds k,
// for a k below the last lane's value.
__device__ inline int first_lane_above(int32_t e, int32_t k) {
    int lo = 0;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1)
        if (__shfl(e, lo + step - 1) <= k) lo += step;
    return lo;
}

// ---------------------------------------------------------------------
// BFS level state (device resident).  Launch p of the level kernel runs
// level k (or, in solo mode, levels k, k+1, ... in work-group 0): every
// work-group sums the per-block counters launch p-1 wrote, derives the state
// S_k from S_{k-1}, picks the direction of step k and runs it; work-group 0
// also stores the state (state[p & 1]) and the level log.
// ---------------------------------------------------------------------
constexpr int kFeat = LGMI_POLICY_FEATURES;  // per-level policy features
constexpr int kHidden = LGMI_POLICY_HIDDEN;
constexpr int kHistBins = 8;                 // log4-degree histogram of a frontier

struct BfsCtrl {
    int32_t level;          // k: the frontier holds the vertices at distance k
    int32_t dir;            // direction of step k -> k+1 (0 top-down, 1 bottom-up)
    int32_t n_front;        // |F_k|
    int32_t n_prev;         // |F_{k-1}|
    int32_t prev_dir;
    int32_t policy_mode;    // LGMI_POLICY_*
    int64_t m_front;        // sum of traversal-direction degrees over F_k
    int64_t m_unvisited;    // sum of scan-direction degrees over unvisited vertices
    int64_t n_unvisited;
    uint32_t front_hist[kHistBins];   // degree histogram of F_k (policy feature)
    float alpha, beta;
    int32_t nv;
    int32_t phase;          // launch (level) index that stored this state
    int64_t nnz_scan;
    int64_t m_light;        // traversal-direction edges of F_k's rows of at most kHeavyRow entries
    int32_t qovf;           // 1: some work-group's queue of F_k overflowed (or there is none)
    int32_t variant;        // step k: top-down in 1 range, 2 queue, 3 solo mode; 0 otherwise
    uint64_t t_main;        // wall clock at the start of level k's launch
};

// Words shared between the launches of one search and the host.
struct BfsShared {
    int32_t done_level;     // launch that found the search ended (INT32_MAX while running)
    int32_t src0;           // the source of a single-source search (-1: several)
    int32_t done_levels;    // levels of the ended search (the empty frontier's k)
    int32_t pad;
    BfsCtrl state[2];       // the state launch p stored, by p & 1
};

struct LevelRec {
    float feat[kFeat];      // features of the state at the start of the step
    int32_t dir;            // direction used by the step
    int32_t n_front;
    int64_t m_front;
    int64_t m_unvisited;
    int32_t n_next;
    int32_t variant;        // top-down: 1 range mode, 2 queue mode, 3 solo mode
    int64_t scanned;        // adjacency entries read (bottom-up) / expanded (top-down)
    int64_t cand;           // bottom-up: unvisited vertices examined
    uint64_t t_start;       // first block start (wall clock ticks)
    uint64_t t_end;         // last block end
    uint64_t t_prev_end;    // end of the previous step (launch gap included)
    int32_t launch;         // index of the launch that ran the step
    int32_t pad;
};

struct PolicyWeights {      // FP32 MLP 16 -> 64 (ReLU) -> 2
    float w1[kHidden * kFeat];
    float b1[kHidden];
    float w2[2 * kHidden];
    float b2[2];
    float mu[kFeat];        // feature standardisation
    float sd[kFeat];
};

// Per-block counters of one step (plain stores: no same-address atomics
// across work-groups); every work-group of the next level's launch sums
// them.
constexpr int kPart = 16;
enum PartSlot { P_NEXT = 0, P_OUT = 1, P_IN = 2, P_SCAN = 3, P_CAND = 4, P_HIST = 5,  // P_HIST .. +7
                P_LIGHT = 13, P_QOVF = 14 };

struct BfsWork {
    int32_t nv = 0;
    int32_t nwords = 0;
    int32_t grid = 0;
    int32_t* dist = nullptr;
    int32_t* parent = nullptr;
    int32_t* srcbuf = nullptr;   // multi-source lists
    unsigned long long* part = nullptr;  // [2][grid][kPart], by level parity
    uint32_t* bm[3] = {nullptr, nullptr, nullptr};
    uint32_t* vis = nullptr;
    int32_t* qbuf = nullptr;     // [2][grid][kQCap]: each work-group's new vertices, by level parity
    BfsShared* sh = nullptr;
    BfsShared* h_sh = nullptr;   // pinned host mirror
    int32_t* h_done = nullptr;   // pinned host-mapped [done_level, done_levels], written by the kernels
    int32_t* d_done = nullptr;   // its device address
    LevelRec* log = nullptr;
    int32_t max_log = 0;
    PolicyWeights* policy = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool time_steps = false;              // HIP events around every step launch
    hipGraphExec_t gexec[32] = {};        // captured level sequences (parents, dir_in, learned, forced)
    uint64_t policy_version = 0;          // version of the weights on the device
    std::vector<hipEvent_t> step_ev;      // launch i runs between step_ev[i] and step_ev[i+1]
    int last_launches = 0;
    double last_steps_event_ms = 0.0;
    std::vector<float> last_launch_ms;
    int last_levels = 0;
    int last_launches_used = 0;        // launch index that finished the last search
    int last_dir_in = 0;
    int last_parents = 0;
    hipStream_t own_stream = nullptr;     // the stream of a concurrent-search workspace (destroyed with it)
    ~BfsWork();
};

struct BcWork;
struct BcGraph;
// Rows longer than kHeavyRow are split into work items of kHeavyRow
// entries so that a hub vertex is expanded by many wavefronts at once.
constexpr int32_t kHeavyRow = 256;

struct DevCSR {
    int32_t* row_ptr = nullptr;
    int32_t* col = nullptr;
    uint32_t* nz_bm = nullptr;   // bit v set when the row of v is non-empty
    int32_t* first = nullptr;    // [nv]: col[row_ptr[v]], or -1 for an empty row
    int32_t* item_v = nullptr;   // heavy-row work items: vertex
    int32_t* item_beg = nullptr; //   first entry
    int32_t* item_end = nullptr; //   and one past the last (at most kHeavyRow entries)
    int32_t n_items = 0;
    // Range table of the rows longer than kHeavyRow: the vertices are cut
    // into `range_grid` ranges of range_size ids, and rtab[h][r] is the
    // first entry of heavy row h whose neighbour id is >= r * range_size.
    int32_t* heavy_id = nullptr; // [nv]: heavy row index of v, or -1
    int32_t* rtab = nullptr;     // [n_heavy][range_grid + 1]
    int32_t n_heavy = 0;
    int32_t range_grid = 0;
    int32_t range_size = 0;      // 0: no range table
};

// Work-group queues and ranges of the BFS range mode.
constexpr int kQCap = 4096;         // new vertices a work-group records per level (the range-mode frontier limit)
constexpr int kMaxGrid = 1024;      // BFS work-groups at most
constexpr int32_t kRangeMax = 24576; // vertices per work-group range at most (LDS)

}  // namespace lgmi

struct lgmi_graph {
    int32_t nv = 0;
    int64_t nnz = 0;
    int directed = 0;
    lgmi::DevCSR out;        // out-neighbours (fadjlist)
    lgmi::DevCSR in;         // in-neighbours (badjlist); aliases `out` when undirected
    int32_t* deg_total = nullptr;  // degree(g, v) as the reference defines it
    hipStream_t stream = nullptr;
    std::mutex mu;
    lgmi::BfsWork* bfs = nullptr;
    std::vector<lgmi::BfsWork*> bfs_pool;   // workspaces of lgmi_bfs_run_many, one stream each
    lgmi::BcWork* bc = nullptr;
    lgmi::BcGraph* bcg = nullptr;   // betweenness: host row pointers, heavy-row segments
    int cu_count = 0;
    int bfs_forced = -1;           // -1 policy, 0/1 forced direction (label runs)
    lgmi_bc_stats bc_last{};       // statistics of the last betweenness call
};

namespace lgmi {
// bfs.hip
void bfs_prepare(lgmi_graph* g);
void bfs_run(lgmi_graph* g, const int32_t* h_srcs, int32_t nsrc, int32_t single_src, int dir_in,
             int want_parents, int forced_dir, lgmi_bfs_stats* stats, const int32_t* h_walls = nullptr,
             int32_t nwalls = 0);
void bfs_fetch(lgmi_graph* g, int32_t* dist, int32_t* parent);
void bfs_run_many(lgmi_graph* g, const int32_t* roots, int32_t nroots, int dir_in, int32_t* dist, int32_t* parent);
void bfs_free(BfsWork* w);
// bc.hip
// mode 0: betweenness into bc; mode 1: stress centrality, Int64 results
// written into the same nv * 8 bytes; mode 2: per-source distance
// statistics into dist_stats (nsrc x 3: sum of distances, vertices reached,
// largest distance), bc unused
void bc_run(lgmi_graph* g, const int32_t* srcs, int32_t nsrc, int64_t k_scale, int skip_degree0,
            int normalize, int endpoints, double* bc, lgmi_bc_stats* stats, int mode,
            uint64_t* dist_stats = nullptr);
void bc_free(BcWork* w);
void bc_graph_free(BcGraph* b);
void bc_set_launch(int mode);
// capi.hip
void finish_csr(lgmi_graph* g, DevCSR& d, const int32_t* rp);
void graph_release(lgmi_graph* g);
// ingest.hip
lgmi_graph* graph_from_edges(const int32_t* h_src, const int32_t* h_dst, int64_t m, int directed, int32_t nv_min,
                             int64_t* ne_out);
// conn.hip
void connected_components(lgmi_graph* g, int32_t* label);
// out[i] = in[0] + ... + in[i-1] on the device; returns the total (waits for
// the stream)
int64_t exclusive_scan(hipStream_t s, const int32_t* in, int32_t* out, int32_t n);
int64_t period_from_levels(lgmi_graph* g, const int32_t* d_dist);
int32_t neighborhood(lgmi_graph* g, int32_t v, int64_t d, int dir_in, int32_t* verts, int32_t* dists);
// global policy state (host copy), set by lgmi_bfs_set_policy
struct PolicyHost {
    int mode = LGMI_POLICY_HEURISTIC;
    float alpha = 15.0f, beta = 18.0f;
    bool have_mlp = false;
    PolicyWeights w{};
    uint64_t version = 1;
};
PolicyHost& policy_host();
std::mutex& policy_mutex();
// process-wide launch thresholds (lgmi_set_tuning), read once per call
lgmi_tuning tuning_get();
}  // namespace lgmi
