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
// whose n * nnz exceeds the per-task target are split.  Tasks are ordered by
// decreasing work and claimed by persistent 1024-thread work-groups from an
// atomic counter.  A task runs in one of two modes:
//  * block mode: the whole work-group runs one single-source search at a
//    time; sigma (FP64) lives in LDS when n fits (160 KB holds 20 000
//    vertices), else in global scratch;
//  * wave mode (n <= wave_max_n, an autotuned threshold): each of the 16
//    wavefronts runs its own source with its own LDS slice, synchronising
//    only inside the wavefront.
// Per search:
//  forward, level by level: the frontier expands into unvisited neighbours,
//    claimed with a CAS on a per-vertex level stamp (seq << 12 | level, so no
//    array is cleared between sources); then every new vertex pulls sigma
//    from its neighbours on the previous level (no atomics; integer-valued
//    sums are exact in FP64);
//  backward, level by level from the deepest: every vertex pulls
//    sum (1 + delta[w]) / sigma[w] over its successors, forms delta, adds it
//    to the task's partial betweenness and stores its own coefficient in the
//    sigma slot.
// Partial arrays are per task (per wavefront in wave mode, folded in a fixed
// order), and a second kernel adds the tasks of a graph in task order and
// applies _rescale!, so results are deterministic.
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
constexpr int kLevelBits = 12;      // stamp = seq << 12 | level
constexpr uint32_t kLevelMask = (1u << kLevelBits) - 1;

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
    int32_t lds_max_n;        // sigma fits in LDS for n <= this (block mode)
    int32_t max_n;
    double* part;             // task partial betweenness
    int32_t* task_ctr;
    // per-block scratch (global)
    uint32_t* stamp;          // [blocks][kBW][scr]
    int32_t* queue;           // [blocks][kBW][scr]
    int32_t* loff;            // [blocks][kBW][scr + 2]
    double* gsig;             // [blocks][max_n] (block mode without LDS room)
    double* wpart;            // [blocks][kBW][wave_max_n]
    int64_t scr;              // per-wavefront scratch length (>= max_n / kBW and >= wave_max_n)
    int32_t* err;             // set when a search exceeds the level range
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

// One single-source Brandes search by a team.  `t` = thread index in the
// team.  qcount: LDS counter (block mode) or unused (wave mode uses ballot).
template <int TEAM>
__device__ void brandes_source(const int32_t* rp, const int32_t* col, int32_t n, int32_t s, uint32_t seq,
                               double* sig, uint32_t* stamp, int32_t* queue, int32_t* loff, double* part,
                               int* qcount, int t, int32_t* err) {
    const uint32_t base = seq << kLevelBits;
    if (t == 0) {
        stamp[s] = base;
        sig[s] = 1.0;
        queue[0] = s;
        loff[0] = 0;
        loff[1] = 1;
        if (TEAM == kBB) *qcount = 1;
    }
    team_sync<TEAM>();
    int32_t L = 0;
    int32_t qn = 1;                  // queue length (team-uniform after sync)
    for (;;) {
        const int32_t lo = loff[L], hi = qn;
        if (lo == hi) break;
        if (L + 1 >= (int32_t)kLevelMask) {
            if (t == 0) atomicOr(err, 1);
            break;
        }
        const uint32_t next = base | (uint32_t)(L + 1);
        // discovery: claim unvisited neighbours of the frontier
        int32_t my_new = 0;
        if (TEAM == kBB) {
            for (int32_t i = lo + t; i < hi; i += TEAM) {
                const int32_t u = queue[i];
                for (int32_t e = rp[u]; e < rp[u + 1]; e++) {
                    const int32_t v = col[e];
                    uint32_t old = stamp[v];
                    if ((old >> kLevelBits) != seq) {
                        if (__hip_atomic_compare_exchange_strong(&stamp[v], &old, next, __ATOMIC_RELAXED,
                                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
                            const int pos = atomicAdd(qcount, 1);
                            queue[pos] = v;
                        }
                    }
                }
            }
            team_sync<TEAM>();
            qn = *qcount;
        } else {
            // wave mode: ballot-aggregated appends, no shared counter
            const int lane = t;
            int32_t cnt = hi;
            for (int32_t i0 = lo; i0 < hi; i0 += 64) {
                const int32_t i = i0 + lane;
                const int32_t u = i < hi ? queue[i] : -1;
                const int32_t eb = u >= 0 ? rp[u] : 0, ee = u >= 0 ? rp[u + 1] : 0;
                int32_t e = eb;
                for (;;) {
                    bool claimed = false;
                    int32_t v = -1;
                    if (e < ee) {
                        v = col[e];
                        uint32_t old = stamp[v];
                        if ((old >> kLevelBits) != seq)
                            claimed = __hip_atomic_compare_exchange_strong(&stamp[v], &old, next, __ATOMIC_RELAXED,
                                                                           __ATOMIC_RELAXED,
                                                                           __HIP_MEMORY_SCOPE_WAVEFRONT);
                        e++;
                    }
                    const uint64_t m = __ballot(claimed);
                    if (m) {
                        if (claimed) queue[cnt + __popcll(m & ((1ull << lane) - 1ull))] = v;
                        cnt += __popcll(m);
                    }
                    if (__ballot(e < ee) == 0) break;
                }
            }
            qn = cnt;
            team_sync<TEAM>();
        }
        (void)my_new;
        // sigma of the new level: pull from neighbours on level L
        const uint32_t cur = base | (uint32_t)L;
        for (int32_t i = hi + t; i < qn; i += TEAM) {
            const int32_t v = queue[i];
            double sv = 0.0;
            for (int32_t e = rp[v]; e < rp[v + 1]; e++) {
                const int32_t u = col[e];
                if (stamp[u] == cur) sv += sig[u];
            }
            sig[v] = sv;
        }
        if (t == 0) loff[L + 2] = qn;
        L++;
        team_sync<TEAM>();
    }
    // backward, deepest level first; level 0 (the source) adds nothing
    for (int32_t l = L; l >= 1; l--) {
        const int32_t lo = loff[l], hi = loff[l + 1];
        const uint32_t succ = base | (uint32_t)(l + 1);
        for (int32_t i = lo + t; i < hi; i += TEAM) {
            const int32_t v = queue[i];
            double acc = 0.0;
            for (int32_t e = rp[v]; e < rp[v + 1]; e++) {
                const int32_t w = col[e];
                if (stamp[w] == succ) acc += sig[w];
            }
            const double sv = sig[v];
            const double d = sv * acc;
            part[v] += d;
            sig[v] = (1.0 + d) / sv;
        }
        team_sync<TEAM>();
    }
}

__global__ __launch_bounds__(kBB) void bc_batch_kernel(BatchArgs a) {
    extern __shared__ double s_sig[];
    __shared__ int s_task;
    __shared__ int s_qcount;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    uint32_t* stamp_b = a.stamp + (size_t)blockIdx.x * kBW * a.scr;
    int32_t* queue_b = a.queue + (size_t)blockIdx.x * kBW * a.scr;
    int32_t* loff_b = a.loff + (size_t)blockIdx.x * kBW * (a.scr + 2);
    for (;;) {
        if (tid == 0) s_task = atomicAdd(a.task_ctr, 1);
        __syncthreads();
        const int task = s_task;
        __syncthreads();
        if (task >= a.ntasks) break;
        const int32_t gi = a.task_graph[task];
        const int32_t n = a.nv[gi];
        const int32_t s0 = a.task_s0[task], s1 = a.task_s1[task];
        const int32_t* rp = a.rp_cat + a.row_off[gi];
        const int32_t* col = a.col_cat + a.col_off[gi];
        double* part = a.part + a.task_part[task];
        if (n <= a.wave_max_n) {
            // ---------------- wave mode ----------------
            double* sig = s_sig + (size_t)wid * a.wave_max_n;
            uint32_t* stamp = stamp_b + (size_t)wid * a.scr;
            int32_t* queue = queue_b + (size_t)wid * a.scr;
            int32_t* loff = loff_b + (size_t)wid * (a.scr + 2);
            double* wp = a.wpart + ((size_t)blockIdx.x * kBW + wid) * a.wave_max_n;
            for (int32_t v = lane; v < n; v += 64) { stamp[v] = 0; wp[v] = 0.0; }
            team_sync<64>();
            uint32_t seq = 0;
            for (int32_t s = s0 + wid; s < s1; s += kBW) {
                seq++;
                brandes_source<64>(rp, col, n, s, seq, sig, stamp, queue, loff, wp, nullptr, lane, a.err);
            }
            __syncthreads();
            // fold the wavefront partials into the task partial, fixed order
            for (int32_t v = tid; v < n; v += kBB) {
                double acc = 0.0;
                for (int q = 0; q < kBW; q++) acc += a.wpart[((size_t)blockIdx.x * kBW + q) * a.wave_max_n + v];
                part[v] = acc;
            }
            __syncthreads();
        } else {
            // ---------------- block mode ----------------
            double* sig = n <= a.lds_max_n ? s_sig : a.gsig + (size_t)blockIdx.x * a.max_n;
            for (int32_t v = tid; v < n; v += kBB) { stamp_b[v] = 0; part[v] = 0.0; }
            __syncthreads();
            uint32_t seq = 0;
            for (int32_t s = s0; s < s1; s++) {
                seq++;
                brandes_source<kBB>(rp, col, n, s, seq, sig, stamp_b, queue_b, loff_b, part, &s_qcount, tid, a.err);
            }
            __syncthreads();
        }
    }
}

// bc of every graph: sum of its tasks' partials in task order, endpoints
// term, _rescale! scale.  One thread per output vertex.
__global__ void bc_batch_finish(int32_t ngraphs, const int32_t* nv, const int64_t* bc_off, const int32_t* gtask0,
                                const int32_t* gtask_n, const int32_t* gtasks, const int64_t* task_part,
                                const double* part, const double* scale, int endpoints, double* out, int64_t total) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        // find the graph of output i (binary search over bc_off)
        int lo = 0, hi = ngraphs - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (bc_off[mid] <= i) lo = mid; else hi = mid - 1;
        }
        const int g = lo;
        const int64_t v = i - bc_off[g];
        double s = 0.0;
        for (int k = 0; k < gtask_n[g]; k++) s += part[task_part[gtasks[gtask0[g] + k]] + v];
        if (endpoints) {
            // every vertex of a graph with an edge is a processed source:
            // it adds n-1 to itself and 1 to each other vertex
            s += 0.0;
        }
        out[i] = s * scale[g];
    }
}

struct BatchTuning {
    int32_t wave_max_n = -1;        // -1: pick from LDS capacity
    double task_work_target = 0.0;  // 0: automatic
};

BatchTuning& tuning() {
    static BatchTuning t;
    return t;
}

}  // namespace
}  // namespace lgmi

using namespace lgmi;

extern "C" int lgmi_betweenness_batch_tuning(int32_t wave_max_n, double task_work_target) {
    try {
        if (wave_max_n < -1 || task_work_target < 0) throw Error(LGMI_E_ARG, "bad tuning values");
        tuning().wave_max_n = wave_max_n;
        tuning().task_work_target = task_work_target;
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
        const int32_t lds_max_n = (int32_t)(dyn_lds / sizeof(double));
        int32_t wave_max_n = tuning().wave_max_n >= 0 ? tuning().wave_max_n : (int32_t)std::min<int64_t>(
            lds_max_n / kBW, 1024);
        wave_max_n = std::min<int32_t>(wave_max_n, lds_max_n / kBW);
        (void)static_lds;

        // ---- tasks: split graphs by work, largest first ----
        std::vector<double> work((size_t)ngraphs);
        double total_work = 0.0;
        for (int32_t g = 0; g < ngraphs; g++) {
            const int32_t* rp = row_ptr_cat + row_off[g];
            work[(size_t)g] = (double)nv[g] * (double)(rp[nv[g]] + nv[g]);
            total_work += work[(size_t)g];
        }
        const int nblocks = std::max(1, cu);
        const double target = tuning().task_work_target > 0 ? tuning().task_work_target
                                                             : std::max(1.0, total_work / ((double)nblocks * 6.0));
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
        std::vector<int32_t> order(tasks.size());
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return tasks[x].w > tasks[y].w; });
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
            a.task_ctr = (int32_t*)dalloc(sizeof(int32_t) * 2);
            a.err = a.task_ctr + 1;
            LGMI_HIP(hipMemsetAsync(a.task_ctr, 0, sizeof(int32_t) * 2, st));
            a.scr = std::max<int64_t>(max_n, 1);
            // block mode uses wavefront 0's slice for the whole block
            a.stamp = (uint32_t*)dalloc(sizeof(uint32_t) * (size_t)nblocks * kBW * a.scr);
            a.queue = (int32_t*)dalloc(sizeof(int32_t) * (size_t)nblocks * kBW * a.scr);
            a.loff = (int32_t*)dalloc(sizeof(int32_t) * (size_t)nblocks * kBW * (a.scr + 2));
            a.gsig = (double*)dalloc(sizeof(double) * (size_t)nblocks * std::max(max_n, 1));
            a.wpart = (double*)dalloc(sizeof(double) * (size_t)nblocks * kBW * std::max(wave_max_n, 1));
            double* d_out = (double*)dalloc(sizeof(double) * (size_t)std::max<int64_t>(out_total, 1));
            const int64_t* d_bcoff = (const int64_t*)up(bc_off, sizeof(int64_t) * ngraphs);
            const int32_t* d_gt0 = (const int32_t*)up(gtask0.data(), sizeof(int32_t) * ngraphs);
            const int32_t* d_gtn = (const int32_t*)up(gtask_n.data(), sizeof(int32_t) * ngraphs);
            const int32_t* d_gts = (const int32_t*)up(gtasks.data(), sizeof(int32_t) * std::max(ntasks, 1));
            const double* d_scale = (const double*)up(scale.data(), sizeof(double) * ngraphs);
            LGMI_HIP(hipStreamSynchronize(st));
            LGMI_HIP(hipFuncSetAttribute((const void*)bc_batch_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)dyn_lds));
            LGMI_HIP(hipEventCreate(&e0));
            LGMI_HIP(hipEventCreate(&e1));
            LGMI_HIP(hipEventRecord(e0, st));
            bc_batch_kernel<<<nblocks, kBB, dyn_lds, st>>>(a);
            LGMI_HIP(hipGetLastError());
            int32_t total_out = (int32_t)std::min<int64_t>(out_total, 1 << 30);
            bc_batch_finish<<<std::max(1, std::min(cu * 8, (int)((out_total + 255) / 256))), 256, 0, st>>>(
                ngraphs, a.nv, d_bcoff, d_gt0, d_gtn, d_gts, a.task_part, a.part, d_scale, 0, d_out, out_total);
            (void)total_out;
            LGMI_HIP(hipGetLastError());
            LGMI_HIP(hipEventRecord(e1, st));
            int32_t herr = 0;
            LGMI_HIP(hipMemcpyAsync(&herr, a.err, sizeof(int32_t), hipMemcpyDeviceToHost, st));
            LGMI_HIP(hipMemcpyAsync(bc_cat, d_out, sizeof(double) * (size_t)out_total, hipMemcpyDeviceToHost, st));
            LGMI_HIP(hipStreamSynchronize(st));
            if (herr) throw Error(LGMI_E_HIP, "a search exceeded 4095 levels");
            float ms = 0.f;
            LGMI_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (device_ms) *device_ms = ms;
        } catch (...) {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            for (void* p : bufs) dev_free(p);
            (void)hipStreamDestroy(st);
            throw;
        }
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
