// capi.hip -- graph handle, error reporting and the extern "C" entry points
// of liblgmi.so (declared in include/lgmi.h).
#include "lgmi_internal.h"

#include <thread>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace lgmi {

namespace {
thread_local std::string t_last_error;

__global__ void nz_bitmap_kernel(const int32_t* rp, int32_t nv, int32_t nwords, uint32_t* nz) {
    const int32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwords) return;
    uint32_t bits = 0;
    for (int b = 0; b < 32; b++) {
        const int64_t v = (int64_t)w * 32 + b;
        if (v < nv && rp[v + 1] > rp[v]) bits |= 1u << b;
    }
    nz[w] = bits;
}

// first[v] = the first (smallest) neighbour of v, -1 for an empty row: the
// entry a bottom-up step tests first, read from one contiguous array.
__global__ void first_entry_kernel(const int32_t* rp, const int32_t* col, int32_t nv, int32_t* first) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const int32_t b = rp[v];
    first[v] = rp[v + 1] > b ? col[b] : -1;
}

// rtab[h][r] = first entry of heavy row h with neighbour id >= r * rs (a
// binary search over the sorted row); one thread per (row, boundary).
__global__ void range_table_kernel(const int32_t* rp, const int32_t* col, const int32_t* heavy_v, int32_t n_heavy,
                                   int32_t ngrid, int32_t rs, int32_t* tab) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n_heavy * (ngrid + 1)) return;
    const int32_t h = (int32_t)(i / (ngrid + 1)), r = (int32_t)(i % (ngrid + 1));
    const int32_t v = heavy_v[h];
    int32_t lo = rp[v], hi = rp[v + 1];
    const int64_t key = (int64_t)r * rs;
    while (lo < hi) {
        const int32_t mid = lo + (hi - lo) / 2;
        if ((int64_t)col[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    tab[i] = lo;
}

// Counts the entries (v, u) of a CSR whose mirror is missing: v must be in
// row u of `orp`/`ocol` (the in-CSR of a directed graph).  A wavefront per
// row, one entry per lane, a binary search of v in the sorted row u.  For an
// undirected graph (upper = 1, orp/ocol the CSR itself) only the entries
// with u > v are searched, and the entries above and below the diagonal are
// counted (cnt[1], cnt[2]): all upper mirrors present and equal counts make
// the lower entries exactly those mirrors.  cnt[0] gets the misses.
__global__ void mirror_check_kernel(const int32_t* rp, const int32_t* col, int32_t nv, const int32_t* orp,
                                    const int32_t* ocol, int upper, unsigned long long* cnt) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    unsigned long long miss = 0, above = 0, below = 0;
    for (int64_t v = wave; v < nv; v += nwaves) {
        const int32_t b0 = rp[v], e0 = rp[v + 1];
        for (int32_t k = b0 + lane; k < e0; k += 64) {
            const int32_t u = col[k];
            if (upper) {
                if (u < (int32_t)v) { below++; continue; }
                if (u == (int32_t)v) continue;
                above++;
            }
            int32_t lo = orp[u], hi = orp[u + 1];
            const int32_t end = hi;
            while (lo < hi) {
                const int32_t m = lo + (hi - lo) / 2;
                if (ocol[m] < (int32_t)v) lo = m + 1;
                else hi = m;
            }
            if (lo >= end || ocol[lo] != (int32_t)v) miss++;
        }
    }
    if (miss) atomicAdd(&cnt[0], miss);
    if (above) atomicAdd(&cnt[1], above);
    if (below) atomicAdd(&cnt[2], below);
}

int g_device_count = -1;
std::mutex g_dev_mu;
}  // namespace

void set_last_error(const std::string& msg) { t_last_error = msg; }

void* dev_alloc(size_t bytes) {
    const char* f = std::getenv("LGMI_FAIL_ALLOC");
    if (f && f[0] == '1') throw Error(LGMI_E_ALLOC, "device allocation refused (LGMI_FAIL_ALLOC=1)");
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(LGMI_E_ALLOC, std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    return p;
}

void dev_free(void* p) {
    if (p) (void)hipFree(p);
}

void ensure_device() {
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if (g_device_count < 0) {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        (void)hipGetLastError();
        g_device_count = n;
    }
    if (g_device_count <= 0) throw Error(LGMI_E_NODEV, "no GPU device visible");
}

int device_cu_count() {
    int dev = 0;
    LGMI_HIP(hipGetDevice(&dev));
    int cu = 0;
    LGMI_HIP(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
    return cu;
}

namespace {
lgmi_tuning tuning_defaults() {
    lgmi_tuning t{};
    t.wave_row = 32;
    t.lane_rounds = 4;
    t.range_front = 2048;
    t.queue_front = 4096;
    t.range_light = 32768;
    t.queue_min_edges = 4096;
    t.queue_max_edges = 262144;
    t.bc_seg_len = 1024;
    t.solo_front = 4096;
    t.solo_edges = 2048;
    return t;
}
std::mutex g_tune_mu;
lgmi_tuning g_tune = tuning_defaults();
}  // namespace

lgmi_tuning tuning_get() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    return g_tune;
}

PolicyHost& policy_host() {
    static PolicyHost p;
    return p;
}
std::mutex& policy_mutex() {
    static std::mutex m;
    return m;
}

// Host check of a CSR: row_ptr from 0 to nnz and monotone, neighbour ids in
// range, rows sorted ascending without duplicates.  Large graphs are checked
// by up to 16 threads over vertex ranges; the violation reported is the one
// at the smallest vertex, as a single pass would report it.
static void check_csr(int32_t nv, int64_t nnz, const int32_t* rp, const int32_t* col, const char* what) {
    if (!rp || (nnz > 0 && !col)) throw Error(LGMI_E_ARG, std::string(what) + ": null array");
    if (rp[0] != 0 || (int64_t)rp[nv] != nnz) throw Error(LGMI_E_ARG, std::string(what) + ": row_ptr[0] must be 0 and row_ptr[nv] == nnz");
    // 0 ok, 1 not monotone, 2 id out of range, 3 unsorted or duplicate; the first violation of [v0, v1)
    auto scan = [&](int32_t v0, int32_t v1, int32_t& at) -> int {
        for (int32_t v = v0; v < v1; v++) {
            const int32_t b = rp[v], e = rp[v + 1];
            at = v;
            if (e < b || b < 0 || (int64_t)e > nnz) return 1;
            for (int32_t k = b; k < e; k++) {
                const int32_t u = col[k];
                if (u < 0 || u >= nv) return 2;
                if (k > b && col[k - 1] >= u) return 3;
            }
        }
        return 0;
    };
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = (int)std::min<int64_t>({16, (int64_t)hw, nnz / (1 << 20) + 1, (int64_t)std::max(nv, 1)});
    std::vector<int> code((size_t)nt, 0);
    std::vector<int32_t> at((size_t)nt, 0);
    if (nt <= 1) {
        code[0] = scan(0, nv, at[0]);
    } else {
        // vertex ranges of about equal entry counts: thread t starts at the
        // first vertex whose row begins at or after t * nnz / nt (row_ptr is
        // only trusted after the monotonicity check, so the split points are
        // clamped into order; a wrong split costs balance, not correctness)
        std::vector<int32_t> cut((size_t)nt + 1, 0);
        cut[(size_t)nt] = nv;
        for (int t = 1; t < nt; t++) {
            const int64_t target = nnz * t / nt;
            int32_t lo = cut[(size_t)t - 1], hi = nv;
            while (lo < hi) {
                const int32_t mid = lo + (hi - lo) / 2;
                if ((int64_t)rp[mid] < target) lo = mid + 1;
                else hi = mid;
            }
            cut[(size_t)t] = lo;
        }
        std::vector<std::thread> th;
        th.reserve((size_t)nt);
        for (int t = 0; t < nt; t++)
            th.emplace_back([&, t]() { code[(size_t)t] = scan(cut[(size_t)t], cut[(size_t)t + 1], at[(size_t)t]); });
        for (auto& x : th) x.join();
    }
    for (int t = 0; t < nt; t++) {
        if (code[(size_t)t] == 1) throw Error(LGMI_E_ARG, std::string(what) + ": row_ptr not monotone");
        if (code[(size_t)t] == 2) throw Error(LGMI_E_BOUNDS, std::string(what) + ": neighbour id out of range");
        if (code[(size_t)t] == 3) throw Error(LGMI_E_ARG, std::string(what) + ": rows must be sorted ascending without duplicates");
    }
}

static void upload_csr(lgmi_graph* g, DevCSR& d, const int32_t* rp, const int32_t* col, int64_t nnz) {
    d.row_ptr = dev_alloc_n<int32_t>((size_t)g->nv + 1);
    d.col = dev_alloc_n<int32_t>((size_t)std::max<int64_t>(nnz, 1));
    LGMI_HIP(hipMemcpyAsync(d.row_ptr, rp, sizeof(int32_t) * ((size_t)g->nv + 1), hipMemcpyHostToDevice, g->stream));
    if (nnz > 0) LGMI_HIP(hipMemcpyAsync(d.col, col, sizeof(int32_t) * (size_t)nnz, hipMemcpyHostToDevice, g->stream));
    finish_csr(g, d, rp);
}

// The derived structures of a CSR whose row_ptr and col are on the device
// (rp: a host copy of row_ptr): non-empty-row bitmap, heavy-row items and
// the range table of the BFS.
void finish_csr(lgmi_graph* g, DevCSR& d, const int32_t* rp) {
    const int32_t nwords = 2 * ((g->nv + 63) / 64);
    d.nz_bm = dev_alloc_n<uint32_t>((size_t)std::max(nwords, 1));
    if (nwords > 0) {
        nz_bitmap_kernel<<<(nwords + 255) / 256, 256, 0, g->stream>>>(d.row_ptr, g->nv, nwords, d.nz_bm);
        LGMI_HIP(hipGetLastError());
    }
    d.first = dev_alloc_n<int32_t>((size_t)std::max(g->nv, 1));
    if (g->nv > 0) {
        first_entry_kernel<<<(unsigned)((g->nv + 255) / 256), 256, 0, g->stream>>>(d.row_ptr, d.col, g->nv, d.first);
        LGMI_HIP(hipGetLastError());
    }
    // rows longer than kHeavyRow -> items of at most kHeavyRow entries
    std::vector<int32_t> iv, ib, ie;
    for (int32_t v = 0; v < g->nv; v++) {
        const int32_t b = rp[v], e = rp[v + 1];
        if (e - b > kHeavyRow)
            for (int32_t x = b; x < e; x += kHeavyRow) {
                iv.push_back(v);
                ib.push_back(x);
                ie.push_back(std::min(e, x + kHeavyRow));
            }
    }
    if (iv.size() > (size_t)0x7fffffff) throw Error(LGMI_E_TOO_LARGE, "too many heavy-row items");
    // range table for the BFS range mode: one range of vertex ids per
    // work-group (the BFS grid is one work-group per CU), when a range fits LDS
    {
        const int32_t ngrid = std::min(std::max(1, g->cu_count), kMaxGrid);
        int64_t rs = ((int64_t)g->nv + ngrid - 1) / ngrid;
        rs = (rs + 63) / 64 * 64;
        std::vector<int32_t> hv, hid((size_t)g->nv, -1);
        for (int32_t v = 0; v < g->nv; v++)
            if (rp[v + 1] - rp[v] > kHeavyRow) { hid[(size_t)v] = (int32_t)hv.size(); hv.push_back(v); }
        if (g->nv > 0 && rs <= kRangeMax && !hv.empty()) {
            d.range_grid = ngrid;
            d.range_size = (int32_t)rs;
            d.n_heavy = (int32_t)hv.size();
            d.heavy_id = dev_alloc_n<int32_t>((size_t)g->nv);
            d.rtab = dev_alloc_n<int32_t>((size_t)d.n_heavy * (ngrid + 1));
            int32_t* dhv = dev_alloc_n<int32_t>(hv.size());
            try {
                LGMI_HIP(hipMemcpyAsync(d.heavy_id, hid.data(), sizeof(int32_t) * hid.size(), hipMemcpyHostToDevice, g->stream));
                LGMI_HIP(hipMemcpyAsync(dhv, hv.data(), sizeof(int32_t) * hv.size(), hipMemcpyHostToDevice, g->stream));
                const int64_t n = (int64_t)d.n_heavy * (ngrid + 1);
                range_table_kernel<<<(unsigned)((n + 255) / 256), 256, 0, g->stream>>>(d.row_ptr, d.col, dhv, d.n_heavy,
                                                                                      ngrid, d.range_size, d.rtab);
                LGMI_HIP(hipGetLastError());
                LGMI_HIP(hipStreamSynchronize(g->stream));
            } catch (...) {
                dev_free(dhv);
                throw;
            }
            dev_free(dhv);
        }
    }
    d.n_items = (int32_t)iv.size();
    if (d.n_items > 0) {
        d.item_v = dev_alloc_n<int32_t>(iv.size());
        d.item_beg = dev_alloc_n<int32_t>(ib.size());
        d.item_end = dev_alloc_n<int32_t>(ie.size());
        LGMI_HIP(hipMemcpyAsync(d.item_v, iv.data(), sizeof(int32_t) * iv.size(), hipMemcpyHostToDevice, g->stream));
        LGMI_HIP(hipMemcpyAsync(d.item_beg, ib.data(), sizeof(int32_t) * ib.size(), hipMemcpyHostToDevice, g->stream));
        LGMI_HIP(hipMemcpyAsync(d.item_end, ie.data(), sizeof(int32_t) * ie.size(), hipMemcpyHostToDevice, g->stream));
        LGMI_HIP(hipStreamSynchronize(g->stream));   // the host vectors go out of scope
    }
}

static void graph_free(lgmi_graph* g) {
    if (!g) return;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    bfs_free(g->bfs);
    for (BfsWork* w : g->bfs_pool) {
        if (w->own_stream) (void)hipStreamSynchronize(w->own_stream);
        bfs_free(w);
    }
    g->bfs_pool.clear();
    bc_free(g->bc);
    bc_graph_free(g->bcg);
    if (g->in.row_ptr != g->out.row_ptr) {
        dev_free(g->in.row_ptr); dev_free(g->in.col); dev_free(g->in.nz_bm); dev_free(g->in.first);
        dev_free(g->in.item_v); dev_free(g->in.item_beg); dev_free(g->in.item_end);
        dev_free(g->in.heavy_id); dev_free(g->in.rtab);
    }
    dev_free(g->out.row_ptr); dev_free(g->out.col); dev_free(g->out.nz_bm); dev_free(g->out.first);
    dev_free(g->out.item_v); dev_free(g->out.item_beg); dev_free(g->out.item_end);
    dev_free(g->out.heavy_id); dev_free(g->out.rtab);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

void graph_release(lgmi_graph* g) { graph_free(g); }

}  // namespace lgmi

using namespace lgmi;

#define LGMI_API_BEGIN try {
#define LGMI_API_END                                                  \
    }                                                                 \
    catch (const Error& e) {                                          \
        set_last_error(e.what());                                     \
        return e.code;                                                \
    }                                                                 \
    catch (const std::bad_alloc&) {                                   \
        set_last_error("host allocation failed");                     \
        return LGMI_E_ALLOC;                                          \
    }                                                                 \
    catch (const std::exception& e) {                                 \
        set_last_error(e.what());                                     \
        return LGMI_E_HIP;                                            \
    }                                                                 \
    return LGMI_OK;

extern "C" {

const char* lgmi_last_error(void) { return t_last_error.c_str(); }

int lgmi_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) { (void)hipGetLastError(); return 0; }
    return n;
}

int lgmi_synchronize(void) {
    LGMI_API_BEGIN
    ensure_device();
    LGMI_HIP(hipDeviceSynchronize());
    LGMI_API_END
}

int lgmi_graph_create(int32_t nv, int64_t nnz, const int32_t* row_ptr, const int32_t* col,
                      int64_t in_nnz, const int32_t* in_row_ptr, const int32_t* in_col,
                      int directed, lgmi_graph_t* out) {
    LGMI_API_BEGIN
    if (!out) throw Error(LGMI_E_ARG, "out is NULL");
    *out = nullptr;
    if (nv < 0) throw Error(LGMI_E_ARG, "nv < 0");
    if (nnz > 0x7fffffffLL || in_nnz > 0x7fffffffLL) throw Error(LGMI_E_TOO_LARGE, "more than 2^31-1 adjacency entries");
    if (nnz < 0) throw Error(LGMI_E_ARG, "nnz < 0");
    check_csr(nv, nnz, row_ptr, col, "out-CSR");
    if (directed) {
        if (!in_row_ptr) throw Error(LGMI_E_ARG, "a directed graph needs its in-neighbour CSR");
        if (in_nnz != nnz) throw Error(LGMI_E_ARG, "in_nnz must equal nnz");
        check_csr(nv, in_nnz, in_row_ptr, in_col, "in-CSR");
    }
    ensure_device();
    lgmi_graph* g = new lgmi_graph();
    try {
        g->nv = nv;
        g->nnz = nnz;
        g->directed = directed ? 1 : 0;
        LGMI_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
        g->cu_count = device_cu_count();
        upload_csr(g, g->out, row_ptr, col, nnz);
        if (directed) upload_csr(g, g->in, in_row_ptr, in_col, in_nnz);
        else g->in = g->out;
        // an undirected CSR must be symmetric and a directed graph's in-CSR
        // the transpose of its out-CSR (the bottom-up BFS steps, the range
        // table and betweenness read the reverse adjacency)
        if (nnz > 0) {
            unsigned long long* d_cnt = dev_alloc_n<unsigned long long>(3);
            unsigned long long cnt[3] = {0, 0, 0};
            try {
                LGMI_HIP(hipMemsetAsync(d_cnt, 0, sizeof(cnt), g->stream));
                const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(((int64_t)nv + 3) / 4,
                                                                                (int64_t)g->cu_count * 32));
                mirror_check_kernel<<<blocks, 256, 0, g->stream>>>(g->out.row_ptr, g->out.col, nv, g->in.row_ptr,
                                                                   g->in.col, directed ? 0 : 1, d_cnt);
                LGMI_HIP(hipGetLastError());
                LGMI_HIP(hipMemcpyAsync(cnt, d_cnt, sizeof(cnt), hipMemcpyDeviceToHost, g->stream));
                LGMI_HIP(hipStreamSynchronize(g->stream));
            } catch (...) {
                dev_free(d_cnt);
                throw;
            }
            dev_free(d_cnt);
            const unsigned long long bad = cnt[0] + (cnt[1] > cnt[2] ? cnt[1] - cnt[2] : cnt[2] - cnt[1]);
            if (bad)
                throw Error(LGMI_E_ARG, directed ? "the in-CSR is not the transpose of the out-CSR ("
                                                       + std::to_string(bad) + " entries unmatched)"
                                                 : "an undirected graph's CSR must be symmetric ("
                                                       + std::to_string(bad) + " entries unmatched)");
        }
        LGMI_HIP(hipStreamSynchronize(g->stream));
    } catch (...) {
        graph_free(g);
        throw;
    }
    *out = g;
    LGMI_API_END
}

void lgmi_graph_destroy(lgmi_graph_t g) { graph_free(g); }

int lgmi_graph_from_edges(const int32_t* src, const int32_t* dst, int64_t m, int directed, int32_t nv_min,
                          lgmi_graph_t* out, int64_t* ne) {
    LGMI_API_BEGIN
    if (!out || !ne || m < 0 || (m > 0 && (!src || !dst)) || nv_min < 0) throw Error(LGMI_E_ARG, "bad arguments");
    *out = nullptr;
    ensure_device();
    *out = graph_from_edges(src, dst, m, directed, nv_min, ne);
    LGMI_API_END
}

int lgmi_graph_size(lgmi_graph_t g, int32_t* nv, int64_t* nnz) {
    LGMI_API_BEGIN
    if (!g || !nv || !nnz) throw Error(LGMI_E_ARG, "bad arguments");
    *nv = g->nv;
    *nnz = g->nnz;
    LGMI_API_END
}

int lgmi_graph_csr(lgmi_graph_t g, int32_t* row_ptr, int32_t* col, int32_t* in_row_ptr, int32_t* in_col) {
    LGMI_API_BEGIN
    if (!g || !row_ptr || !col) throw Error(LGMI_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(g->mu);
    LGMI_HIP(hipMemcpy(row_ptr, g->out.row_ptr, sizeof(int32_t) * ((size_t)g->nv + 1), hipMemcpyDeviceToHost));
    if (g->nnz > 0) LGMI_HIP(hipMemcpy(col, g->out.col, sizeof(int32_t) * (size_t)g->nnz, hipMemcpyDeviceToHost));
    if (in_row_ptr && in_col) {
        LGMI_HIP(hipMemcpy(in_row_ptr, g->in.row_ptr, sizeof(int32_t) * ((size_t)g->nv + 1), hipMemcpyDeviceToHost));
        if (g->nnz > 0) LGMI_HIP(hipMemcpy(in_col, g->in.col, sizeof(int32_t) * (size_t)g->nnz, hipMemcpyDeviceToHost));
    }
    LGMI_API_END
}

static void check_src(lgmi_graph_t g, int32_t s) {
    if (s < 0 || s >= g->nv) throw Error(LGMI_E_BOUNDS, "source vertex " + std::to_string(s) + " out of range");
}

int lgmi_gdistances(lgmi_graph_t g, const int32_t* srcs, int32_t nsrc, int32_t* dist) {
    LGMI_API_BEGIN
    if (!g || !dist || nsrc < 0 || (nsrc > 0 && !srcs)) throw Error(LGMI_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(g->mu);
    std::vector<int32_t> s(srcs, srcs + nsrc);
    for (int32_t x : s) check_src(g, x);
    std::sort(s.begin(), s.end());
    s.erase(std::unique(s.begin(), s.end()), s.end());
    if (s.size() == 1) bfs_run(g, nullptr, 1, s[0], 0, 0, -1, nullptr);
    else bfs_run(g, s.data(), (int32_t)s.size(), -1, 0, 0, -1, nullptr);
    bfs_fetch(g, dist, nullptr);
    LGMI_API_END
}

int lgmi_bfs_parents(lgmi_graph_t g, int32_t src, int dir_in, int32_t* parent) {
    LGMI_API_BEGIN
    if (!g || !parent) throw Error(LGMI_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(g->mu);
    check_src(g, src);
    bfs_run(g, nullptr, 1, src, dir_in ? 1 : 0, 1, -1, nullptr);
    bfs_fetch(g, nullptr, parent);
    LGMI_API_END
}

int lgmi_bfs(lgmi_graph_t g, int32_t src, int dir_in, int32_t* dist, int32_t* parent) {
    LGMI_API_BEGIN
    if (!g) throw Error(LGMI_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(g->mu);
    check_src(g, src);
    bfs_run(g, nullptr, 1, src, dir_in ? 1 : 0, parent ? 1 : 0, -1, nullptr);
    bfs_fetch(g, dist, parent);
    LGMI_API_END
}

int lgmi_bfs_run(lgmi_graph_t g, int32_t src, int dir_in, int want_parents, lgmi_bfs_stats* stats) {
    LGMI_API_BEGIN
    if (!g) throw Error(LGMI_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(g->mu);
    check_src(g, src);
    const int forced = g->bfs ? g->bfs_forced : -1;
    bfs_run(g, nullptr, 1, src, dir_in ? 1 : 0, want_parents ? 1 : 0, forced, stats);
    LGMI_API_END
}

int lgmi_bfs_run_many(lgmi_graph_t g, const int32_t* roots, int32_t nroots, int dir_in, int32_t* dist,
                      int32_t* parent) {
    LGMI_API_BEGIN
    if (!g || nroots < 0 || (nroots > 0 && (!roots || !dist))) throw Error(LGMI_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(g->mu);
    for (int32_t i = 0; i < nroots; i++) check_src(g, roots[i]);
    bfs_run_many(g, roots, nroots, dir_in ? 1 : 0, dist, parent);
    LGMI_API_END
}

int lgmi_bfs_fetch(lgmi_graph_t g, int32_t* dist, int32_t* parent) {
    LGMI_API_BEGIN
    if (!g) throw Error(LGMI_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(g->mu);
    bfs_fetch(g, dist, parent);
    LGMI_API_END
}

int lgmi_bfs_force_direction(lgmi_graph_t g, int32_t dir) {
    LGMI_API_BEGIN
    if (!g || dir < -1 || dir > 1) throw Error(LGMI_E_ARG, "dir must be -1, 0 or 1");
    std::lock_guard<std::mutex> lk(g->mu);
    bfs_prepare(g);
    g->bfs_forced = dir;
    LGMI_API_END
}

int lgmi_bfs_set_policy(int mode, float alpha, float beta, const float* mlp, int32_t nweights) {
    LGMI_API_BEGIN
    if (mode != LGMI_POLICY_HEURISTIC && mode != LGMI_POLICY_LEARNED && mode != LGMI_POLICY_TOPDOWN)
        throw Error(LGMI_E_ARG, "unknown policy mode");
    if (!(alpha > 0.f) || !(beta > 0.f)) throw Error(LGMI_E_ARG, "alpha and beta must be positive");
    std::lock_guard<std::mutex> lk(policy_mutex());
    PolicyHost& p = policy_host();
    if (mlp) {
        if (nweights != LGMI_POLICY_NWEIGHTS) throw Error(LGMI_E_ARG, "MLP needs LGMI_POLICY_NWEIGHTS floats");
        static_assert(sizeof(PolicyWeights) == sizeof(float) * LGMI_POLICY_NWEIGHTS, "layout");
        PolicyWeights raw;
        std::memcpy(&raw, mlp, sizeof(PolicyWeights));
        for (int i = 0; i < kFeat; i++)
            if (!(raw.sd[i] > 0.f)) throw Error(LGMI_E_ARG, "feature scales must be positive");
        // Fold the standardisation into the first layer for the device BFS:
        // W1'[j][i] = W1[j][i] / sd[i], b1'[j] = b1[j] - sum_i W1'[j][i] mu[i].
        p.w = raw;
        for (int j = 0; j < kHidden; j++) {
            double b = raw.b1[j];
            for (int i = 0; i < kFeat; i++) {
                const double wf = (double)raw.w1[j * kFeat + i] / (double)raw.sd[i];
                p.w.w1[j * kFeat + i] = (float)wf;
                b -= wf * (double)raw.mu[i];
            }
            p.w.b1[j] = (float)b;
        }
        for (int i = 0; i < kFeat; i++) { p.w.mu[i] = 0.f; p.w.sd[i] = 1.f; }
        p.have_mlp = true;
    } else if (mode == LGMI_POLICY_LEARNED && !p.have_mlp) {
        throw Error(LGMI_E_ARG, "learned policy needs MLP weights");
    }
    p.mode = mode;
    p.alpha = alpha;
    p.beta = beta;
    p.version++;
    LGMI_API_END
}

int lgmi_betweenness(lgmi_graph_t g, const int32_t* srcs, int32_t nsrc, int64_t k_scale, int skip_degree0,
                     int normalize, int endpoints, double* bc) {
    LGMI_API_BEGIN
    if (!g || !bc || nsrc < 0 || (nsrc > 0 && !srcs) || k_scale < 0) throw Error(LGMI_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(g->mu);
    for (int32_t i = 0; i < nsrc; i++) check_src(g, srcs[i]);
    bc_run(g, srcs, nsrc, k_scale, skip_degree0, normalize, endpoints, bc, &g->bc_last, 0);
    LGMI_API_END
}

int lgmi_connected_components(lgmi_graph_t g, int32_t* label) {
    LGMI_API_BEGIN
    if (!g || !label) throw Error(LGMI_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(g->mu);
    connected_components(g, label);
    LGMI_API_END
}

int lgmi_has_path(lgmi_graph_t g, int32_t u, int32_t v, const int32_t* exclude, int32_t nexclude, int* result) {
    LGMI_API_BEGIN
    if (!g || !result || nexclude < 0 || (nexclude > 0 && !exclude)) throw Error(LGMI_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(g->mu);
    check_src(g, u);
    check_src(g, v);
    std::vector<int32_t> ex(exclude, exclude + nexclude);
    for (int32_t x : ex) check_src(g, x);
    std::sort(ex.begin(), ex.end());
    ex.erase(std::unique(ex.begin(), ex.end()), ex.end());
    // bfs.jl:153-157: an excluded end is no path; u == v is a path
    if (std::binary_search(ex.begin(), ex.end(), u) || std::binary_search(ex.begin(), ex.end(), v)) {
        *result = 0;
        return LGMI_OK;
    }
    if (u == v) {
        *result = 1;
        return LGMI_OK;
    }
    bfs_run(g, nullptr, 1, u, 0, 0, -1, nullptr, ex.data(), (int32_t)ex.size());
    int32_t dv = -1;
    LGMI_HIP(hipMemcpyAsync(&dv, g->bfs->dist + v, sizeof(int32_t), hipMemcpyDeviceToHost, g->stream));
    LGMI_HIP(hipStreamSynchronize(g->stream));
    *result = dv >= 0 ? 1 : 0;
    LGMI_API_END
}

int lgmi_neighborhood(lgmi_graph_t g, int32_t v, int64_t d, int dir_in, int32_t* verts, int32_t* dists,
                      int32_t* count) {
    LGMI_API_BEGIN
    if (!g || !verts || !count) throw Error(LGMI_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(g->mu);
    check_src(g, v);
    *count = neighborhood(g, v, d, dir_in, verts, dists);
    LGMI_API_END
}

int lgmi_period(lgmi_graph_t g, int64_t* period) {
    LGMI_API_BEGIN
    if (!g || !period) throw Error(LGMI_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(g->mu);
    if (!g->directed) throw Error(LGMI_E_ARG, "period is defined for directed graphs");
    if (g->nv == 0) throw Error(LGMI_E_ARG, "Graph must be strongly connected");
    // is_strongly_connected: vertex 1 reaches every vertex and every vertex reaches it
    lgmi_bfs_stats st{};
    bfs_run(g, nullptr, 1, 0, 1, 0, -1, &st);
    if (st.reached != g->nv) throw Error(LGMI_E_ARG, "Graph must be strongly connected");
    bfs_run(g, nullptr, 1, 0, 0, 0, -1, &st);
    if (st.reached != g->nv) throw Error(LGMI_E_ARG, "Graph must be strongly connected");
    *period = period_from_levels(g, g->bfs->dist);
    LGMI_API_END
}

int lgmi_distance_stats(lgmi_graph_t g, const int32_t* srcs, int32_t nsrc, int64_t* stats) {
    LGMI_API_BEGIN
    if (!g || !stats || nsrc < 0 || (nsrc > 0 && !srcs)) throw Error(LGMI_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(g->mu);
    for (int32_t i = 0; i < nsrc; i++) check_src(g, srcs[i]);
    if (nsrc == 0) return LGMI_OK;
    bc_run(g, srcs, nsrc, 0, 0, 0, 0, nullptr, &g->bc_last, 2, reinterpret_cast<uint64_t*>(stats));
    LGMI_API_END
}

int lgmi_stress(lgmi_graph_t g, const int32_t* srcs, int32_t nsrc, int64_t* stress) {
    LGMI_API_BEGIN
    if (!g || !stress || nsrc < 0 || (nsrc > 0 && !srcs)) throw Error(LGMI_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(g->mu);
    for (int32_t i = 0; i < nsrc; i++) check_src(g, srcs[i]);
    bc_run(g, srcs, nsrc, 0, 1, 0, 0, reinterpret_cast<double*>(stress), &g->bc_last, 1);
    LGMI_API_END
}

int lgmi_betweenness_set_launch(int mode) {
    LGMI_API_BEGIN
    if (mode != 0 && mode != 1) throw Error(LGMI_E_ARG, "launch mode must be 0 or 1");
    bc_set_launch(mode);
    LGMI_API_END
}

int lgmi_set_tuning(const lgmi_tuning* t) {
    LGMI_API_BEGIN
    const lgmi_tuning v = t ? *t : tuning_defaults();
    if (v.wave_row < 2 || v.wave_row > 257 || v.lane_rounds < 0 || v.lane_rounds > 64 || v.range_front < 0 ||
        v.range_front > 2048 || v.queue_front < 0 || v.queue_front > 4096 || v.range_light < 0 ||
        v.queue_min_edges < 0 || v.queue_max_edges < 0 || v.bc_seg_len < 64 || v.bc_seg_len > 65536 ||
        v.solo_front < 0 || v.solo_front > 4096 || v.solo_edges < 0 || v.solo_edges > 262144)
        throw Error(LGMI_E_ARG, "tuning value out of range");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune = v;
    LGMI_API_END
}

int lgmi_get_tuning(lgmi_tuning* t) {
    LGMI_API_BEGIN
    if (!t) throw Error(LGMI_E_ARG, "bad arguments");
    *t = tuning_get();
    LGMI_API_END
}

int lgmi_betweenness_stats(lgmi_graph_t g, lgmi_bc_stats* stats) {
    LGMI_API_BEGIN
    if (!g || !stats) throw Error(LGMI_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(g->mu);
    *stats = g->bc_last;
    LGMI_API_END
}

}  // extern "C"
