// ingest.hip -- edge list -> sorted, de-duplicated CSR on the device, gfx950.
//
// Replaces the edge-list constructors
//   SimpleGraph(edge_list)      src/SimpleGraphs/simplegraph.jl:208-244
//   SimpleDiGraph(edge_list)    src/SimpleGraphs/simpledigraph.jl:212-253
// with cleanupedges! (simplegraph.jl:163-179, simpledigraph.jl:174-189):
// nv is the largest vertex id used, an edge with an id below 1 (below 0
// here) is skipped, an undirected edge enters both rows (a self-loop once),
// rows are sorted and duplicates removed, and ne counts an undirected edge
// once and a self-loop once.
//
// Each edge becomes one 64-bit key (row << b | column, b = bits of nv) per
// row it enters; invalid slots get the all-ones key, which sorts last.  A
// least-significant-digit radix sort over the low 2b bits orders the keys
// (8-bit digits; per digit pass: per-tile digit counts, an exclusive scan
// of the counts in digit-major order, then a stable scatter that ranks each
// key among the equal digits of its wavefront with eight ballots and among
// the earlier wavefronts and chunks of its tile through LDS counters); a
// key equal to its predecessor is a duplicate; an exclusive scan of the
// keep flags gives each kept key its column slot; the row pointers come
// from the row boundaries of the kept keys.  A directed graph sorts a
// second key set (column << b | row) for the in-neighbour CSR.  The host
// receives only the row pointers (the handle's derived structures are built
// from them); columns stay on the device.
#include <cstring>

#include "lgmi_internal.h"

#include <algorithm>
#include <vector>

namespace lgmi {

namespace {

constexpr int kIngBlock = 256;
constexpr uint64_t kNoKey = ~0ull;

__global__ void max_id(const int32_t* src, const int32_t* dst, int64_t m, int32_t* out) {
    int32_t mx = -1;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t s = src[e], d = dst[e];
        if (s >= 0 && d >= 0) mx = max(mx, max(s, d));
    }
    for (int o = 32; o >= 1; o >>= 1) mx = max(mx, __shfl_xor(mx, o));
    if ((threadIdx.x & 63) == 0 && mx >= 0) atomicMax(out, mx);
}

// keys[2e], keys[2e+1] for an undirected edge; keys[e] (reverse=0) or the
// transposed key (reverse=1) for a directed one.
__global__ void make_keys(const int32_t* src, const int32_t* dst, int64_t m, int undirected, int reverse, int b,
                          uint64_t* keys) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t s = src[e], d = dst[e];
        const bool ok = s >= 0 && d >= 0;
        const uint64_t fwd = ((uint64_t)s << b) | (uint64_t)d, bwd = ((uint64_t)d << b) | (uint64_t)s;
        if (undirected) {
            keys[2 * e] = ok ? fwd : kNoKey;
            keys[2 * e + 1] = ok && s != d ? bwd : kNoKey;
        } else {
            keys[e] = ok ? (reverse ? bwd : fwd) : kNoKey;
        }
    }
}

__global__ void keep_flags(const uint64_t* keys, int64_t n, int32_t* flag) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        flag[i] = keys[i] != kNoKey && (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// Kept key i (slot pos[i]) -> col[pos[i]]; the first kept key of a row u
// sets row_ptr for u and every empty row between the previous kept row and
// u; self-loops are counted.
__global__ void scatter_csr(const uint64_t* keys, const int32_t* flag, const int32_t* pos, int64_t n, int b,
                            int32_t nv, int32_t* col, int32_t* row_ptr, unsigned long long* loops) {
    const uint64_t mask = (1ull << b) - 1;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (!flag[i]) continue;
        const uint64_t k = keys[i];
        const int32_t u = (int32_t)(k >> b), v = (int32_t)(k & mask);
        const int32_t p = pos[i];
        col[p] = v;
        if (u == v) atomicAdd(loops, 1ull);
        // the previous kept key (flags are 0 for duplicates and the tail)
        int32_t uprev = -1;
        if (p > 0) {
            int64_t j = i - 1;
            while (!flag[j]) j--;
            uprev = (int32_t)(keys[j] >> b);
        }
        if (uprev != u)
            for (int32_t w = uprev + 1; w <= u; w++) row_ptr[w] = p;
    }
}

// ---- radix sort of 64-bit keys, 8 bits per pass ----
constexpr int kRsBlock = 256;                 // threads per work-group
constexpr int kRsPer = 16;                    // keys per thread per tile
constexpr int kRsTile = kRsBlock * kRsPer;    // keys per tile
constexpr int kRsBins = 256;                  // digit values
constexpr int kRsWaves = kRsBlock / 64;
static_assert(kRsBins == kRsBlock, "one digit per thread in the count and carry steps");

// counts[d * tiles + t] = keys of tile t with digit d at `shift`.
__global__ __launch_bounds__(kRsBlock) void rs_count(const uint64_t* keys, int64_t n, int shift, int32_t tiles,
                                                     int32_t* counts) {
    __shared__ int32_t s_h[kRsBins];
    const int tid = threadIdx.x;
    s_h[tid] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRsTile;
    #pragma unroll 4
    for (int j = 0; j < kRsPer; j++) {
        const int64_t i = base + (int64_t)j * kRsBlock + tid;
        if (i < n) atomicAdd(&s_h[(int)((keys[i] >> shift) & (kRsBins - 1))], 1);
    }
    __syncthreads();
    counts[(int64_t)tid * tiles + blockIdx.x] = s_h[tid];
}

// Stable scatter of tile t's keys to their sorted positions: the tile's
// keys are taken in index order, 256 at a time (one per thread); a key's
// position is the scanned count of its digit for this tile, plus the keys
// of that digit in earlier chunks of the tile, in earlier wavefronts of the
// chunk, and in lower lanes of its wavefront (lanes with an equal digit are
// found with one ballot per digit bit).
__global__ __launch_bounds__(kRsBlock) void rs_scatter(const uint64_t* in, uint64_t* out, int64_t n, int shift,
                                                       int32_t tiles, const int32_t* offs) {
    __shared__ int32_t s_base[kRsBins];
    __shared__ int32_t s_wcnt[kRsWaves][kRsBins];
    const int tid = threadIdx.x, lane = tid & 63, wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    s_base[tid] = offs[(int64_t)tid * tiles + blockIdx.x];
    const int64_t base = (int64_t)blockIdx.x * kRsTile;
    const uint64_t below = (1ull << lane) - 1ull;
    for (int j = 0; j < kRsPer; j++) {
        #pragma unroll
        for (int w = 0; w < kRsWaves; w++) s_wcnt[w][tid] = 0;
        __syncthreads();
        const int64_t i = base + (int64_t)j * kRsBlock + tid;
        const bool valid = i < n;
        const uint64_t k = valid ? in[i] : 0ull;
        const int d = (int)((k >> shift) & (kRsBins - 1));
        uint64_t same = __ballot(valid);
        #pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            const uint64_t m = __ballot((d >> bit) & 1);
            same &= ((d >> bit) & 1) ? m : ~m;
        }
        const int rank = __popcll(same & below);
        if (valid && rank == 0) s_wcnt[wid][d] = __popcll(same);
        __syncthreads();
        if (valid) {
            int32_t pre = 0;
            for (int w = 0; w < wid; w++) pre += s_wcnt[w][d];
            out[s_base[d] + pre + rank] = k;
        }
        __syncthreads();
        int32_t add = 0;
        #pragma unroll
        for (int w = 0; w < kRsWaves; w++) add += s_wcnt[w][tid];
        s_base[tid] += add;
    }
}

// Sorts n keys by their low end_bit bits; the result is in *keys or *alt
// (returned); both buffers are overwritten.
static uint64_t* radix_sort(hipStream_t s, uint64_t* keys, uint64_t* alt, int64_t n, int end_bit) {
    if (n <= 1) return keys;
    const int64_t tiles64 = (n + kRsTile - 1) / kRsTile;
    if (tiles64 * kRsBins > 0x7fffffffLL) throw Error(LGMI_E_TOO_LARGE, "too many keys for the radix sort");
    const int32_t tiles = (int32_t)tiles64;
    int32_t* counts = dev_alloc_n<int32_t>((size_t)tiles * kRsBins);
    int32_t* offs = nullptr;
    try {
        offs = dev_alloc_n<int32_t>((size_t)tiles * kRsBins);
        uint64_t *src = keys, *dst = alt;
        for (int shift = 0; shift < end_bit; shift += 8) {
            rs_count<<<tiles, kRsBlock, 0, s>>>(src, n, shift, tiles, counts);
            LGMI_HIP(hipGetLastError());
            exclusive_scan(s, counts, offs, tiles * kRsBins);
            rs_scatter<<<tiles, kRsBlock, 0, s>>>(src, dst, n, shift, tiles, offs);
            LGMI_HIP(hipGetLastError());
            std::swap(src, dst);
        }
        LGMI_HIP(hipStreamSynchronize(s));
        dev_free(counts);
        dev_free(offs);
        return src;
    } catch (...) {
        dev_free(counts);
        dev_free(offs);
        throw;
    }
}

__global__ void tail_rows(int32_t* row_ptr, int32_t from, int32_t nv, int32_t total) {
    for (int32_t w = from + blockIdx.x * blockDim.x + threadIdx.x; w <= nv; w += gridDim.x * blockDim.x)
        row_ptr[w] = total;
}

__global__ void last_row(const uint64_t* keys, const int32_t* flag, int64_t n, int b, int32_t* out) {
    if (blockIdx.x || threadIdx.x) return;
    int64_t j = n - 1;
    while (j >= 0 && !flag[j]) j--;
    *out = j < 0 ? -1 : (int32_t)(keys[j] >> b);
}

}  // namespace

// One CSR from `n` keys of `b`-bit halves: sorts, de-duplicates and fills
// d.row_ptr / d.col (allocated here).  Returns the entry count; *loops gets
// the self-loops.
static int64_t keys_to_csr(lgmi_graph* g, uint64_t* keys, int64_t n, int b, DevCSR& d, int64_t* loops,
                           std::vector<int32_t>& rp_host) {
    hipStream_t s = g->stream;
    const int32_t nv = g->nv;
    uint64_t* alt = dev_alloc_n<uint64_t>((size_t)std::max<int64_t>(n, 1));
    uint64_t* sorted = nullptr;
    int32_t *flag = nullptr, *pos = nullptr, *lastu = nullptr;
    unsigned long long* dloops = nullptr;
    try {
        sorted = radix_sort(s, keys, alt, n, std::min(64, 2 * b));
        flag = dev_alloc_n<int32_t>((size_t)std::max<int64_t>(n, 1));
        pos = dev_alloc_n<int32_t>((size_t)std::max<int64_t>(n, 1));
        const int grid = std::max(1, (int)std::min<int64_t>(g->cu_count * 8, (n + kIngBlock - 1) / kIngBlock));
        keep_flags<<<grid, kIngBlock, 0, s>>>(sorted, n, flag);
        LGMI_HIP(hipGetLastError());
        const int64_t total = n > 0 ? exclusive_scan(s, flag, pos, (int32_t)n) : 0;
        if (total > 0x7fffffffLL) throw Error(LGMI_E_TOO_LARGE, "more than 2^31-1 adjacency entries");
        d.row_ptr = dev_alloc_n<int32_t>((size_t)nv + 1);
        d.col = dev_alloc_n<int32_t>((size_t)std::max<int64_t>(total, 1));
        dloops = dev_alloc_n<unsigned long long>(1);
        lastu = dev_alloc_n<int32_t>(1);
        LGMI_HIP(hipMemsetAsync(dloops, 0, sizeof(unsigned long long), s));
        LGMI_HIP(hipMemsetAsync(d.row_ptr, 0, sizeof(int32_t) * ((size_t)nv + 1), s));
        int32_t h_last = -1;
        if (n > 0) {
            scatter_csr<<<grid, kIngBlock, 0, s>>>(sorted, flag, pos, n, b, nv, d.col, d.row_ptr, dloops);
            last_row<<<1, 64, 0, s>>>(sorted, flag, n, b, lastu);
            LGMI_HIP(hipGetLastError());
            LGMI_HIP(hipMemcpyAsync(&h_last, lastu, sizeof(int32_t), hipMemcpyDeviceToHost, s));
            LGMI_HIP(hipStreamSynchronize(s));
        }
        // rows after the last non-empty one
        tail_rows<<<std::max(1, std::min(g->cu_count * 4, (nv + 1 - (h_last + 1) + 255) / 256)), 256, 0, s>>>(
            d.row_ptr, h_last + 1, nv, (int32_t)total);
        LGMI_HIP(hipGetLastError());
        unsigned long long h_loops = 0;
        rp_host.resize((size_t)nv + 1);
        LGMI_HIP(hipMemcpyAsync(&h_loops, dloops, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        LGMI_HIP(hipMemcpyAsync(rp_host.data(), d.row_ptr, sizeof(int32_t) * ((size_t)nv + 1), hipMemcpyDeviceToHost, s));
        LGMI_HIP(hipStreamSynchronize(s));
        *loops = (int64_t)h_loops;
        dev_free(alt); dev_free(flag); dev_free(pos); dev_free(dloops); dev_free(lastu);
        return total;
    } catch (...) {
        dev_free(alt); dev_free(flag); dev_free(pos); dev_free(dloops); dev_free(lastu);
        throw;
    }
}

lgmi_graph* graph_from_edges(const int32_t* h_src, const int32_t* h_dst, int64_t m, int directed, int32_t nv_min,
                             int64_t* ne_out) {
    lgmi_graph* g = new lgmi_graph();
    int32_t *src = nullptr, *dst = nullptr, *dmax = nullptr;
    uint64_t* keys = nullptr;
    try {
        g->directed = directed ? 1 : 0;
        LGMI_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
        g->cu_count = device_cu_count();
        hipStream_t s = g->stream;
        src = dev_alloc_n<int32_t>((size_t)std::max<int64_t>(m, 1));
        dst = dev_alloc_n<int32_t>((size_t)std::max<int64_t>(m, 1));
        dmax = dev_alloc_n<int32_t>(1);
        if (m > 0) {
            LGMI_HIP(hipMemcpyAsync(src, h_src, sizeof(int32_t) * (size_t)m, hipMemcpyHostToDevice, s));
            LGMI_HIP(hipMemcpyAsync(dst, h_dst, sizeof(int32_t) * (size_t)m, hipMemcpyHostToDevice, s));
        }
        const int32_t minus1 = -1;
        LGMI_HIP(hipMemcpyAsync(dmax, &minus1, sizeof(int32_t), hipMemcpyHostToDevice, s));
        const int grid = std::max(1, (int)std::min<int64_t>(g->cu_count * 8, (m + kIngBlock - 1) / kIngBlock));
        if (m > 0) max_id<<<grid, kIngBlock, 0, s>>>(src, dst, m, dmax);
        LGMI_HIP(hipGetLastError());
        int32_t mx = -1;
        LGMI_HIP(hipMemcpyAsync(&mx, dmax, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        LGMI_HIP(hipStreamSynchronize(s));
        const int32_t nv = std::max(nv_min, mx + 1);
        g->nv = nv;
        // 2^b > nv: every valid key is below 2^(2b) - 1, which the all-ones
        // key reads as in the sorted bits, so it sorts after them
        int b = 1;
        while ((1LL << b) <= (int64_t)nv) b++;
        const int64_t n = directed ? m : 2 * m;
        if (n > 0x7fffffffLL) throw Error(LGMI_E_TOO_LARGE, "more than 2^31-1 adjacency entries");
        keys = dev_alloc_n<uint64_t>((size_t)std::max<int64_t>(n, 1));
        std::vector<int32_t> rp;
        int64_t loops = 0;
        if (n > 0) make_keys<<<grid, kIngBlock, 0, s>>>(src, dst, m, directed ? 0 : 1, 0, b, keys);
        LGMI_HIP(hipGetLastError());
        const int64_t nnz = keys_to_csr(g, keys, n, b, g->out, &loops, rp);
        g->nnz = nnz;
        finish_csr(g, g->out, rp.data());
        if (directed) {
            if (n > 0) make_keys<<<grid, kIngBlock, 0, s>>>(src, dst, m, 0, 1, b, keys);
            LGMI_HIP(hipGetLastError());
            std::vector<int32_t> irp;
            int64_t iloops = 0;
            keys_to_csr(g, keys, n, b, g->in, &iloops, irp);
            finish_csr(g, g->in, irp.data());
            *ne_out = nnz;
        } else {
            g->in = g->out;
            *ne_out = (nnz + loops) / 2;
        }
        LGMI_HIP(hipStreamSynchronize(s));
        dev_free(src); dev_free(dst); dev_free(dmax); dev_free(keys);
    } catch (...) {
        dev_free(src); dev_free(dst); dev_free(dmax); dev_free(keys);
        graph_release(g);
        throw;
    }
    return g;
}

}  // namespace lgmi
