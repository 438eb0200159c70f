// gen.cpp -- workload generators and the sorted Int32 CSR builder.
//
// Host-only C++ compiled into liblgmi.so.  The generators restate the
// reference's algorithms (LightGraphs.jl v1.3.5) with this project's own
// counter-based RNG (SplitMix64), so graph instances are reproducible from
// a seed on any machine but are not the instances Julia's MersenneTwister
// would draw:
//   kronecker          src/SimpleGraphs/generators/randgraphs.jl:1126-1152
//   erdos_renyi(n, ne) src/SimpleGraphs/generators/randgraphs.jl:25-40
//   barabasi_albert    src/SimpleGraphs/generators/randgraphs.jl:391-512
//   watts_strogatz     src/SimpleGraphs/generators/randgraphs.jl:262-313
//   grid([r, c])       src/SimpleGraphs/generators/staticgraphs.jl:452-474
//                      with cartesian_product, src/operators.jl:534-551
// The CSR builder gives the storage invariant of SimpleGraph/SimpleDiGraph
// (sorted, de-duplicated adjacency; src/SimpleGraphs/simplegraph.jl:445-461,
// simpledigraph.jl:414-430), 0-based.
#include "../../include/lgmi_gen.h"

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <unordered_set>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Stateless draw number `i` of stream `stream` under `seed`.
inline uint64_t draw(uint64_t seed, uint64_t stream, uint64_t i) {
    return splitmix64(splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + i);
}

inline double to_unit(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

// Uniform integer in [0, n) by 64x64->128 multiply-shift.
inline uint64_t to_range(uint64_t x, uint64_t n) {
    return (uint64_t)(((__uint128_t)x * (__uint128_t)n) >> 64);
}

// Sequential stream.
struct Rng {
    uint64_t seed, stream, i = 0;
    Rng(uint64_t s, uint64_t st) : seed(s), stream(st) {}
    uint64_t next() { return draw(seed, stream, i++); }
    double unit() { return to_unit(next()); }
    uint64_t range(uint64_t n) { return to_range(next(), n); }
};

int hw_threads() {
    unsigned n = std::thread::hardware_concurrency();
    if (n == 0) n = 1;
    if (n > 16) n = 16;
    return (int)n;
}

template <class F>
void parallel_for(int64_t n, F f) {
    int nt = hw_threads();
    if (n < 65536 || nt == 1) { f(0, n); return; }
    std::vector<std::thread> th;
    int64_t chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; t++) {
        int64_t lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) break;
        th.emplace_back([=] { f(lo, hi); });
    }
    for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

// Graph500 Kronecker sampler (randgraphs.jl:1126-1152).  For each of the
// M = edgefactor * 2^scale samples and each bit level: the row bit is set
// with probability 1-(A+B); the column bit is set with probability
// 1-C/(1-(A+B)) when the row bit is set and 1-A/(A+B) otherwise
// (:1134-1136).  Vertex ids are relabelled by a random permutation (:1140-
// 1141).  The edge shuffle (:1143-1144) does not change the graph and is
// omitted.  Writes M endpoint pairs, 0-based; returns M.
int64_t lgmi_gen_kronecker(int32_t scale, int32_t edgefactor, double A, double B, double C,
                           uint64_t seed, int32_t* src, int32_t* dst) {
    if (scale < 1 || scale > 30 || edgefactor < 1) return -1;
    const int64_t N = int64_t(1) << scale;
    const int64_t M = (int64_t)edgefactor * N;
    const double ab = A + B;
    const double c_norm = C / (1.0 - (A + B));
    const double a_norm = A / (A + B);
    parallel_for(M, [&](int64_t lo, int64_t hi) {
        for (int64_t e = lo; e < hi; e++) {
            int64_t i = 0, j = 0;
            for (int ib = 0; ib < scale; ib++) {
                double r1 = to_unit(draw(seed, 1, (uint64_t)e * 64 + 2 * ib));
                double r2 = to_unit(draw(seed, 1, (uint64_t)e * 64 + 2 * ib + 1));
                int ii = r1 > ab;
                int jj = r2 > (ii ? c_norm : a_norm);
                i |= (int64_t)ii << ib;
                j |= (int64_t)jj << ib;
            }
            src[e] = (int32_t)i;
            dst[e] = (int32_t)j;
        }
    });
    // randperm(rng, N): Fisher-Yates on the seeded stream.
    std::vector<int32_t> p((size_t)N);
    for (int64_t k = 0; k < N; k++) p[k] = (int32_t)k;
    Rng rng(seed, 2);
    for (int64_t k = N - 1; k > 0; k--) {
        int64_t r = (int64_t)rng.range((uint64_t)k + 1);
        std::swap(p[k], p[r]);
    }
    parallel_for(M, [&](int64_t lo, int64_t hi) {
        for (int64_t e = lo; e < hi; e++) { src[e] = p[src[e]]; dst[e] = p[dst[e]]; }
    });
    return M;
}

// SimpleGraph{T}(nv, ne; seed) (randgraphs.jl:25-40): draw uniform vertex
// pairs and keep non-loop pairs not yet present until ne edges exist.  For
// ne > 2/3 of the maximum the reference builds the complement of a graph
// with maxe-ne edges; that branch is restated too.  Writes ne pairs (u < v).
int64_t lgmi_gen_erdos_renyi(int32_t n, int64_t ne, uint64_t seed, int32_t* src, int32_t* dst) {
    if (n < 0 || ne < 0) return -1;
    const int64_t maxe = (int64_t)n * (n - 1) / 2;
    if (ne > maxe) return -1;
    const bool complement = ne > (2 * maxe) / 3;
    const int64_t target = complement ? maxe - ne : ne;
    std::unordered_set<uint64_t> have;
    have.reserve((size_t)target * 2 + 16);
    Rng rng(seed, 3);
    while ((int64_t)have.size() < target) {
        uint64_t s = rng.range((uint64_t)n), d = rng.range((uint64_t)n);
        if (s == d) continue;
        uint64_t u = std::min(s, d), v = std::max(s, d);
        have.insert((u << 32) | v);
    }
    int64_t k = 0;
    if (!complement) {
        std::vector<uint64_t> keys(have.begin(), have.end());
        std::sort(keys.begin(), keys.end());
        for (uint64_t key : keys) { src[k] = (int32_t)(key >> 32); dst[k] = (int32_t)(key & 0xffffffffu); k++; }
    } else {
        for (int64_t u = 0; u < n; u++)
            for (int64_t v = u + 1; v < n; v++)
                if (!have.count(((uint64_t)u << 32) | (uint64_t)v)) { src[k] = (int32_t)u; dst[k] = (int32_t)v; k++; }
    }
    return k;
}

// barabasi_albert(n, k) = barabasi_albert(n, k, k) (randgraphs.jl:391-392):
// start from k isolated vertices; since the start graph has no edges, vertex
// k (0-based) is attached to all k existing vertices (sample!(1:n0-1, k) of
// exactly k items, :472-480).  Every later vertex draws distinct targets
// from the degree-weighted endpoint list (:494-510).  Returns the edge count
// k + (n - k - 1) * k.
int64_t lgmi_gen_barabasi_albert(int32_t n, int32_t k, uint64_t seed, int32_t* src, int32_t* dst) {
    if (k < 1 || n < k) return -1;
    if (n == k) return 0;
    Rng rng(seed, 4);
    int64_t ne = 0;
    int32_t n0 = k + 1;                          // n0 += 1 (0-based new vertex id k)
    std::vector<int32_t> weighted;
    weighted.reserve((size_t)2 * ((size_t)(n - n0) * k + k));
    // sample!(rng, collect(1:n0-1), k) with k == n0-1 draws all of them; the
    // order does not change the graph.
    for (int32_t t = 0; t < k; t++) { src[ne] = k; dst[ne] = t; ne++; }
    for (int64_t e = 0; e < ne; e++) {           // edges(g) order: (min, max)
        weighted.push_back(std::min(src[e], dst[e]));
        weighted.push_back(std::max(src[e], dst[e]));
    }
    std::vector<uint8_t> picked((size_t)n, 0);
    std::vector<int32_t> targets((size_t)k);
    for (int32_t source = n0; source < n; source++) {
        int i = 0;
        while (i < k) {
            int32_t target = weighted[rng.range(weighted.size())];
            if (!picked[target]) { targets[i++] = target; picked[target] = 1; }
        }
        for (int32_t t : targets) {
            src[ne] = source; dst[ne] = t; ne++;
            weighted.push_back(source);
            weighted.push_back(t);
            picked[t] = 0;
        }
    }
    return ne;
}

// watts_strogatz(n, k, beta) (randgraphs.jl:262-313), undirected.  Phase 1
// links every s to its i-th clockwise neighbour for i = 1..k/2; phase 2
// visits (i, s) in the same order and, with probability beta and while s is
// not connected to everyone, redraws a target until it either hits the
// original target (keep) or a new non-neighbour (rewire).  The reference
// draws the candidate from the global RNG (:302); here both draws use the
// seeded stream so that the graph is reproducible.  Returns the edge count
// (n*k/2 unless k >= n-1).
int64_t lgmi_gen_watts_strogatz(int32_t n, int32_t k, double beta, uint64_t seed,
                                int32_t* src, int32_t* dst) {
    if (k >= n || k < 0) return -1;
    std::vector<std::vector<int32_t>> adj((size_t)n);
    auto has = [&](int32_t a, int32_t b) {
        const auto& v = adj[a];
        return std::binary_search(v.begin(), v.end(), b);
    };
    auto add = [&](int32_t a, int32_t b) {
        if (a == b || has(a, b)) return false;
        auto& va = adj[a]; va.insert(std::lower_bound(va.begin(), va.end(), b), b);
        auto& vb = adj[b]; vb.insert(std::lower_bound(vb.begin(), vb.end(), a), a);
        return true;
    };
    auto rem = [&](int32_t a, int32_t b) {
        auto& va = adj[a]; auto ia = std::lower_bound(va.begin(), va.end(), b);
        if (ia == va.end() || *ia != b) return false;
        va.erase(ia);
        auto& vb = adj[b]; vb.erase(std::lower_bound(vb.begin(), vb.end(), a));
        return true;
    };
    if (k == n - 1 && k % 2 == 0) {              // complete graph (:268-270)
        int64_t ne = 0;
        for (int32_t a = 0; a < n; a++) for (int32_t b = a + 1; b < n; b++) { src[ne] = a; dst[ne] = b; ne++; }
        return ne;
    }
    for (auto& v : adj) v.reserve((size_t)k + 4);
    auto target = [&](int32_t s, int32_t i) { return (int32_t)(((int64_t)s + i) % n); };
    for (int32_t i = 1; i <= k / 2; i++)
        for (int32_t s = 0; s < n; s++) add(s, target(s, i));
    Rng rng(seed, 5);
    for (int32_t i = 1; i <= k / 2; i++) {
        for (int32_t s = 0; s < n; s++) {
            if (!(rng.unit() < beta && (int64_t)adj[s].size() < n - 1)) continue;
            int32_t t = target(s, i);
            while (true) {
                int32_t d = (int32_t)rng.range((uint64_t)n);
                if (d == s) continue;
                if (d == t) break;
                if (add(s, d)) { rem(s, t); break; }
            }
        }
    }
    int64_t ne = 0;
    for (int32_t a = 0; a < n; a++)
        for (int32_t b : adj[a]) if (a < b) { src[ne] = a; dst[ne] = b; ne++; }
    return ne;
}

// grid([r, c]) = cartesian_product(path_graph(c), path_graph(r)) with id
// (i-1)*r + j (operators.jl:534-551, staticgraphs.jl:452-474): vertex
// (i, j), i in 0..c-1, j in 0..r-1 gets id i*r + j.  Returns 2rc - r - c.
int64_t lgmi_gen_grid2d(int32_t r, int32_t c, int32_t* src, int32_t* dst) {
    if (r <= 0 || c <= 0) return 0;
    int64_t ne = 0;
    for (int64_t i = 0; i + 1 < c; i++)          // edges of the first factor
        for (int64_t j = 0; j < r; j++) { src[ne] = (int32_t)(i * r + j); dst[ne] = (int32_t)((i + 1) * r + j); ne++; }
    for (int64_t j = 0; j + 1 < r; j++)          // edges of the second factor
        for (int64_t i = 0; i < c; i++) { src[ne] = (int32_t)(i * r + j); dst[ne] = (int32_t)(i * r + j + 1); ne++; }
    return ne;
}

// Sorted, de-duplicated CSR from an edge list.  symmetrize=1 stores every
// edge in both rows (SimpleGraph fadjlist); drop_self_loops=1 removes u==v.
// row_ptr has nv+1 entries; col must hold col_cap entries (2*ne suffices
// when symmetrising, ne otherwise).  Returns 0, or a negative code when the
// result needs more than col_cap entries or 2^31-1 entries.
int lgmi_csr_build(int32_t nv, int64_t ne, const int32_t* src, const int32_t* dst, int symmetrize,
                   int drop_self_loops, int32_t* row_ptr, int32_t* col, int64_t col_cap,
                   int64_t* nnz_out) {
    if (nv < 0 || ne < 0) return -1;
    for (int64_t e = 0; e < ne; e++)
        if (src[e] < 0 || src[e] >= nv || dst[e] < 0 || dst[e] >= nv) return -2;
    std::vector<int64_t> cnt((size_t)nv + 1, 0);
    for (int64_t e = 0; e < ne; e++) {
        int32_t u = src[e], v = dst[e];
        if (u == v && drop_self_loops) continue;
        cnt[u + 1]++;
        if (symmetrize && u != v) cnt[v + 1]++;
    }
    for (int32_t v = 0; v < nv; v++) cnt[v + 1] += cnt[v];
    const int64_t total = cnt[nv];
    std::vector<int32_t> tmp((size_t)std::max<int64_t>(total, 1));
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t e = 0; e < ne; e++) {
        int32_t u = src[e], v = dst[e];
        if (u == v && drop_self_loops) continue;
        tmp[pos[u]++] = v;
        if (symmetrize && u != v) tmp[pos[v]++] = u;
    }
    std::vector<int64_t> uniq((size_t)nv, 0);
    parallel_for(nv, [&](int64_t lo, int64_t hi) {
        for (int64_t v = lo; v < hi; v++) {
            auto b = tmp.begin() + cnt[v], e = tmp.begin() + cnt[v + 1];
            std::sort(b, e);
            uniq[v] = std::unique(b, e) - b;
        }
    });
    int64_t nnz = 0;
    for (int32_t v = 0; v < nv; v++) nnz += uniq[v];
    if (nnz > col_cap) return -3;
    if (nnz > 0x7fffffffLL) return -4;
    row_ptr[0] = 0;
    for (int32_t v = 0; v < nv; v++) row_ptr[v + 1] = (int32_t)(row_ptr[v] + uniq[v]);
    parallel_for(nv, [&](int64_t lo, int64_t hi) {
        for (int64_t v = lo; v < hi; v++)
            std::memcpy(col + row_ptr[v], tmp.data() + cnt[v], sizeof(int32_t) * (size_t)uniq[v]);
    });
    *nnz_out = nnz;
    return 0;
}

// Transpose (in-neighbour CSR, the badjlist of a SimpleDiGraph); rows stay
// sorted because rows of the input are visited in ascending order.
int lgmi_csr_transpose(int32_t nv, const int32_t* row_ptr, const int32_t* col, int32_t* t_row_ptr,
                       int32_t* t_col) {
    const int64_t nnz = row_ptr[nv];
    std::vector<int64_t> cnt((size_t)nv + 1, 0);
    for (int64_t e = 0; e < nnz; e++) {
        if (col[e] < 0 || col[e] >= nv) return -2;
        cnt[col[e] + 1]++;
    }
    for (int32_t v = 0; v < nv; v++) cnt[v + 1] += cnt[v];
    for (int32_t v = 0; v <= nv; v++) t_row_ptr[v] = (int32_t)cnt[v];
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int32_t u = 0; u < nv; u++)
        for (int32_t e = row_ptr[u]; e < row_ptr[u + 1]; e++) t_col[pos[col[e]]++] = u;
    return 0;
}

}  // extern "C"
