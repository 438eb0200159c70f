/*
 * lgmi_gen.h -- workload generators and CSR construction exported by
 * liblgmi.so (host code; no GPU involved).
 *
 * These are not on the accelerated path.  They build the graphs the
 * benchmarks and tests run on, and the sorted Int32 CSR that the Julia shim
 * produces from a fadjlist.  Each generator restates a reference generator
 * with this project's counter-based SplitMix64 RNG; all vertex ids are
 * 0-based.
 */
#ifndef LGMI_GEN_H
#define LGMI_GEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* kronecker(SCALE, edgefactor, A, B, C; seed)
 * src/SimpleGraphs/generators/randgraphs.jl:1126-1152.
 * Writes edgefactor * 2^scale pairs into src/dst; returns that count. */
int64_t lgmi_gen_kronecker(int32_t scale, int32_t edgefactor, double A, double B, double C,
                           uint64_t seed, int32_t* src, int32_t* dst);

/* SimpleGraph(nv, ne; seed), also erdos_renyi(nv, ne)
 * src/SimpleGraphs/generators/randgraphs.jl:25-40, 149-151.
 * Writes ne pairs (u < v); returns ne. */
int64_t lgmi_gen_erdos_renyi(int32_t n, int64_t ne, uint64_t seed, int32_t* src, int32_t* dst);

/* barabasi_albert(n, k; seed)
 * src/SimpleGraphs/generators/randgraphs.jl:391-392, 417-426, 449-512.
 * Needs room for k + (n-k-1)*k pairs; returns the pair count. */
int64_t lgmi_gen_barabasi_albert(int32_t n, int32_t k, uint64_t seed, int32_t* src, int32_t* dst);

/* watts_strogatz(n, k, beta; seed), undirected
 * src/SimpleGraphs/generators/randgraphs.jl:262-313.
 * Needs room for n*k/2 pairs (n*(n-1)/2 when k == n-1); returns the count. */
int64_t lgmi_gen_watts_strogatz(int32_t n, int32_t k, double beta, uint64_t seed,
                                int32_t* src, int32_t* dst);

/* grid([r, c]) -- src/SimpleGraphs/generators/staticgraphs.jl:452-474.
 * Needs room for 2rc - r - c pairs; returns the count. */
int64_t lgmi_gen_grid2d(int32_t r, int32_t c, int32_t* src, int32_t* dst);

/* Sorted, de-duplicated CSR (the fadjlist invariant,
 * src/SimpleGraphs/simplegraph.jl:445-461).  Returns 0 or a negative code:
 * -1 bad size, -2 endpoint out of range, -3 col_cap too small, -4 more than
 * 2^31-1 entries. */
int lgmi_csr_build(int32_t nv, int64_t ne, const int32_t* src, const int32_t* dst, int symmetrize,
                   int drop_self_loops, int32_t* row_ptr, int32_t* col, int64_t col_cap,
                   int64_t* nnz_out);

/* In-neighbour CSR (the badjlist of a SimpleDiGraph,
 * src/SimpleGraphs/simpledigraph.jl:8-22); rows sorted ascending. */
int lgmi_csr_transpose(int32_t nv, const int32_t* row_ptr, const int32_t* col, int32_t* t_row_ptr,
                       int32_t* t_col);

/* The edge lines of one graph in the LG text format, read as
 * _lg_read_one_graph reads them (src/persistence/lg.jl:35-52).  At most
 * nlines lines of text[0..len) are read: an empty line (after its line break
 * is removed) is skipped but counts; any other line gives one pair, its first
 * two comma-separated fields parsed as Int64 (blanks around commas and
 * numbers allowed, optional sign, 0x/0o/0b prefixes; further fields are
 * ignored).  The pairs, as written (1-based, not range-checked), go to
 * src/dst, which hold nlines entries; *npairs = their count, *consumed = the
 * bytes of text the lines took (where the next header starts).  Returns 0,
 * or -1 for a line with a single field, -2 for a field that is not an
 * integer, -3 for one beyond Int64 (*bad_line = the 1-based line, counted
 * from the first edge line), -4 for bad arguments. */
int lgmi_lg_parse_edges(const char* text, int64_t len, int64_t nlines, int64_t* src, int64_t* dst,
                        int64_t* npairs, int64_t* consumed, int64_t* bad_line);

#ifdef __cplusplus
}
#endif

#endif /* LGMI_GEN_H */
