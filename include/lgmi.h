/*
 * lgmi.h -- C ABI of liblgmi.so: MI355X (gfx950) breadth-first search and
 * exact Brandes betweenness for LightGraphs.jl's SimpleGraph/SimpleDiGraph.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns
 * LGMI_OK (0) or a negative LGMI_E_* code, and leaves a message for
 * lgmi_last_error() on failure.  No C++ exception crosses the boundary.
 *
 * Vertex ids are 0-based Int32; a graph is a CSR whose rows are sorted
 * ascending (the fadjlist/badjlist invariant of the reference,
 * src/SimpleGraphs/simplegraph.jl:445-461, simpledigraph.jl:414-430).
 * Unreached distances and parents come back as -1; the Julia shim maps them
 * to typemax(T) and 0 and adds 1 to every id.
 *
 * Host arrays are borrowed for the duration of a call.  Calls on one graph
 * handle are serialised by the handle; distinct handles may be used from
 * distinct threads.  Results are complete when a call returns.
 */
#ifndef LGMI_H
#define LGMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LGMI_OK            0
#define LGMI_E_ARG        -1  /* ArgumentError in the shim                 */
#define LGMI_E_BOUNDS     -2  /* BoundsError: a vertex id out of range     */
#define LGMI_E_TOO_LARGE  -3  /* more than 2^31-1 adjacency entries        */
#define LGMI_E_NODEV      -4  /* no GPU visible                            */
#define LGMI_E_ALLOC      -5  /* device allocation failed                  */
#define LGMI_E_HIP        -6  /* any other HIP runtime error               */

#define LGMI_POLICY_HEURISTIC 0   /* Beamer alpha/beta direction switch    */
#define LGMI_POLICY_LEARNED   1   /* FP32 MLP 16 -> 64 (ReLU) -> 2 logits  */
#define LGMI_POLICY_TOPDOWN   2   /* always top-down (reference behaviour) */
#define LGMI_POLICY_FEATURES  16
#define LGMI_POLICY_HIDDEN    64
/* MLP parameter vector layout (FP32): W1[64x16] row-major, b1[64],
 * W2[2x64] row-major, b2[2], feature mean[16], feature scale[16]. */
#define LGMI_POLICY_NWEIGHTS  (64 * 16 + 64 + 2 * 64 + 2 + 16 + 16)

typedef struct lgmi_graph* lgmi_graph_t;

/* Statistics of the last BFS run on a handle. */
typedef struct {
    int32_t levels;            /* number of BFS steps executed               */
    int32_t reached;           /* vertices with a finite distance            */
    int32_t td_steps;          /* steps run top-down                         */
    int32_t bu_steps;          /* steps run bottom-up                        */
    double device_ms;          /* HIP-event time, init to last step          */
    int64_t edges_reached;     /* sum over reached vertices of out-degree    */
    int32_t launches;          /* step launches (a few after the last level  */
                               /* exit at once)                              */
    int32_t pad;
    double steps_event_ms;     /* sum of per-launch HIP-event times of the   */
                               /* step kernel (0 unless step timing is on)   */
} lgmi_bfs_stats;

/* One BFS step (level L -> L+1) of the last run. */
typedef struct {
    float feat[16];            /* policy features of the state at step start */
    int32_t dir;               /* 0 top-down, 1 bottom-up                    */
    int32_t n_front;           /* |F_L|                                      */
    int64_t m_front;           /* traversal-direction edges of F_L           */
    int64_t m_unvisited;       /* scan-direction edges of unvisited vertices */
    int32_t n_next;            /* |F_{L+1}|                                  */
    int32_t variant;           /* top-down: 1 range mode (each work-group    */
                               /* owns a vertex range), 2 queue mode (the    */
                               /* frontier list is shared out by edges),     */
                               /* 3 solo mode (one work-group, several       */
                               /* levels per launch); 0 otherwise            */
    int64_t scanned;           /* bottom-up: adjacency entries read;         */
                               /* top-down: edges expanded                   */
    int64_t cand;              /* bottom-up: unvisited vertices examined     */
    float kernel_us;           /* first block start to last block end        */
    float step_us;             /* previous step end to this step end         */
    float event_ms;            /* HIP-event time of the launch, shared       */
                               /* equally among the levels it ran; 0 if off  */
    int32_t launch;            /* index of the level-kernel launch that ran  */
                               /* the step (solo launches run several)       */
} lgmi_bfs_step;

/* Statistics of the last betweenness run. */
typedef struct {
    int32_t groups;            /* 64-source lane groups processed            */
    int32_t max_levels;        /* deepest BFS level over all groups          */
    int32_t launch;            /* 2: one persistent launch replayed from a   */
                               /* captured hipGraph, 1: one persistent       */
                               /* launch, 0: one launch per level           */
    int32_t participants;      /* persistent launch: work-groups that took   */
                               /* part (those resident together at its       */
                               /* start); 0 otherwise                        */
    double device_ms;          /* HIP-event time of the whole computation    */
    double forward_ms;         /* sigma passes                                */
    double backward_ms;        /* delta passes                                */
    int32_t blocks;            /* persistent launch: work-groups launched    */
    int32_t pad;
} lgmi_bc_stats;

/* ----------------------------------------------------------------------
 * Graph handle.  Replaces the reference's fadjlist/badjlist access
 * (outneighbors/inneighbors, src/SimpleGraphs/SimpleGraphs.jl:78-79,92-93;
 * badj, simplegraph.jl:376-377, simpledigraph.jl:381-382) with a device CSR.
 * in_row_ptr/in_col are NULL for an undirected graph (the in-CSR is then the
 * out-CSR); for a directed graph they are the badjlist as a CSR.  The host
 * checks bounds, sorted rows and duplicates; the device then checks that an
 * undirected CSR is symmetric and that the in-CSR is the transpose of the
 * out-CSR (LGMI_E_ARG otherwise), an O(nnz log d) pass.
 * ---------------------------------------------------------------------- */
int lgmi_graph_create(int32_t nv, int64_t nnz, const int32_t* row_ptr, const int32_t* col,
                      int64_t in_nnz, const int32_t* in_row_ptr, const int32_t* in_col,
                      int directed, lgmi_graph_t* out);
void lgmi_graph_destroy(lgmi_graph_t g);

/* SimpleGraph(edge_list) / SimpleDiGraph(edge_list) built on the device
 *   src/SimpleGraphs/simplegraph.jl:163-244 (cleanupedges!, constructor)
 *   src/SimpleGraphs/simpledigraph.jl:174-253
 * from m edges (src[i], dst[i]), 0-based; an edge with a negative id is
 * skipped (the reference skips ids below 1).  nv = max(nv_min, largest id
 * used + 1) (the reference: nv_min = 0).  Undirected: each edge enters both
 * rows, a self-loop once; rows sorted, duplicates removed; *ne = the
 * reference's ne (an edge once, a self-loop once).  Directed: out- and
 * in-CSR, *ne = the number of distinct edges.  The CSR is built and kept on
 * the device; *out is a handle like lgmi_graph_create's. */
int lgmi_graph_from_edges(const int32_t* src, const int32_t* dst, int64_t m, int directed, int32_t nv_min,
                          lgmi_graph_t* out, int64_t* ne);
/* nv and the adjacency-entry count of a handle; the CSR copied to the host
 * (in_row_ptr / in_col may be NULL; for an undirected graph they repeat the
 * out-CSR). */
int lgmi_graph_size(lgmi_graph_t g, int32_t* nv, int64_t* nnz);
int lgmi_graph_csr(lgmi_graph_t g, int32_t* row_ptr, int32_t* col, int32_t* in_row_ptr, int32_t* in_col);

/* gdistances(g, source) / gdistances!(g, source, vert_level)
 *   src/traversals/bfs.jl:95-141
 * Parallel.gdistances(g, sources) / Parallel.gdistances!(...)
 *   src/Parallel/traversals/gdistances.jl:36-139
 * Distances along out-edges from all of srcs[0..nsrc) at once; dist[v] = -1
 * when v is unreached.  nsrc == 0 gives all -1.  Duplicate sources are
 * harmless.  LGMI_E_BOUNDS when a source is outside [0, nv). */
int lgmi_gdistances(lgmi_graph_t g, const int32_t* srcs, int32_t nsrc, int32_t* dist);

/* bfs_parents(g, s; dir)  -- src/traversals/bfs.jl:35-66
 * bfs_tree / Parallel.bfs_tree(!) use these parents
 *   (bfs.jl:76, src/Parallel/traversals/bfs.jl:78-104).
 * dir_in = 0 follows out-edges (dir = :out), 1 follows in-edges.
 * parent[src] = src; parent[v] = the smallest-id vertex u one level closer
 * to src with an edge u -> v, which is the vertex the reference's sorted
 * frontier discovers v from; unreached = -1. */
int lgmi_bfs_parents(lgmi_graph_t g, int32_t src, int dir_in, int32_t* parent);

/* Distances and parents of one single-source BFS in one pass (either
 * output may be NULL). */
int lgmi_bfs(lgmi_graph_t g, int32_t src, int dir_in, int32_t* dist, int32_t* parent);

/* Benchmark form: runs a single-source BFS and leaves dist/parent in the
 * handle's device buffers (no host copy).  want_parents = 0 skips the
 * parent array (gdistances).  stats may be NULL. */
int lgmi_bfs_run(lgmi_graph_t g, int32_t src, int dir_in, int want_parents, lgmi_bfs_stats* stats);
/* Copies the last run's results to the host; either pointer may be NULL. */
int lgmi_bfs_fetch(lgmi_graph_t g, int32_t* dist, int32_t* parent);

/* Many single-source searches in one call: the loop a caller writes over
 * gdistances / bfs_parents (src/traversals/bfs.jl:35-66, 95-141), one
 * search per root, each with the results of lgmi_bfs for that root.  Up to
 * eight searches run at once, each on a search state, HIP stream and host
 * thread of its own, so that one search's level launches fill the time that
 * another's launch boundaries and small levels leave the GPU.  dist and parent are
 * caller-owned DEVICE arrays of nroots x nv Int32: row i holds the
 * distances (-1 unreached) and parents (parent[root] = root, -1
 * unreached) of roots[i]; parent NULL: distances only.  The call returns
 * when every row is written; work the caller queued on dist / parent must
 * be complete before it.  LGMI_E_BOUNDS when a root is outside [0, nv).
 * The first call on a handle allocates the eight search states (about
 * nv / 2 bytes each plus 8 MB of queues); lgmi_bfs_fetch does not see
 * these searches. */
int lgmi_bfs_run_many(lgmi_graph_t g, const int32_t* roots, int32_t nroots, int dir_in, int32_t* dist,
                      int32_t* parent);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code that must not call hipMalloc: tight serving loops, harnesses
 * that forbid it): the single-source BFS of lgmi_bfs_run / lgmi_bfs with the
 * results written to caller-owned DEVICE buffers -- dist (nv Int32, -1
 * unreached) and parent (nv Int32, parent[src] = src, -1 unreached; NULL:
 * distances only) -- and every scratch buffer of the search (frontier and
 * visited bitmaps, per-work-group counters and queues, the level state, the
 * policy weights) in one fixed device array of LGMI_NOALLOC_SCRATCH_BYTES
 * that each GPU holds from the library's first use until the process ends.
 * The call allocates nothing (no hipMalloc, hipFree, hipHostMalloc, event or
 * graph creation).  nv must be at most LGMI_NOALLOC_MAX_NV (LGMI_E_TOO_LARGE
 * otherwise).  All calls on one GPU share its array: the library serialises
 * the calls it sees on a device, and calls that overlap in time on one GPU
 * by other means race.  Work the caller queued on dist / parent must be
 * complete before the call (the library runs on its own non-blocking
 * stream).  Results are identical to lgmi_bfs's and complete when the call
 * returns. */
#define LGMI_NOALLOC_MAX_NV         (1 << 24)
#define LGMI_NOALLOC_SCRATCH_BYTES  (17u << 20)
int lgmi_bfs_run_noalloc(lgmi_graph_t g, int32_t src, int dir_in, int32_t* dist, int32_t* parent);

/* Per-step records of the last BFS run (policy features and labels,
 * algorithmic-byte counters).  Writes up to max_steps records and returns
 * the number written (or a negative code). */
int lgmi_bfs_steps(lgmi_graph_t g, int32_t max_steps, lgmi_bfs_step* out);

/* enable = 1 brackets every step launch of later runs with HIP events
 * (per-launch durations in lgmi_bfs_step.event_ms and
 * lgmi_bfs_stats.steps_event_ms); 0 turns it off. */
int lgmi_bfs_time_steps(lgmi_graph_t g, int enable);

/* Forces every step of the next BFS runs to top-down (0), bottom-up (1) or
 * lets the policy decide (-1).  Outputs never change; only time does. */
int lgmi_bfs_force_direction(lgmi_graph_t g, int32_t dir);

/* One probed BFS level: the policy features of the state and the device
 * time of the step run top-down and bottom-up on that identical state. */
typedef struct {
    float feat[16];
    float td_us;
    float bu_us;
    int32_t level;
    int32_t n_front;
} lgmi_probe_rec;

/* Training-label collection for the direction policy (no counterpart in
 * the reference).  Runs a BFS from src; at every level listed in `levels`
 * (strictly increasing) it runs the step both ways on the same state
 * (undoing the first) and records the two HIP-event times.  Levels past the
 * last one are skipped.  Returns the number of records written. */
int lgmi_bfs_probe(lgmi_graph_t g, int32_t src, int dir_in, const int32_t* levels, int32_t nlevels,
                   lgmi_probe_rec* out);

/* Direction-policy MLP on the GPU (FP32, layout LGMI_POLICY_NWEIGHTS).
 * x: n rows of 16 features, y: labels (0 top-down faster, 1 bottom-up
 * faster), w: per-sample weights (NULL = 1).  Loss = weighted softmax
 * cross-entropy / sum of weights.  The feature mean/scale entries of
 * params are inputs and are never trained (their gradient is 0). */
int lgmi_mlp_loss_grad(const float* x, const int32_t* y, const float* w, int32_t n, const float* params,
                       float* loss, float* grad);
/* Full-batch Adam for `steps` steps at learning rate lr; params in/out. */
int lgmi_mlp_train(const float* x, const int32_t* y, const float* w, int32_t n, float* params, int32_t steps,
                   float lr, float* loss_out);
/* Logits (n x 2) of the policy for n feature rows. */
int lgmi_mlp_forward(const float* x, int32_t n, const float* params, float* logits);

/* Direction policy for all later BFS runs (process-wide).
 * mode: LGMI_POLICY_HEURISTIC (alpha, beta), LGMI_POLICY_LEARNED (mlp must
 * hold LGMI_POLICY_NWEIGHTS floats; the heuristic is used while none is
 * loaded) or LGMI_POLICY_TOPDOWN.  No counterpart in the reference, whose
 * BFS is top-down only (bfs.jl:108-122). */
int lgmi_bfs_set_policy(int mode, float alpha, float beta, const float* mlp, int32_t nweights);

/* betweenness_centrality(g, vs, DefaultDistance; normalize, endpoints)
 *   src/centrality/betweenness.jl:46-160 (k-sample form :77-78 with vs
 *   drawn by the caller)
 * Parallel.betweenness_centrality(...; parallel = :distributed | :threads)
 *   src/Parallel/centrality/betweenness.jl:1-75 (the shim passes the
 *   source list each mode actually uses)
 * Exact Brandes over srcs[0..nsrc) in FP64; duplicates count twice.
 * skip_degree0 = 1 skips sources with degree(g, s) == 0 as the serial and
 * :distributed functions do (betweenness.jl:58, Parallel/centrality/
 * betweenness.jl:23); 0 runs every listed source, as :threads does with the
 * positions findall returns (Parallel/centrality/betweenness.jl:53-56).
 * k_scale is the k of _rescale! (the reference's length(vs)).  Results
 * match the reference within 1e-12 relative error (only the floating-point
 * summation order differs).  LGMI_E_TOO_LARGE when the search state of one
 * 64-source lane group (about 1.8 KB per vertex) does not fit in the free
 * device memory. */
int lgmi_betweenness(lgmi_graph_t g, const int32_t* srcs, int32_t nsrc, int64_t k_scale,
                     int skip_degree0, int normalize, int endpoints, double* bc);
/* Replaces, for unweighted graphs:
 *   stress_centrality(g, vs)          src/centrality/stress.jl:33-48 (with
 *                                     _stress_accumulate_basic!, :53-77)
 *   stress_centrality(g, k)           stress.jl:50-51 (the shim draws the
 *                                     sample and passes it as srcs)
 *   Parallel.stress_centrality(...; parallel = :distributed | :threads)
 *                                     src/Parallel/centrality/stress.jl:1-51
 * Sources of degree(g, s) == 0 are skipped (stress.jl:41); duplicates count
 * twice.  stress[v] (nv entries) is the reference's Int64 sum, wrapping on
 * overflow as Julia's Int does; exact. */
int lgmi_stress(lgmi_graph_t g, const int32_t* srcs, int32_t nsrc, int64_t* stress);

/* connected_components!(label, g) / connected_components(g)
 * (src/connectivity.jl:13-95; is_connected :113-114): label[v] = the
 * smallest vertex id (0-based) of v's component -- weak components for a
 * directed graph, as the reference's all_neighbors gives -- which is the
 * label the reference's outer loop assigns.  The shim forms the component
 * lists from it (components, :52-64). */
int lgmi_connected_components(lgmi_graph_t g, int32_t* label);

/* has_path(g, u, v; exclude_vertices) (src/traversals/bfs.jl:150-172):
 * *result = 1 when v is reachable from u over out-neighbours without
 * entering an excluded vertex, or u == v; 0 when not, or when u or v is
 * excluded. */
int lgmi_has_path(lgmi_graph_t g, int32_t u, int32_t v, const int32_t* exclude, int32_t nexclude,
                  int* result);

/* neighborhood_dists(g, v, d; dir) / neighborhood(g, v, d; dir) with unit
 * lengths (src/connectivity.jl:634-702): the vertices within distance d of
 * v (d < 0: none) in the reference's FIFO queue order -- by level, and
 * within a level by the queue position of the first vertex of the previous
 * level that reaches them, then by adjacency order -- over out-neighbours,
 * or in-neighbours when dir_in.  verts (and dists when not NULL) need room
 * for nv entries; *count gets the number written. */
int lgmi_neighborhood(lgmi_graph_t g, int32_t v, int64_t d, int dir_in, int32_t* verts, int32_t* dists,
                      int32_t* count);

/* period(g) for a directed graph (src/connectivity.jl:495-516): LGMI_E_ARG
 * ("Graph must be strongly connected") unless vertex 0 reaches every vertex
 * and every vertex reaches it; else the gcd over the edges outside the BFS
 * tree from vertex 0 of level[src] - level[dst] + 1 (1 when there is a
 * self-loop). */
int lgmi_period(lgmi_graph_t g, int64_t* period);

/* Shortest-path distance statistics of every listed source (unit edge
 * lengths), 64 sources per lane group, forward pass only: for source
 * srcs[i], stats[3i] = the sum of the distances to the vertices it reaches,
 * stats[3i+1] = the number of vertices it reaches (itself included),
 * stats[3i+2] = the largest of those distances.  Exact.  The reference
 * computes one dijkstra_shortest_paths per source for each of
 *   closeness_centrality             src/centrality/closeness.jl:32-57
 *   radiality_centrality             src/centrality/radiality.jl:34-47
 *   eccentricity, diameter, radius, center, periphery
 *                                    src/distance.jl:74-174
 * and their Parallel forms (src/Parallel/centrality/{closeness,radiality}.jl,
 * src/Parallel/distance.jl); the host forms those results from these three
 * numbers (an unreachable vertex is typemax(Int) in the reference's
 * distance vector, which the host restores). */
int lgmi_distance_stats(lgmi_graph_t g, const int32_t* srcs, int32_t nsrc, int64_t* stats);

/* Statistics of the last betweenness (or stress) run on g. */
int lgmi_betweenness_stats(lgmi_graph_t g, lgmi_bc_stats* stats);

/* How lgmi_betweenness launches its work (process-wide; no counterpart in
 * the reference, whose source loop runs on the host):
 *   0 (default) -- all lane groups in one persistent launch, captured in a
 *                  hipGraph and replayed while the arguments stay the same,
 *                  whenever the groups fit the slots at once; one launch
 *                  per level otherwise, and as the fallback when the
 *                  persistent launch cannot finish;
 *   1           -- one launch per level always.
 * Both give bit-identical results. */
int lgmi_betweenness_set_launch(int mode);

/* All-source betweenness for many independent undirected graphs in one
 * persistent launch (each graph as betweenness_centrality(g) with
 * vs = vertices(g)).  Graph i has nv[i] vertices, its row pointer at
 * row_ptr_cat[row_off[i] .. row_off[i]+nv[i]] (nv[i]+1 entries, 0-based
 * within the graph) and its adjacency at col_cat[col_off[i] ..]; its
 * nv[i] results go to bc_cat[bc_off[i] ..].  device_ms may be NULL. */
int lgmi_betweenness_batch(int32_t ngraphs, const int32_t* nv, const int64_t* row_off,
                           const int32_t* row_ptr_cat, const int64_t* col_off,
                           const int32_t* col_cat, const int64_t* bc_off, int normalize,
                           int endpoints, double* bc_cat, double* device_ms);
/* Tuning of lgmi_betweenness_batch (process-wide; no reference counterpart):
 * graphs with at most wave_max_n vertices run one source per wavefront,
 * larger ones one source per work-group (-1: the largest n whose 16 sigma
 * slices fit in LDS, at most 1024); task_work_target is the n * (nnz + n)
 * work above which a graph's sources are split over several tasks
 * (0: the largest graph's work / 128, doubled until the tasks' partial
 * arrays, one FP64 value per vertex each, fit in 8 GB); graphs with more
 * than lds_max_n
 * vertices keep the per-vertex search state in global memory instead of LDS
 * (-1: as many as LDS holds; values above that are clamped); lane_group
 * (1, 2, 4 or 8; 0: 2) is the number of lanes that expand one adjacency
 * row, each taking `slots` entries per round (2, 4 or 8, or 16 with one
 * lane per row; 0: 8), rows
 * longer than 2 * lane_group * slots being expanded by a whole wavefront. */
int lgmi_betweenness_batch_tuning(int32_t wave_max_n, double task_work_target, int32_t lds_max_n,
                                  int32_t lane_group, int32_t slots);

/* Launch thresholds of the BFS steps and of the betweenness heavy-row
 * segments (process-wide).  No counterpart in the reference, whose threaded
 * BFS has one such knob, queue_segment_size (src/Parallel/traversals/
 * gdistances.jl:135-139).  The autotuner (lgmi/tuning.py) writes its
 * choice to lgmi/tuning.json, which the Python package applies before it
 * creates a graph handle; the Julia shim does the same.  No value changes
 * a result, only time.  Out-of-range values give LGMI_E_ARG. */
typedef struct {
    int32_t wave_row;          /* top-down sweep: rows of at least this many  */
                               /* entries are expanded by a whole wavefront   */
                               /* (default 32; 2..257)                        */
    int32_t lane_rounds;       /* bottom-up: rounds of four entries a         */
                               /* candidate scans in its own lane before a    */
                               /* whole wavefront takes its row (default 4;   */
                               /* 0..64)                                      */
    int32_t range_front;       /* range mode: frontier vertices at most       */
                               /* (default 2048; 0..2048, 0 turns it off)     */
    int32_t queue_front;       /* queue mode: frontier vertices at most       */
                               /* (default 4096; 0..4096, 0 turns it off)     */
    int64_t range_light;       /* range mode: frontier edges in rows of at    */
                               /* most 256 entries, at most (default 32768)   */
    int64_t queue_min_edges;   /* queue mode: frontier edges at least         */
    int64_t queue_max_edges;   /* queue mode: frontier edges at most          */
    int32_t bc_seg_len;        /* betweenness: rows longer than this are cut  */
                               /* into segments of this many entries (default */
                               /* 1024; 64..65536; fixed for a graph at its   */
                               /* first betweenness call)                     */
    int32_t solo_front;        /* solo mode: frontier vertices at most        */
                               /* (default 4096; 0..4096, 0 turns it off)     */
    int64_t solo_edges;        /* solo mode: frontier edges at most (default  */
                               /* 2048; 0..262144).  In solo mode one work-   */
                               /* group runs consecutive small top-down       */
                               /* levels inside one launch                    */
} lgmi_tuning;
/* t = NULL restores the built-in defaults. */
int lgmi_set_tuning(const lgmi_tuning* t);
int lgmi_get_tuning(lgmi_tuning* t);

/* Thread-local message of the last failed call ("" if none). */
const char* lgmi_last_error(void);

/* Number of visible GPUs (0 when none; never fails). */
int lgmi_device_count(void);

/* Blocks until all work the library queued has finished. */
int lgmi_synchronize(void);

#ifdef __cplusplus
}
#endif

#endif /* LGMI_H */
