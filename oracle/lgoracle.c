/*
 * lgoracle.c -- CPU oracle for the BFS / exact-Brandes path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference algorithms (LightGraphs.jl v1.3.5, Julia) that the HIP library
 * replaces.  Only tests/, the smoke() check and the CPU-baseline leg of
 * bench.py may load it, and only as the checker or as the timed CPU
 * baseline; the product path never calls it.
 *
 * Conventions: vertices are 0-based here (the reference is 1-based); an
 * "unreached" distance is left untouched (serial gdistances!) or written as
 * -1 where the reference writes typemax(T); an unreached parent is -1 where
 * the reference writes 0.  Adjacency arrives as a CSR (row_ptr, col) whose
 * rows are sorted ascending, which is the reference's fadjlist/badjlist
 * invariant (src/SimpleGraphs/simplegraph.jl:445-461,
 * src/SimpleGraphs/simpledigraph.jl:414-430).
 *
 * Parity pinning (what checks each function; the tests live in
 * tests/test_oracle.py):
 *   orc_gdistances       reference vectors test/traversals/bfs.jl:31-33,
 *                        test/parallel/traversals/gdistances.jl:7-17; networkx
 *                        single_source_shortest_path_length on random graphs.
 *   orc_bfs_parents      test/traversals/bfs.jl:12 ([1,1,1,3]); brute-force
 *                        min-predecessor closed form on random graphs.
 *   orc_par_gdistances   equals orc_gdistances with typemax prefill (the
 *                        reference's own test asserts that equality).
 *   orc_par_bfs_tree     test/parallel/traversals/bfs.jl:26 ([1,1,1,3]); tree
 *                        validity (depth = BFS distance) on random graphs.
 *   orc_par_gdistances_threaded  equals orc_gdistances on the reference
 *                        vectors above and on random graphs, for 1 to 16
 *                        threads and segment sizes 1 to 1000.
 *   orc_par_bfs_tree_threaded  test/parallel/traversals/bfs.jl:26; tree
 *                        validity on random graphs for 1 to 16 threads.
 *   orc_dijkstra_allpaths test/shortestpaths/dijkstra.jl:81-97 (pathcounts,
 *                        predecessors, closest_vertices); sigma vs. brute-force
 *                        path enumeration on tiny graphs.
 *   orc_betweenness      fixture test/testdata/graph-50-500-bc.txt; endpoints
 *                        values test/centrality/betweenness.jl:16-17; self-loop
 *                        and doctest cases; networkx golden vectors; brute
 *                        force over all shortest paths on tiny graphs.
 *   orc_betweenness_threaded  equals orc_betweenness when vs = 1:n (the
 *                        reference test's only use); index-not-vertex quirk
 *                        pinned by a hand-derived case.
 *   orc_stress           fixture test/testdata/graph-50-500-sc.txt (exact
 *                        Int64); test/centrality/stress.jl:17-24 case
 *                        [4, 2, 4, 10, 0]; doctests stress.jl:18-30; a
 *                        direct restatement of the same recursion from the
 *                        enumerated shortest-path DAG on random graphs.
 */
#define _POSIX_C_SOURCE 200809L   /* pthread_barrier_t */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <pthread.h>

/* ------------------------------------------------------------------ */
/* helpers                                                              */
/* ------------------------------------------------------------------ */

static int cmp_i32(const void* a, const void* b) {
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

/* degree(g, v): out-degree for undirected graphs (fadjlist length, a
 * self-loop counted once as stored), in+out for directed graphs
 * (src/core.jl:124-128, trait dispatch on IsDirected). */
static int64_t degree_of(int32_t v, const int32_t* row_ptr, const int32_t* in_row_ptr,
                         int directed) {
    int64_t d = (int64_t)row_ptr[v + 1] - row_ptr[v];
    if (directed) d += (int64_t)in_row_ptr[v + 1] - in_row_ptr[v];
    return d;
}

/* ------------------------------------------------------------------ */
/* gdistances! -- src/traversals/bfs.jl:95-124                          */
/* ------------------------------------------------------------------ */
/* Writes level numbers only for reached vertices; vert_level must be
 * pre-filled by the caller (bfs.jl:141 fills typemax).  Frontier is sorted
 * after each level (bfs.jl:121); that affects no output. */
int orc_gdistances(int32_t nv, const int32_t* row_ptr, const int32_t* col,
                   const int32_t* srcs, int32_t nsrc, int32_t* vert_level) {
    uint8_t* visited = (uint8_t*)calloc((size_t)nv + 1, 1);
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nv + (size_t)nsrc + 1));
    int32_t* nxt = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nv + 1));
    if (!visited || !cur || !nxt) { free(visited); free(cur); free(nxt); return -1; }
    int64_t ncur = 0, nnxt = 0;
    for (int32_t i = 0; i < nsrc; i++) {            /* bfs.jl:103-107 */
        int32_t s = srcs[i];
        if (s < 0 || s >= nv) { free(visited); free(cur); free(nxt); return -2; }
        vert_level[s] = 0;
        visited[s] = 1;
        cur[ncur++] = s;
    }
    int32_t n_level = 1;
    while (ncur > 0) {                               /* bfs.jl:108 */
        for (int64_t k = 0; k < ncur; k++) {
            int32_t v = cur[k];
            for (int32_t e = row_ptr[v]; e < row_ptr[v + 1]; e++) {
                int32_t i = col[e];
                if (!visited[i]) {                   /* bfs.jl:111-115 */
                    nxt[nnxt++] = i;
                    vert_level[i] = n_level;
                    visited[i] = 1;
                }
            }
        }
        n_level++;
        int32_t* t = cur; cur = nxt; nxt = t;
        ncur = nnxt; nnxt = 0;
        qsort(cur, (size_t)ncur, sizeof(int32_t), cmp_i32);   /* bfs.jl:121 */
    }
    free(visited); free(cur); free(nxt);
    return 0;
}

/* ------------------------------------------------------------------ */
/* _bfs_parents -- src/traversals/bfs.jl:38-66                          */
/* ------------------------------------------------------------------ */
/* (row_ptr, col) is the CSR of neighborfn: out-neighbours for dir=:out,
 * in-neighbours otherwise (bfs.jl:35-36).  parent[src] = src, unreached
 * = -1 (the reference's 0 in 1-based ids). */
int orc_bfs_parents(int32_t nv, const int32_t* row_ptr, const int32_t* col,
                    int32_t src, int32_t* parents) {
    if (src < 0 || src >= nv) return -2;
    uint8_t* visited = (uint8_t*)calloc((size_t)nv, 1);
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nv + 1));
    int32_t* nxt = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nv + 1));
    if (!visited || !cur || !nxt) { free(visited); free(cur); free(nxt); return -1; }
    for (int32_t i = 0; i < nv; i++) parents[i] = -1;  /* zeros(T, nv) */
    int64_t ncur = 0, nnxt = 0;
    visited[src] = 1; cur[ncur++] = src; parents[src] = src;   /* bfs.jl:46-50 */
    while (ncur > 0) {
        for (int64_t k = 0; k < ncur; k++) {
            int32_t v = cur[k];
            for (int32_t e = row_ptr[v]; e < row_ptr[v + 1]; e++) {
                int32_t i = col[e];
                if (!visited[i]) {                   /* bfs.jl:54-58 */
                    nxt[nnxt++] = i;
                    parents[i] = v;
                    visited[i] = 1;
                }
            }
        }
        int32_t* t = cur; cur = nxt; nxt = t;
        ncur = nnxt; nnxt = 0;
        qsort(cur, (size_t)ncur, sizeof(int32_t), cmp_i32);   /* bfs.jl:63 */
    }
    free(visited); free(cur); free(nxt);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Parallel.gdistances! -- src/Parallel/traversals/gdistances.jl:36-116  */
/* ------------------------------------------------------------------ */
/* Level-synchronous sweep with typemax pre-fill (:46).  The threaded
 * segment claiming (:71-101) only changes which thread expands a vertex;
 * the level of every vertex is unique, so a serial sweep over the
 * partitioned queues gives the same vert_level.  Unreached = -1. */
int orc_par_gdistances(int32_t nv, const int32_t* row_ptr, const int32_t* col,
                       const int32_t* srcs, int32_t nsrc, int32_t* vert_level) {
    for (int32_t i = 0; i < nv; i++) vert_level[i] = -1;   /* fill!(typemax) */
    uint8_t* visited = (uint8_t*)calloc((size_t)nv + 1, 1);
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nv + (size_t)nsrc + 1));
    int32_t* nxt = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nv + 1));
    if (!visited || !cur || !nxt) { free(visited); free(cur); free(nxt); return -1; }
    int64_t ncur = 0, nnxt = 0;
    for (int32_t i = 0; i < nsrc; i++) {            /* :55-58 */
        int32_t s = srcs[i];
        if (s < 0 || s >= nv) { free(visited); free(cur); free(nxt); return -2; }
        visited[s] = 1; vert_level[s] = 0; cur[ncur++] = s;
    }
    int32_t n_level = 0;
    while (ncur > 0) {
        n_level++;
        for (int64_t k = 0; k < ncur; k++) {
            int32_t v = cur[k];
            if (!(visited[v] && vert_level[v] == n_level - 1)) continue;  /* :88 */
            for (int32_t e = row_ptr[v]; e < row_ptr[v + 1]; e++) {
                int32_t i = col[e];
                if (!visited[i]) { vert_level[i] = n_level; visited[i] = 1; nxt[nnxt++] = i; }
            }
        }
        int32_t* t = cur; cur = nxt; nxt = t;
        ncur = nnxt; nnxt = 0;
    }
    free(visited); free(cur); free(nxt);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Parallel.bfs_tree! -- src/Parallel/traversals/bfs.jl:49-92           */
/* ------------------------------------------------------------------ */
/* Serial restatement of the frontier loop: bfskernel visits `level` in
 * queue order and atomic_cas!(parents[v], 0, src) first-wins (:59).  The
 * threaded reference is non-deterministic; this serial order is one of its
 * legal outcomes.  (row_ptr, col) = neighbors(g, .) (:56), i.e.
 * out-neighbours.  Unset parents stay -1 (the reference's 0). */
int orc_par_bfs_tree(int32_t nv, const int32_t* row_ptr, const int32_t* col,
                     int32_t source, int32_t* parents) {
    if (source < 0 || source >= nv) return -2;
    int32_t* q = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nv + 1));
    if (!q) return -1;
    for (int32_t i = 0; i < nv; i++) parents[i] = -1;
    int64_t head = 0, tail = 0;
    parents[source] = source;                        /* :84 */
    q[tail++] = source;                              /* :85 */
    while (head != tail) {                           /* :86 */
        int64_t lo = head, hi = tail;                /* :87-88 */
        head = tail;
        for (int64_t k = lo; k < hi; k++) {          /* bfskernel :55-64 */
            int32_t src = q[k];
            for (int32_t e = row_ptr[src]; e < row_ptr[src + 1]; e++) {
                int32_t v = col[e];
                if (parents[v] == -1) { parents[v] = src; q[tail++] = v; }
            }
        }
    }
    free(q);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Threaded Parallel.gdistances! and Parallel.bfs_tree! (CPU baseline)   */
/* ------------------------------------------------------------------ */
/* The two functions above give the reference's outputs; these two run the
 * reference's threaded schedule on `nthreads` POSIX threads so that it can
 * be timed.  The reference's benign data races (unsynchronised reads and
 * writes of visited, vert_level, the queue entries and the segment
 * counters) become relaxed atomic loads and stores of the same locations:
 * the same schedule, without undefined behaviour in C.  Julia's `@threads`
 * per level becomes a pool of threads that meet at a barrier twice per
 * level. */
#define RLX __ATOMIC_RELAXED

typedef struct {
    int32_t nv; const int32_t* row_ptr; const int32_t* col;
    int nt; int32_t seg;
    int32_t** cur; int64_t* ncur;      /* cur_level_t, length(cur_level_t[t]) */
    int32_t** nxt; int64_t* nnxt;      /* next_level_t */
    int64_t* front;                    /* cur_front_t, 16 int64 apart */
    uint8_t* explored;                 /* queue_explored_t, 64 B apart */
    uint8_t* visited; int32_t* level;
    pthread_barrier_t bar; int done; int32_t n_level; int rc;
} pgd_pool;

typedef struct { pgd_pool* p; int tid; } pgd_arg;

/* One level of the @threads body, gdistances.jl:67-102, for thread `tid`:
 * its own queue first, then the others' (:71), each one segment claim at a
 * time (:76-78); a claimed run continues until it meets an entry already
 * taken (:81-85).  Taken entries are overwritten with -1 (the reference's
 * zero(T) in 1-based ids). */
static void pgd_level(pgd_pool* p, int tid) {
    const int nt = p->nt;
    const int32_t n_level = p->n_level;
    int32_t* next = p->nxt[tid];
    int64_t nn = p->nnxt[tid];
    for (int k = 0; k < nt; k++) {
        const int t = (tid + k) % nt;
        if (__atomic_load_n(&p->explored[(size_t)t * 64], RLX)) continue;   /* :72 */
        int32_t* cur = p->cur[t];
        const int64_t len = p->ncur[t];
        for (;;) {
            int64_t lf = __atomic_fetch_add(&p->front[(size_t)t * 16], (int64_t)p->seg, RLX);   /* :77-78 */
            if (lf >= len) break;                                            /* :80 */
            while (lf < len) {
                const int32_t v = __atomic_load_n(&cur[lf], RLX);
                if (v < 0) break;                                            /* :81 */
                __atomic_store_n(&cur[lf], -1, RLX);                         /* :83 */
                lf++;
                if (!(__atomic_load_n(&p->visited[v], RLX) &&
                      __atomic_load_n(&p->level[v], RLX) == n_level - 1)) continue;   /* :88 */
                for (int32_t e = p->row_ptr[v]; e < p->row_ptr[v + 1]; e++) {  /* :89-97 */
                    const int32_t i = p->col[e];
                    if (!__atomic_load_n(&p->visited[i], RLX)) {
                        __atomic_store_n(&p->level[i], n_level, RLX);
                        __atomic_store_n(&p->visited[i], (uint8_t)1, RLX);
                        next[nn++] = i;
                    }
                }
            }
        }
        __atomic_store_n(&p->explored[(size_t)t * 64], (uint8_t)1, RLX);     /* :101 */
    }
    p->nnxt[tid] = nn;
}

/* Between levels (:105-113), on thread 0 while the others wait: swap the
 * queues, reset the segment counters, mark empty queues explored. */
static void pgd_swap(pgd_pool* p) {
    int empty = 1;
    for (int t = 0; t < p->nt; t++) {
        int32_t* tmp = p->cur[t]; p->cur[t] = p->nxt[t]; p->nxt[t] = tmp;
        p->ncur[t] = p->nnxt[t]; p->nnxt[t] = 0;
        p->front[(size_t)t * 16] = 0;
        p->explored[(size_t)t * 64] = p->ncur[t] == 0;
        empty = empty && p->ncur[t] == 0;
    }
    p->done = empty;
    p->n_level++;
}

static void* pgd_worker(void* arg) {
    pgd_arg* a = (pgd_arg*)arg;
    pgd_pool* p = a->p;
    for (;;) {
        pthread_barrier_wait(&p->bar);          /* level start: queues ready */
        if (p->done) break;
        pgd_level(p, a->tid);
        if (pthread_barrier_wait(&p->bar) == PTHREAD_BARRIER_SERIAL_THREAD) pgd_swap(p);
    }
    return NULL;
}

/* Parallel.gdistances!(g, sources, vert_level; queue_segment_size) --
 * src/Parallel/traversals/gdistances.jl:36-116.  Unreached = -1.  Two
 * threads that find a vertex together both push it (the reference's race,
 * :91-95), but one thread pushes a vertex at most once in the whole search
 * (it sees its own store to visited), so each thread's queue holds at most
 * max(nv, nsrc) entries. */
int orc_par_gdistances_threaded(int32_t nv, const int32_t* row_ptr, const int32_t* col,
                                const int32_t* srcs, int32_t nsrc, int nthreads, int32_t seg,
                                int32_t* vert_level) {
    if (nthreads < 1) nthreads = 1;
    if (seg < 1) seg = 1;
    for (int32_t i = 0; i < nsrc; i++) if (srcs[i] < 0 || srcs[i] >= nv) return -2;
    pgd_pool p;
    memset(&p, 0, sizeof(p));
    p.nv = nv; p.row_ptr = row_ptr; p.col = col; p.nt = nthreads; p.seg = seg;
    const size_t qcap = (size_t)(nv > nsrc ? nv : nsrc) + 1;
    p.cur = (int32_t**)calloc((size_t)nthreads, sizeof(int32_t*));
    p.nxt = (int32_t**)calloc((size_t)nthreads, sizeof(int32_t*));
    p.ncur = (int64_t*)calloc((size_t)nthreads, sizeof(int64_t));
    p.nnxt = (int64_t*)calloc((size_t)nthreads, sizeof(int64_t));
    p.front = (int64_t*)calloc((size_t)nthreads * 16, sizeof(int64_t));
    p.explored = (uint8_t*)calloc((size_t)nthreads * 64, 1);
    p.visited = (uint8_t*)calloc((size_t)nv + 1, 1);
    pgd_arg* args = (pgd_arg*)calloc((size_t)nthreads, sizeof(pgd_arg));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    int ok = p.cur && p.nxt && p.ncur && p.nnxt && p.front && p.explored && p.visited && args && th;
    for (int t = 0; ok && t < nthreads; t++) {
        p.cur[t] = (int32_t*)malloc(sizeof(int32_t) * qcap);
        p.nxt[t] = (int32_t*)malloc(sizeof(int32_t) * qcap);
        ok = p.cur[t] && p.nxt[t];
    }
    int rc = ok ? 0 : -1;
    if (ok) {
        p.level = vert_level;
        for (int32_t i = 0; i < nv; i++) vert_level[i] = -1;      /* :46 */
        for (int32_t i = 0; i < nsrc; i++) { p.visited[srcs[i]] = 1; vert_level[srcs[i]] = 0; }   /* :55-58 */
        /* partition_sources! (:8-19): contiguous parts of `sources` */
        for (int t = 0; t < nthreads; t++) {
            const int64_t lo = (int64_t)nsrc * t / nthreads, hi = (int64_t)nsrc * (t + 1) / nthreads;
            for (int64_t k = lo; k < hi; k++) p.cur[t][p.ncur[t]++] = srcs[k];
            p.explored[(size_t)t * 64] = hi == lo;
        }
        p.done = nsrc == 0;
        p.n_level = 1;
        pthread_barrier_init(&p.bar, NULL, (unsigned)nthreads);
        for (int t = 1; t < nthreads; t++) {
            args[t].p = &p; args[t].tid = t;
            pthread_create(&th[t], NULL, pgd_worker, &args[t]);
        }
        args[0].p = &p; args[0].tid = 0;
        pgd_worker(&args[0]);
        for (int t = 1; t < nthreads; t++) pthread_join(th[t], NULL);
        pthread_barrier_destroy(&p.bar);
    }
    for (int t = 0; p.cur && p.nxt && t < nthreads; t++) { free(p.cur[t]); free(p.nxt[t]); }
    free(p.cur); free(p.nxt); free(p.ncur); free(p.nnxt); free(p.front); free(p.explored);
    free(p.visited); free(args); free(th);
    return rc;
}

typedef struct {
    const int32_t* row_ptr; const int32_t* col; int nt;
    int32_t* q; int64_t head; int64_t tail;       /* ThreadQueue (bfs.jl:15-19) */
    int32_t* parents;
    pthread_barrier_t bar; int done;
    int64_t lo, hi;                               /* the level being expanded */
} pbt_pool;

typedef struct { pbt_pool* p; int tid; } pbt_arg;

static void* pbt_worker(void* arg) {
    pbt_arg* a = (pbt_arg*)arg;
    pbt_pool* p = a->p;
    for (;;) {
        pthread_barrier_wait(&p->bar);
        if (p->done) break;
        /* bfskernel (bfs.jl:49-65): `@threads for src in level` gives each
         * thread one contiguous share of the level (Julia's static schedule) */
        const int64_t n = p->hi - p->lo;
        const int64_t a0 = p->lo + n * a->tid / p->nt, a1 = p->lo + n * (a->tid + 1) / p->nt;
        for (int64_t k = a0; k < a1; k++) {
            const int32_t src = p->q[k];
            for (int32_t e = p->row_ptr[src]; e < p->row_ptr[src + 1]; e++) {
                const int32_t v = p->col[e];
                int32_t expect = -1;              /* atomic_cas!(parents[v], zero(T), src) :59 */
                if (__atomic_load_n(&p->parents[v], RLX) == -1 &&
                    __atomic_compare_exchange_n(&p->parents[v], &expect, src, 0, RLX, RLX)) {
                    const int64_t off = __atomic_fetch_add(&p->tail, 1, RLX);   /* push! :27-31 */
                    __atomic_store_n(&p->q[off], v, RLX);
                }
            }
        }
        if (pthread_barrier_wait(&p->bar) == PTHREAD_BARRIER_SERIAL_THREAD) {
            /* bfs_tree! loop (:86-90): next level = q[head:tail) */
            p->lo = p->head = p->hi;
            p->hi = p->tail;
            p->done = p->lo == p->hi;
        }
    }
    return NULL;
}

/* Parallel.bfs_tree!(next, g, source, parents) -- src/Parallel/traversals/
 * bfs.jl:78-92 on `nthreads` threads.  The parents are one legal outcome of
 * the reference's first-CAS-wins race; unset = -1.  Each vertex enters the
 * queue once (the CAS), so nv entries suffice.  The reference's
 * `isempty` test (:38-40) is the level-empty test here. */
int orc_par_bfs_tree_threaded(int32_t nv, const int32_t* row_ptr, const int32_t* col,
                              int32_t source, int nthreads, int32_t* parents) {
    if (nthreads < 1) nthreads = 1;
    if (source < 0 || source >= nv) return -2;
    pbt_pool p;
    memset(&p, 0, sizeof(p));
    p.row_ptr = row_ptr; p.col = col; p.nt = nthreads; p.parents = parents;
    p.q = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nv + 1));
    pbt_arg* args = (pbt_arg*)calloc((size_t)nthreads, sizeof(pbt_arg));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    if (!p.q || !args || !th) { free(p.q); free(args); free(th); return -1; }
    for (int32_t i = 0; i < nv; i++) parents[i] = -1;
    parents[source] = source;                       /* :84 */
    p.q[0] = source; p.tail = 1;                    /* :85 */
    p.lo = 0; p.hi = 1; p.head = 1;
    pthread_barrier_init(&p.bar, NULL, (unsigned)nthreads);
    for (int t = 1; t < nthreads; t++) {
        args[t].p = &p; args[t].tid = t;
        pthread_create(&th[t], NULL, pbt_worker, &args[t]);
    }
    args[0].p = &p; args[0].tid = 0;
    pbt_worker(&args[0]);
    for (int t = 1; t < nthreads; t++) pthread_join(th[t], NULL);
    pthread_barrier_destroy(&p.bar);
    free(p.q); free(args); free(th);
    return 0;
}
#undef RLX

/* ------------------------------------------------------------------ */
/* Binary-heap priority queue with a key index.                         */
/* ------------------------------------------------------------------ */
/* Restates the published algorithm of DataStructures.jl's PriorityQueue
 * (compat "0.17, 0.18", /root/reference Project.toml:19; used at
 * src/shortestpaths/dijkstra.jl:73,80,87,106,116): a 1-based binary
 * min-heap of (key, priority) pairs, percolate-up on insert, root swap with
 * the last element and percolate-down on dequeue, comparisons on priority
 * only with strict less-than.  The heap decides only the dequeue order of
 * equal-distance vertices, which changes no distance or path count. */
typedef struct { int32_t key; int64_t pri; } pq_item;
typedef struct { pq_item* xs; int64_t n; int32_t* index; } pq_t;

static void pq_up(pq_t* pq, int64_t i) {
    pq_item x = pq->xs[i];
    while (i > 1) {
        int64_t j = i >> 1;
        pq_item xj = pq->xs[j];
        if (x.pri < xj.pri) { pq->index[xj.key] = (int32_t)i; pq->xs[i] = xj; i = j; }
        else break;
    }
    pq->index[x.key] = (int32_t)i;
    pq->xs[i] = x;
}

static void pq_down(pq_t* pq, int64_t i) {
    pq_item x = pq->xs[i];
    int64_t l;
    while ((l = 2 * i) <= pq->n) {
        int64_t r = l + 1;
        int64_t j = (r > pq->n || pq->xs[l].pri < pq->xs[r].pri) ? l : r;
        if (pq->xs[j].pri < x.pri) { pq->index[pq->xs[j].key] = (int32_t)i; pq->xs[i] = pq->xs[j]; i = j; }
        else break;
    }
    pq->index[x.key] = (int32_t)i;
    pq->xs[i] = x;
}

static void pq_set(pq_t* pq, int32_t key, int64_t pri) {
    if (pq->index[key] > 0) {                        /* existing key */
        int64_t i = pq->index[key];
        int64_t old = pq->xs[i].pri;
        pq->xs[i].pri = pri;
        if (old < pri) pq_down(pq, i); else pq_up(pq, i);
    } else {
        pq->n++;
        pq->xs[pq->n].key = key; pq->xs[pq->n].pri = pri;
        pq->index[key] = (int32_t)pq->n;
        pq_up(pq, pq->n);
    }
}

static int32_t pq_pop(pq_t* pq) {
    pq_item x = pq->xs[1];
    pq_item y = pq->xs[pq->n];
    pq->n--;
    if (pq->n > 0) {
        pq->xs[1] = y;
        pq->index[y.key] = 1;
        pq_down(pq, 1);
    }
    pq->index[x.key] = 0;
    return x.key;
}

/* ------------------------------------------------------------------ */
/* dijkstra_shortest_paths(g, s, DefaultDistance; allpaths, trackvertices)
 * -- src/shortestpaths/dijkstra.jl:59-141                               */
/* ------------------------------------------------------------------ */
/* Workspace reused across sources.  preds[v] is stored in the slot range
 * of v's in-neighbour row (in_row_ptr): every predecessor is an
 * in-neighbour, so the row bounds the count. */
typedef struct {
    int32_t nv;
    int64_t* dists;      /* -1 = typemax */
    uint8_t* visited;
    double* pathcounts;
    int32_t* pred_cnt;
    int32_t* preds;      /* nnz(in) slots */
    int32_t* closest;    /* closest_vertices, length nv */
    int32_t nclosest;
    pq_t pq;
    double* delta;
} dijk_ws;

static int dijk_ws_init(dijk_ws* w, int32_t nv, int64_t in_nnz) {
    memset(w, 0, sizeof(*w));
    w->nv = nv;
    w->dists = (int64_t*)malloc(sizeof(int64_t) * ((size_t)nv + 1));
    w->visited = (uint8_t*)malloc((size_t)nv + 1);
    w->pathcounts = (double*)malloc(sizeof(double) * ((size_t)nv + 1));
    w->pred_cnt = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nv + 1));
    w->preds = (int32_t*)malloc(sizeof(int32_t) * ((size_t)in_nnz + 1));
    w->closest = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nv + 1));
    w->pq.xs = (pq_item*)malloc(sizeof(pq_item) * ((size_t)nv + 2));
    w->pq.index = (int32_t*)calloc((size_t)nv + 1, sizeof(int32_t));
    w->delta = (double*)malloc(sizeof(double) * ((size_t)nv + 1));
    if (!w->dists || !w->visited || !w->pathcounts || !w->pred_cnt || !w->preds ||
        !w->closest || !w->pq.xs || !w->pq.index || !w->delta) return -1;
    return 0;
}

static void dijk_ws_free(dijk_ws* w) {
    free(w->dists); free(w->visited); free(w->pathcounts); free(w->pred_cnt);
    free(w->preds); free(w->closest); free(w->pq.xs); free(w->pq.index); free(w->delta);
}

static void dijkstra_allpaths(dijk_ws* w, const int32_t* row_ptr, const int32_t* col,
                              const int32_t* in_row_ptr, int32_t s) {
    int32_t nv = w->nv;
    for (int32_t i = 0; i < nv; i++) {               /* :67-72 */
        w->dists[i] = -1; w->visited[i] = 0; w->pathcounts[i] = 0.0; w->pred_cnt[i] = 0;
    }
    w->pq.n = 0;
    w->dists[s] = 0; w->visited[s] = 1; w->pathcounts[s] = 1.0;   /* :76-81 */
    pq_set(&w->pq, s, 0);
    w->nclosest = 0;
    while (w->pq.n > 0) {                            /* :86 */
        int32_t u = pq_pop(&w->pq);
        w->closest[w->nclosest++] = u;               /* :89-91 */
        int64_t d = w->dists[u];
        for (int32_t e = row_ptr[u]; e < row_ptr[u + 1]; e++) {
            int32_t v = col[e];
            int64_t alt = d + 1;                     /* DefaultDistance == 1 */
            if (!w->visited[v]) {                    /* :97-106 */
                w->visited[v] = 1;
                w->dists[v] = alt;
                w->pathcounts[v] += w->pathcounts[u];
                w->preds[in_row_ptr[v]] = u; w->pred_cnt[v] = 1;
                pq_set(&w->pq, v, alt);
            } else if (alt < w->dists[v]) {          /* :107-116 */
                w->dists[v] = alt;
                w->pathcounts[v] = w->pathcounts[u];
                w->preds[in_row_ptr[v]] = u; w->pred_cnt[v] = 1;
                pq_set(&w->pq, v, alt);
            } else if (alt == w->dists[v]) {         /* :117-121 */
                w->pathcounts[v] += w->pathcounts[u];
                w->preds[in_row_ptr[v] + w->pred_cnt[v]] = u; w->pred_cnt[v]++;
            }
        }
    }
    for (int32_t v = 0; v < nv; v++)                 /* :126-132 */
        if (!w->visited[v]) w->closest[w->nclosest++] = v;
    w->pathcounts[s] = 1.0;                          /* :134-138 */
    w->pred_cnt[s] = 0;
}

/* Exported form for tests (pins pathcounts / predecessors / closest order
 * against test/shortestpaths/dijkstra.jl:81-97).  preds_out uses the
 * in_row_ptr slots; pred_cnt_out gives the count per vertex. */
int orc_dijkstra_allpaths(int32_t nv, const int32_t* row_ptr, const int32_t* col,
                          const int32_t* in_row_ptr, int32_t s,
                          int64_t* dists, double* pathcounts, int32_t* pred_cnt,
                          int32_t* preds, int32_t* closest) {
    if (s < 0 || s >= nv) return -2;
    dijk_ws w;
    if (dijk_ws_init(&w, nv, in_row_ptr[nv]) != 0) { dijk_ws_free(&w); return -1; }
    dijkstra_allpaths(&w, row_ptr, col, in_row_ptr, s);
    memcpy(dists, w.dists, sizeof(int64_t) * (size_t)nv);
    memcpy(pathcounts, w.pathcounts, sizeof(double) * (size_t)nv);
    memcpy(pred_cnt, w.pred_cnt, sizeof(int32_t) * (size_t)nv);
    memcpy(preds, w.preds, sizeof(int32_t) * (size_t)in_row_ptr[nv]);
    memcpy(closest, w.closest, sizeof(int32_t) * (size_t)nv);
    dijk_ws_free(&w);
    return 0;
}

/* _accumulate_basic! -- src/centrality/betweenness.jl:81-107 */
static void accumulate_basic(double* bc, dijk_ws* w, const int32_t* in_row_ptr, int32_t si) {
    int32_t nv = w->nv;
    for (int32_t i = 0; i < nv; i++) w->delta[i] = 0.0;
    for (int32_t k = w->nclosest - 1; k >= 0; k--) {   /* S = reverse(closest) */
        int32_t x = w->closest[k];
        double coeff = (1.0 + w->delta[x]) / w->pathcounts[x];
        const int32_t* P = w->preds + in_row_ptr[x];
        int32_t np = (x == si) ? 0 : w->pred_cnt[x];     /* P[si] = [] */
        for (int32_t j = 0; j < np; j++) {
            int32_t v = P[j];
            w->delta[v] += (w->pathcounts[v] * coeff);
        }
        if (x != si) bc[x] += w->delta[x];
    }
}

/* _accumulate_endpoints! -- src/centrality/betweenness.jl:109-134 */
static void accumulate_endpoints(double* bc, dijk_ws* w, const int32_t* in_row_ptr, int32_t si) {
    int32_t nv = w->nv;
    for (int32_t i = 0; i < nv; i++) w->delta[i] = 0.0;
    bc[si] += (double)(w->nclosest - 1);              /* length(S) - 1, :122 */
    for (int32_t k = w->nclosest - 1; k >= 0; k--) {
        int32_t x = w->closest[k];
        double coeff = (1.0 + w->delta[x]) / w->pathcounts[x];
        const int32_t* P = w->preds + in_row_ptr[x];
        int32_t np = w->pred_cnt[x];
        for (int32_t j = 0; j < np; j++) {
            int32_t v = P[j];
            w->delta[v] += w->pathcounts[v] * coeff;
        }
        if (x != si) bc[x] += (w->delta[x] + 1.0);
    }
}

/* _rescale! -- src/centrality/betweenness.jl:136-160 */
void orc_rescale(double* bc, int32_t n, int normalize, int directed, int64_t k) {
    int do_scale = 0;
    double scale = 0.0;
    if (normalize) {
        if (n <= 2) do_scale = 0;
        else { do_scale = 1; scale = 1.0 / (double)(((int64_t)n - 1) * ((int64_t)n - 2)); }
    } else {
        if (!directed) { do_scale = 1; scale = 1.0 / 2.0; }
        else do_scale = 0;
    }
    if (do_scale) {
        if (k > 0) scale = scale * (double)n / (double)k;
        for (int32_t i = 0; i < n; i++) bc[i] *= scale;
    }
}

/* betweenness_centrality(g, vs, DefaultDistance; normalize, endpoints)
 * -- src/centrality/betweenness.jl:46-75.  in_row_ptr is the in-neighbour
 * CSR (equal to row_ptr for undirected graphs): it provides in-degree for
 * degree() and the predecessor slots. */
int orc_betweenness(int32_t nv, const int32_t* row_ptr, const int32_t* col,
                    const int32_t* in_row_ptr, int directed,
                    const int32_t* vs, int32_t nvs, int normalize, int endpoints,
                    double* bc) {
    for (int32_t i = 0; i < nv; i++) bc[i] = 0.0;
    for (int32_t i = 0; i < nvs; i++) if (vs[i] < 0 || vs[i] >= nv) return -2;
    dijk_ws w;
    if (dijk_ws_init(&w, nv, in_row_ptr[nv]) != 0) { dijk_ws_free(&w); return -1; }
    for (int32_t i = 0; i < nvs; i++) {
        int32_t s = vs[i];
        if (degree_of(s, row_ptr, in_row_ptr, directed) > 0) {    /* :58 */
            dijkstra_allpaths(&w, row_ptr, col, in_row_ptr, s);
            if (endpoints) accumulate_endpoints(bc, &w, in_row_ptr, s);
            else accumulate_basic(bc, &w, in_row_ptr, s);
        }
    }
    dijk_ws_free(&w);
    orc_rescale(bc, nv, normalize, directed, nvs);   /* k = length(vs) */
    return 0;
}

/* ------------------------------------------------------------------ */
/* threaded_betweenness_centrality -- Parallel/centrality/betweenness.jl:42-75
 * ------------------------------------------------------------------ */
/* vs_active = findall(x -> degree(g, x) > 0, vs) returns positions in vs
 * (:53); the loop then uses those positions as source vertices (:55-56).
 * The restatement keeps that: source = position (0-based) of every entry of
 * vs whose vertex has degree > 0.  Per-thread accumulators are reduced in
 * thread order (:63).  With `faithful_index_quirk` = 0 the sources are the
 * vertices themselves (the :distributed variant, :9-37, and the serial
 * function's semantics), which is what the CPU baseline times. */
typedef struct {
    int32_t nv; const int32_t* row_ptr; const int32_t* col; const int32_t* in_row_ptr;
    const int32_t* srcs; int64_t nsrc; int endpoints; double* acc;
    int64_t* next; pthread_mutex_t* mu; int rc;
} bc_job;

static void* bc_worker(void* arg) {
    bc_job* j = (bc_job*)arg;
    dijk_ws w;
    if (dijk_ws_init(&w, j->nv, j->in_row_ptr[j->nv]) != 0) { dijk_ws_free(&w); j->rc = -1; return NULL; }
    for (;;) {
        int64_t i;
        pthread_mutex_lock(j->mu);
        i = (*j->next)++;
        pthread_mutex_unlock(j->mu);
        if (i >= j->nsrc) break;
        int32_t s = j->srcs[i];
        dijkstra_allpaths(&w, j->row_ptr, j->col, j->in_row_ptr, s);
        if (j->endpoints) accumulate_endpoints(j->acc, &w, j->in_row_ptr, s);
        else accumulate_basic(j->acc, &w, j->in_row_ptr, s);
    }
    dijk_ws_free(&w);
    j->rc = 0;
    return NULL;
}

int orc_betweenness_threaded(int32_t nv, const int32_t* row_ptr, const int32_t* col,
                             const int32_t* in_row_ptr, int directed,
                             const int32_t* vs, int32_t nvs, int normalize, int endpoints,
                             int nthreads, int faithful_index_quirk, double* bc) {
    if (nthreads < 1) nthreads = 1;
    for (int32_t i = 0; i < nvs; i++) if (vs[i] < 0 || vs[i] >= nv) return -2;
    int32_t* act = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nvs + 1));
    double* accs = (double*)calloc((size_t)nthreads * (size_t)nv + 1, sizeof(double));
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    bc_job* jobs = (bc_job*)malloc(sizeof(bc_job) * (size_t)nthreads);
    if (!act || !accs || !th || !jobs) { free(act); free(accs); free(th); free(jobs); return -1; }
    int64_t nact = 0;
    for (int32_t i = 0; i < nvs; i++) {
        if (degree_of(vs[i], row_ptr, in_row_ptr, directed) > 0) {
            int32_t s = faithful_index_quirk ? i : vs[i];
            if (s >= nv) { free(act); free(accs); free(th); free(jobs); return -2; }
            act[nact++] = s;
        }
    }
    pthread_mutex_t mu;
    pthread_mutex_init(&mu, NULL);
    int64_t next = 0;
    for (int t = 0; t < nthreads; t++) {
        jobs[t].nv = nv; jobs[t].row_ptr = row_ptr; jobs[t].col = col; jobs[t].in_row_ptr = in_row_ptr;
        jobs[t].srcs = act; jobs[t].nsrc = nact; jobs[t].endpoints = endpoints;
        jobs[t].acc = accs + (size_t)t * (size_t)nv; jobs[t].next = &next; jobs[t].mu = &mu; jobs[t].rc = 0;
        pthread_create(&th[t], NULL, bc_worker, &jobs[t]);
    }
    int rc = 0;
    for (int t = 0; t < nthreads; t++) { pthread_join(th[t], NULL); if (jobs[t].rc) rc = jobs[t].rc; }
    pthread_mutex_destroy(&mu);
    for (int32_t v = 0; v < nv; v++) {               /* reduce(+, local_betweenness) */
        double s = 0.0;
        for (int t = 0; t < nthreads; t++) s += accs[(size_t)t * (size_t)nv + v];
        bc[v] = s;
    }
    free(act); free(accs); free(th); free(jobs);
    if (rc) return rc;
    orc_rescale(bc, nv, normalize, directed, nvs);
    return 0;
}

/* ------------------------------------------------------------------ */
/* stress_centrality -- src/centrality/stress.jl:33-77                  */
/* ------------------------------------------------------------------ */
/* _stress_accumulate_basic! (:53-77): for w in reverse(closest_vertices),
 * every predecessor v of w gets delta[v] += delta[w] + 1 (with delta[w]
 * not yet multiplied), then delta[w] *= length(P[w]) and stress[w] +=
 * delta[w] for w != s; P[s] is emptied first.  Int64 arithmetic wraps as
 * Julia's Int does (done in uint64_t here). */
static void stress_accumulate(int64_t* stress, dijk_ws* w, const int32_t* in_row_ptr, int32_t si) {
    int32_t nv = w->nv;
    uint64_t* delta = (uint64_t*)w->delta;      /* same 8-byte slots, integer use */
    for (int32_t i = 0; i < nv; i++) delta[i] = 0;
    for (int32_t k = w->nclosest - 1; k >= 0; k--) {
        int32_t x = w->closest[k];
        const int32_t* P = w->preds + in_row_ptr[x];
        int32_t np = (x == si) ? 0 : w->pred_cnt[x];
        for (int32_t j = 0; j < np; j++) delta[P[j]] += delta[x] + 1u;
        delta[x] *= (uint64_t)np;
        if (x != si) stress[x] = (int64_t)((uint64_t)stress[x] + delta[x]);
    }
}

/* stress_centrality(g, vs) -- stress.jl:33-48 (degree-0 sources skipped,
 * :41).  Parallel.stress_centrality (:distributed and :threads,
 * src/Parallel/centrality/stress.jl:9-51) sums the same per-source
 * vectors; integer addition makes the order irrelevant, so with nthreads
 * > 1 the sources are shared among threads and the per-thread sums added. */
typedef struct {
    int32_t nv; const int32_t* row_ptr; const int32_t* col; const int32_t* in_row_ptr;
    const int32_t* srcs; int64_t nsrc; int64_t* acc;
    int64_t* next; pthread_mutex_t* mu; int rc;
} sc_job;

static void* sc_worker(void* arg) {
    sc_job* j = (sc_job*)arg;
    dijk_ws w;
    if (dijk_ws_init(&w, j->nv, j->in_row_ptr[j->nv]) != 0) { dijk_ws_free(&w); j->rc = -1; return NULL; }
    for (;;) {
        int64_t i;
        pthread_mutex_lock(j->mu);
        i = (*j->next)++;
        pthread_mutex_unlock(j->mu);
        if (i >= j->nsrc) break;
        int32_t s = j->srcs[i];
        dijkstra_allpaths(&w, j->row_ptr, j->col, j->in_row_ptr, s);
        stress_accumulate(j->acc, &w, j->in_row_ptr, s);
    }
    dijk_ws_free(&w);
    j->rc = 0;
    return NULL;
}

int orc_stress(int32_t nv, const int32_t* row_ptr, const int32_t* col, const int32_t* in_row_ptr,
               int directed, const int32_t* vs, int32_t nvs, int nthreads, int64_t* stress) {
    if (nthreads < 1) nthreads = 1;
    for (int32_t i = 0; i < nvs; i++) if (vs[i] < 0 || vs[i] >= nv) return -2;
    int32_t* act = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nvs + 1));
    int64_t* accs = (int64_t*)calloc((size_t)nthreads * (size_t)nv + 1, sizeof(int64_t));
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    sc_job* jobs = (sc_job*)malloc(sizeof(sc_job) * (size_t)nthreads);
    if (!act || !accs || !th || !jobs) { free(act); free(accs); free(th); free(jobs); return -1; }
    int64_t nact = 0;
    for (int32_t i = 0; i < nvs; i++)
        if (degree_of(vs[i], row_ptr, in_row_ptr, directed) > 0) act[nact++] = vs[i];
    pthread_mutex_t mu;
    pthread_mutex_init(&mu, NULL);
    int64_t next = 0;
    for (int t = 0; t < nthreads; t++) {
        jobs[t].nv = nv; jobs[t].row_ptr = row_ptr; jobs[t].col = col; jobs[t].in_row_ptr = in_row_ptr;
        jobs[t].srcs = act; jobs[t].nsrc = nact; jobs[t].acc = accs + (size_t)t * (size_t)nv;
        jobs[t].next = &next; jobs[t].mu = &mu; jobs[t].rc = 0;
        pthread_create(&th[t], NULL, sc_worker, &jobs[t]);
    }
    int rc = 0;
    for (int t = 0; t < nthreads; t++) { pthread_join(th[t], NULL); if (jobs[t].rc) rc = jobs[t].rc; }
    pthread_mutex_destroy(&mu);
    for (int32_t v = 0; v < nv; v++) {
        uint64_t x = 0;
        for (int t = 0; t < nthreads; t++) x += (uint64_t)accs[(size_t)t * (size_t)nv + v];
        stress[v] = (int64_t)x;
    }
    free(act); free(accs); free(th); free(jobs);
    return rc;
}

/* Serial single-source BFS used as the CPU-baseline leg for
 * single-source BFS: distances and parents in one pass, restating
 * gdistances! (bfs.jl:95-124) with the parent write of _bfs_parents
 * (bfs.jl:54-58).  Returns the number of edges scanned (for TEPS). */
int64_t orc_bfs_dist_parent(int32_t nv, const int32_t* row_ptr, const int32_t* col,
                            int32_t src, int32_t* dist, int32_t* parents) {
    if (src < 0 || src >= nv) return -2;
    uint8_t* visited = (uint8_t*)calloc((size_t)nv, 1);
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nv + 1));
    int32_t* nxt = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nv + 1));
    if (!visited || !cur || !nxt) { free(visited); free(cur); free(nxt); return -1; }
    for (int32_t i = 0; i < nv; i++) { dist[i] = -1; parents[i] = -1; }
    int64_t ncur = 0, nnxt = 0, scanned = 0;
    visited[src] = 1; cur[ncur++] = src; parents[src] = src; dist[src] = 0;
    int32_t lvl = 1;
    while (ncur > 0) {
        for (int64_t k = 0; k < ncur; k++) {
            int32_t v = cur[k];
            scanned += row_ptr[v + 1] - row_ptr[v];
            for (int32_t e = row_ptr[v]; e < row_ptr[v + 1]; e++) {
                int32_t i = col[e];
                if (!visited[i]) { nxt[nnxt++] = i; parents[i] = v; dist[i] = lvl; visited[i] = 1; }
            }
        }
        lvl++;
        int32_t* t = cur; cur = nxt; nxt = t;
        ncur = nnxt; nnxt = 0;
        qsort(cur, (size_t)ncur, sizeof(int32_t), cmp_i32);
    }
    free(visited); free(cur); free(nxt);
    return scanned;
}
