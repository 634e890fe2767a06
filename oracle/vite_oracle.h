/*
 * CPU oracle: a single-threaded restatement of Vite's Louvain hot path at MPI world
 * size 1 (ECP-ExaGraph/vite).  TEST INFRASTRUCTURE ONLY: tests/, the smoke check and
 * the bench's CPU-baseline leg call it as the checker; the product never links it.
 *
 * Every function states the reference file:line it follows.  Floating-point sums are
 * left folds in the order one OpenMP thread performs them, with no FMA contraction
 * (the reference Makefile has no -march, Makefile:4,7), so the oracle reproduces the
 * reference run at OMP_NUM_THREADS=1 bit for bit.
 */
#ifndef VITE_ORACLE_H
#define VITE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Edge record, 16 bytes: edge.hpp:72-80 (GraphElem tail; GraphWeight weight). */
typedef struct {
    int64_t tail;
    double weight;
} or_edge;

/* CSR graph: graph.hpp:85-86 (edgeListIndexes, edgeList). */
typedef struct {
    int64_t nv, ne;
    int64_t* offs;   /* nv + 1 */
    or_edge* edges;  /* ne */
} or_graph;

void or_graph_free(or_graph* g);
int or_graph_copy(const or_graph* src, or_graph* dst);

/* Vite binary file (distgraph.cpp:149-150,181,220 reader; 1007-1045 writer). */
int or_read_bin(const char* path, or_graph* g);
int or_write_bin(const char* path, const or_graph* g);

/* ---- Louvain phase (louvain.hpp:80-137, louvain.cpp:7-1938) ---- */
enum { OR_MODE_PLAIN = 0, OR_MODE_COLOR = 1, OR_MODE_VORDER = 2 };

typedef struct {
    double lower;         /* `lower` argument                                     */
    double thresh;        /* `thresh` argument                                    */
    int et_type;          /* 0 none; 1..4 as main.cpp:350-365                     */
    double et_delta;      /* ETDelta for types 2/4                                */
    int mode;             /* OR_MODE_*                                            */
    int64_t num_color;    /* numColor as passed (numColors + 1, main.cpp:283)     */
    const int64_t* colors;/* vertexColor, -1 = uncolored                          */
} or_phase_params;

/* Optional per-iteration record. */
typedef struct {
    int max_iters;        /* capacity of the arrays below */
    int n;                /* iterations recorded          */
    double* q;            /* currMod per iteration (NaN when the ET count check breaks) */
    int64_t* moves;       /* vertices with targetComm != currComm per iteration          */
} or_trace;

double or_louvain_phase(const or_graph* g, const or_phase_params* p, int64_t* cvect_out,
                        int* iters_out, or_trace* trace);

/* distSumVertexDegree + distCalcConstantForSecondTerm (louvain.cpp:2161-2218).
 * Returns 1/(sum of degrees). */
double or_init_degrees(const or_graph* g, double* vdeg);

/* One Jacobi sweep of distExecuteLouvainIteration over `order` (louvain.cpp:2281-2466)
 * from the given state.  target/cw/dsize/ddeg are fully written (cw, dsize, ddeg start
 * from zero; vertices not in `order` keep target = curr). */
void or_sweep(const or_graph* g, const int64_t* curr, const int64_t* csize, const double* cdeg,
              const double* vdeg, double constant, const int64_t* order, int64_t norder,
              int64_t* target, double* cw, int64_t* dsize, double* ddeg);

/* distComputeModularity (louvain.cpp:2468-2532) with ABS_MOD_PER_ITER. */
double or_modularity(int64_t nv, const double* cw, const double* cdeg, double constant);

/* ---- Coarsening: distbuildNextLevelGraph (rebuild.cpp:77-504) ---- */
int or_rebuild(const or_graph* g, const int64_t* cvect, or_graph* out);

/* ---- Coloring: distColoringMultiHashMinMax (coloring.cpp:52-247) ---- */
int64_t or_color(const or_graph* g, int64_t nhash, int target_pct, int single_iter,
                 int64_t* colors_out);
uint32_t or_color_hash(uint32_t a, uint32_t seed);

/* ---- Whole run: main.cpp:242-513 ---- */
typedef struct {
    double threshold;        /* -h, default 1e-6 (utils.hpp:71-73)           */
    int threshold_cycling;   /* -i                                           */
    int et_type;             /* -t (0 = off)                                 */
    double et_delta;         /* -a, default 1.0                              */
    int coloring;            /* -c                                           */
    int vertex_order;        /* -d                                           */
    int max_colors;          /* first -c / -d argument, default 8            */
    int single_color_iter;   /* second -c / -d argument                      */
    int one_phase;           /* -p                                           */
} or_run_params;

#define OR_MAX_LEVELS 256
typedef struct {
    int nlevels;                         /* Level lines printed (main.cpp:468)           */
    double level_q[OR_MAX_LEVELS];
    int level_iters[OR_MAX_LEVELS];      /* cumulative, as printed                      */
    int64_t level_nv[OR_MAX_LEVELS];     /* vertices of the graph clustered at the level */
    int64_t level_ne[OR_MAX_LEVELS];
    int ncalls;                          /* Louvain calls made                          */
    double call_q[OR_MAX_LEVELS + 1];
    int call_iters[OR_MAX_LEVELS + 1];
    uint64_t call_cvect_fnv[OR_MAX_LEVELS + 1]; /* FNV-1a of the partition each call returned */
    double final_q;                      /* currMod at loop exit (MODS / time)          */
    int total_iters;
    int64_t num_colors;                  /* coloring result when -c/-d                  */
} or_run_result;

int or_run(const or_graph* g, const or_run_params* rp, or_run_result* res);

/* ---- Generators ---- */
uint64_t or_rng_u64(uint64_t seed, uint64_t stream, uint64_t k);
double or_rgg_radius(int64_t n);
/* Seeded random geometric graph (cell-binned), optional extra uniform edges (percent of
 * the undirected edge count, weight U(0,1)). */
int or_gen_rgg(uint64_t seed, int64_t n, double extra_pct, or_graph* out);
/* Seeded Barabasi-Albert graph (Batagelj-Brandes), m attachments, unit weights. */
int or_gen_ba(uint64_t seed, int64_t n, int64_t m, or_graph* out);
/* Vite's own in-memory RGG at one rank (distgraph.cpp:391-983, utils.hpp:135-330). */
int or_gen_vite_rgg(int64_t n, or_graph* out);
uint32_t or_seed_seq_1(uint32_t s);

/* ---- Ground-truth comparison ---- */
typedef struct {
    int64_t n, nc_truth, nc_found;
    int64_t same_same, same_diff, diff_same;
    double precision, recall, f_score, gini_truth, gini_found;
} or_compare_stats;
/* compare_communities (compare.cpp:58-307) and compute_gini_coeff (compare.cpp:310-337)
 * at one thread: per-community member lists, then every member pair. -1 on bad input. */
int or_compare(int64_t n, const int64_t* truth, const int64_t* found, or_compare_stats* out);

/* FNV-1a of an int64 vector (for compact fixtures). */
uint64_t or_fnv_i64(const int64_t* v, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
