/*
 * vite_amd — C ABI of the MI355X (gfx950) Louvain community-detection path.
 *
 * Every entry point returns an int status (VT_OK = 0, negative on error), never
 * throws, and leaves a message retrievable with vt_last_error().  Host pointers are
 * plain C arrays; no C++ or framework types cross this boundary.  One context owns one
 * HIP stream on one device; calls on a context are synchronous at return and must not
 * overlap from several host threads.
 *
 * Reference correspondence (ECP-ExaGraph/vite, file:line):
 *   vt_graph_upload        DistGraph/Graph CSR (graph.hpp:85-86, distgraph.hpp:77-107);
 *                          Edge = {GraphElem tail; GraphWeight weight} (edge.hpp:72-80)
 *   vt_louvain_phase       distLouvainMethod / ...WithColoring / ...VertexOrder, all nine
 *                          overloads (louvain.hpp:80-137; louvain.cpp:7-1938)
 *   vt_sweep               distExecuteLouvainIteration over all vertices from a given state
 *                          (louvain.cpp:2281-2466), for replay parity
 *   vt_modularity          distComputeModularity (louvain.cpp:2468-2532)
 *   vt_rebuild             distbuildNextLevelGraph (rebuild.hpp:109-112; rebuild.cpp:480-504)
 *   vt_color               distColoringMultiHashMinMax (coloring.hpp:86; coloring.cpp:52-118)
 *   vt_gen_vite_rgg        generateInMemGraph at one rank (distgraph.hpp:116; distgraph.cpp:391-983)
 *   vt_compare_communities compare_communities / compute_gini_coeff (compare.hpp; compare.cpp:58-337)
 *   vt_gen_rgg, vt_gen_ba  the workload generators (no reference counterpart: Vite's RGG uses
 *                          another radius and an O(n^2) pair test; its BA needs NetworKit)
 *   vt_run                 the phase loop of main.cpp:242-513
 *   vt_batch_*             the same loop over many graphs at once (no reference counterpart)
 *   vt_server_*            a request loop running such jobs concurrently (no reference counterpart)
 */
#ifndef VITE_AMD_H
#define VITE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VT_OK 0
#define VT_ERR_ARG (-1)      /* invalid argument or malformed CSR            */
#define VT_ERR_HIP (-2)      /* HIP runtime error                            */
#define VT_ERR_BUDGET (-3)   /* allocation would exceed the context budget   */
#define VT_ERR_LIMIT (-4)    /* size beyond what this build supports         */
#define VT_ERR_STATE (-5)    /* call out of order (e.g. no partition yet)    */

typedef struct vt_ctx vt_ctx;
typedef struct vt_graph vt_graph;

/* 16-byte edge record, layout-identical to the reference Edge (edge.hpp:72-80). */
typedef struct {
    int64_t tail;
    double weight;
} vt_edge;

/* ---- context ---------------------------------------------------------------- */
/* hbm_budget_bytes = 0 means no limit beyond the device. */
int vt_ctx_create(int device, uint64_t hbm_budget_bytes, vt_ctx** out);
int vt_ctx_destroy(vt_ctx* ctx);
const char* vt_last_error(const vt_ctx* ctx);
/* Bytes of live device allocations of this context (graphs + working state); freed blocks
 * the context keeps for reuse are not counted, and are returned to the driver when the
 * budget or the device runs short and at vt_ctx_destroy. */
uint64_t vt_ctx_bytes_in_use(const vt_ctx* ctx);
/* Per-kernel timing with HIP events on the context stream (0 = off). */
int vt_ctx_set_profiling(vt_ctx* ctx, int on);
/* Average duration (ms) and launch count of kernel `name` since profiling was enabled. */
int vt_ctx_kernel_stats(const vt_ctx* ctx, const char* name, double* avg_ms, int64_t* count,
                        double* total_ms);
int vt_ctx_reset_stats(vt_ctx* ctx);

/* Launch shape of the local move.  Rows of degree <= bucket_max_degree[b] (and above the
 * previous bound) run with 4, 8, 16, 32 or 64 lanes per vertex (b = 0..4; each bound at most
 * 2 x lanes - 1) or one 64-lane wave with a 2048-slot table (b = 5, at most 1023); longer rows
 * use tables in HBM.  blocks_per_bucket caps the workgroups one bucket gets per launch.
 * Any valid setting gives bit-identical results; only the speed changes.
 * Default {7, 15, 47, 63, 191, 511}, 65536, {255, 511}, 0, 1024. */
typedef struct {
    int32_t bucket_max_degree[6];
    int32_t blocks_per_bucket;
    /* Table size of the wide bucket (one wave per row, up to 1023 entries): rows up to
     * [0] get 512 LDS slots, up to [1] 1024, the others 2048.  A table stays at most 3/4
     * full, so [0] <= 383 and [1] <= 767 (-1: none); [0] <= [1]. */
    int32_t wide_table_max_degree[2];
    /* 1: the group kernels (rows up to bucket 4) read each neighbour's community, and each
     * candidate community's size and degree, from one 16-byte record per vertex id, rebuilt
     * before every sweep; a candidate community whose id is the neighbour's own id then comes
     * with the neighbour's line.  Pays on graphs whose neighbours are scattered (random
     * reads); 0 (the default): separate arrays. */
    int32_t vertex_records;
    /* Rows above bucket_max_degree[5] (hubs) are grouped by neighbour community with a sort
     * per sweep.  A sweep with at least this many hubs gives each hub a workgroup, which
     * sorts the hub's row (a segmented sort) and moves the hub, unless its largest hub holds
     * more than 1/64 of the hub entries; otherwise it sorts all hub entries at once and runs
     * every community run in a flat launch (a few large hubs would leave most of the chip
     * idle).  0 <= hub_block_min <= 2^30; 0 gives every sweep the per-hub form. */
    int32_t hub_block_min;
} vt_tuning;
int vt_ctx_get_tuning(const vt_ctx* ctx, vt_tuning* out);
int vt_ctx_set_tuning(vt_ctx* ctx, const vt_tuning* t);
/* Autotuner: from the state the first `prep` iterations of a phase-0 run (threshold 1e-3)
 * reach on g, times `reps` local-move sweeps under candidate bucket bounds, block caps and
 * wide-row table sizes (coordinate descent from the current setting), keeps the fastest on
 * ctx and returns it. */
int vt_autotune(vt_ctx* ctx, vt_graph* g, int32_t prep, int32_t reps, vt_tuning* out);

/* ---- graphs ------------------------------------------------------------------- */
/* Copies a host CSR (offsets[nv+1], edges[ne]) to the device.  Rejects offsets that are
 * not monotone from 0 to ne and tails outside [0, nv). */
int vt_graph_upload(vt_ctx* ctx, int64_t nv, int64_t ne, const int64_t* offsets,
                    const vt_edge* edges, vt_graph** out);
int vt_graph_info(const vt_graph* g, int64_t* nv, int64_t* ne);
/* Copies the device CSR back; offsets[nv+1], edges[ne] are caller-allocated. */
int vt_graph_download(vt_ctx* ctx, const vt_graph* g, int64_t* offsets, vt_edge* edges);
int vt_graph_free(vt_ctx* ctx, vt_graph* g);

/* ---- Louvain phase -------------------------------------------------------------- */
#define VT_MODE_PLAIN 0   /* distLouvainMethod                     (louvain.cpp:7-591)     */
#define VT_MODE_COLOR 1   /* distLouvainMethodWithColoring         (louvain.cpp:593-1263)  */
#define VT_MODE_VORDER 2  /* distLouvainMethodVertexOrder          (louvain.cpp:1270-1938) */

/* Bit flags. */
#define VT_FLAG_SEQUENTIAL_SUMS 1  /* 2m and modularity as left folds over vertex order,
                                      as one OpenMP thread of the reference computes them
                                      (slow; for exact whole-run replay on small graphs) */
#define VT_FLAG_NO_GRAPH 2         /* vt_batch_run: launch each iteration's kernels directly
                                      instead of replaying captured HIP graphs            */

typedef struct {
    double lower;          /* `lower` argument                                          */
    double thresh;         /* `thresh` argument                                         */
    int32_t et_type;       /* 0 none; 1 ET(local), 2 ET-prob(local), 3 ET(count check),
                              4 ET-prob(count check) -- main.cpp:350-365                */
    int32_t mode;          /* VT_MODE_*                                                 */
    double et_delta;       /* ETDelta for types 2 and 4                                 */
    int64_t num_color;     /* numColor as the reference receives it (numColors + 1)     */
    const int64_t* colors; /* vertexColor (host, nv entries, -1 = uncolored) or NULL    */
    int32_t flags;         /* VT_FLAG_*                                                 */
    int32_t max_iters;     /* safety cap on iterations, 0 = none                        */
} vt_phase_params;

/* Runs one phase.  cvect_out (host, nv entries) may be NULL: the partition then stays on
 * the device only, attached to `g` for vt_rebuild.  q_out receives the return value of
 * the reference function (prevMod), iters_out its `iters`. */
int vt_louvain_phase(vt_ctx* ctx, vt_graph* g, const vt_phase_params* p, int64_t* cvect_out,
                     int32_t* iters_out, double* q_out);

/* The same phase, one iteration at a time, with its state resident on the device between
 * calls (the locals of distLouvainMethod*, louvain.cpp:14-50).  vt_phase_step runs one
 * iteration of the reference loop body and reports whether the reference would stop;
 * vt_phase_sweep runs only the local move over every vertex from the current state without
 * applying the moves (repeatable: it leaves the state unchanged); vt_phase_end returns what
 * vt_louvain_phase returns and frees the state. */
typedef struct vt_phase vt_phase;
int vt_phase_begin(vt_ctx* ctx, vt_graph* g, const vt_phase_params* p, vt_phase** out);
int vt_phase_step(vt_ctx* ctx, vt_phase* st, double* q_out, int32_t* done_out);
int vt_phase_sweep(vt_ctx* ctx, vt_phase* st);
int vt_phase_end(vt_ctx* ctx, vt_phase* st, int64_t* cvect_out, int32_t* iters_out, double* q_out);
/* Copies the current state to host arrays (nv entries each; any may be NULL): currComm,
 * localCinfo sizes and degrees, vertex degrees, 1/2m, and the last sweep's targetComm and
 * clusterWeight. */
int vt_phase_get_state(vt_ctx* ctx, const vt_phase* st, int64_t* curr, int64_t* csize, double* cdeg, double* vdeg,
                       double* constant, int64_t* target, double* cw);

/* Per-iteration record of the last vt_louvain_phase call on this context. */
int vt_last_phase_trace(const vt_ctx* ctx, int32_t cap, double* q, int64_t* moves, int32_t* n);

/* One local-move sweep (every vertex, Jacobi) from a host-given state: currComm, community
 * sizes and degrees (localCinfo), vertex degrees and 1/2m.  Writes targetComm and
 * clusterWeight; dsize/ddeg (may be NULL) receive localCupdate folded per community in
 * vertex order. */
int vt_sweep(vt_ctx* ctx, const vt_graph* g, const int64_t* curr, const int64_t* csize,
             const double* cdeg, const double* vdeg, double constant, int64_t* target_out,
             double* cw_out, int64_t* dsize_out, double* ddeg_out);

/* distComputeModularity on host arrays (fixed-order device tree, or left fold with
 * VT_FLAG_SEQUENTIAL_SUMS). */
int vt_modularity(vt_ctx* ctx, int64_t nv, const double* cw, const double* cdeg, double constant,
                  int32_t flags, double* q_out);

/* vt_modularity on DEVICE arrays, allocating nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (serving loops, harnesses that forbid
 * hipMalloc).  d_cw and d_cdeg hold nv doubles each and d_q_out one double, all device
 * memory of ctx's device; the result has the same bits as vt_modularity's on the same
 * values and flags.  The partial sums live in a fixed 16.25 KiB device array per device
 * (any nv fits: the tree has at most 1024 blocks), resolved when a context is created on
 * that device and held until the process ends.  Every call on one device shares that
 * array, so calls that overlap on one device (from several contexts or threads) race.
 * Synchronous at return like every other entry point. */
int vt_modularity_noalloc(vt_ctx* ctx, int64_t nv, const double* d_cw, const double* d_cdeg, double constant,
                          int32_t flags, double* d_q_out);

/* Vertex degrees (row left folds) and 1/2m. */
int vt_init_degrees(vt_ctx* ctx, const vt_graph* g, int32_t flags, double* vdeg_out,
                    double* constant_out);

/* ---- coarsening ----------------------------------------------------------------- */
/* Builds the next-level graph from a partition: cvect is a host array (nv entries) or NULL
 * for the partition the last vt_louvain_phase left on `g`. */
int vt_rebuild(vt_ctx* ctx, const vt_graph* g, const int64_t* cvect, vt_graph** out);

/* ---- coloring ----------------------------------------------------------------------- */
int vt_color(vt_ctx* ctx, const vt_graph* g, int64_t nhash, int32_t target_pct,
             int32_t single_iter, int64_t* colors_out, int64_t* next_color_out);

/* ---- ground-truth comparison ------------------------------------------------------ */
/* compare_communities (compare.cpp:58-307) with compute_gini_coeff (compare.cpp:310-337):
 * pair counts of a computed partition against a ground truth, both host arrays of n
 * community ids >= 0 (ids need not be dense: a partition has max id + 1 communities, empty
 * ones included, as in the reference).  The counts are exact; precision, recall, F-score
 * and the Gini coefficients use the reference's expressions and its one-thread order. */
typedef struct {
    int64_t n;
    int64_t nc_truth, nc_found;  /* max id + 1 of each partition                           */
    int64_t same_same;           /* pairs together in both (true positives)                */
    int64_t same_diff;           /* together in the truth, apart in found (false negatives) */
    int64_t diff_same;           /* apart in the truth, together in found (false positives) */
    double precision, recall, f_score;
    double gini_truth, gini_found;
} vt_compare_stats;
int vt_compare_communities(vt_ctx* ctx, int64_t n, const int64_t* truth, const int64_t* found,
                           vt_compare_stats* out);

/* ---- generators (device) --------------------------------------------------------- */
/* Seeded random geometric graph in [0,1]^2 with radius sqrt(2 ln n/(pi n)) and extra_pct
 * percent extra uniform-random edges of weight U(0,1). */
int vt_gen_rgg(vt_ctx* ctx, uint64_t seed, int64_t n, double extra_pct, vt_graph** out);
/* Seeded Barabasi-Albert graph, m attachments per vertex, unit weights. */
int vt_gen_ba(vt_ctx* ctx, uint64_t seed, int64_t n, int64_t m, vt_graph** out);

/* Vite's own in-memory RGG at one rank (generateInMemGraph, distgraph.cpp:391-983): Lehmer
 * LCG x_{k+1} = 16807 x_k mod (2^31 - 1) from x0 (utils.hpp:135-330), X = first n values,
 * Y = next n, radius (sqrt(ln n/(3.14159 n)) + sqrt(2.0736/n))/2, rows sorted by tail.
 * The reference draws x0 as the first word of std::seed_seq{1}. */
int vt_gen_vite_rgg(vt_ctx* ctx, int64_t n, uint32_t x0, vt_graph** out);

/* ---- device primitives of the batched loop ---------------------------------------- */
/* No reference counterpart: the scan and the stable radix sort the batched loop runs inside
 * its captured HIP graph, where the element count is only known on the device.  Exposed so
 * they can be tested on their own.  Host arrays in and out; the count n is handed to the
 * kernels through device memory, and the grids are sized for `capacity` (>= n).
 * vt_dev_sort_pairs: stable sort of n (key, value) pairs by key bits [0, bits), 1 <= bits <= 32.
 * vt_dev_scan: exclusive prefix sums of n int32 values; total_out (may be NULL) gets the sum. */
int vt_dev_sort_pairs(vt_ctx* ctx, int64_t n, int64_t capacity, int32_t bits, const uint32_t* keys,
                      const uint32_t* vals, uint32_t* keys_out, uint32_t* vals_out);
int vt_dev_scan(vt_ctx* ctx, int64_t n, int64_t capacity, const int32_t* in, int32_t* out, int64_t* total_out);

/* ---- whole run (main.cpp:242-513) -------------------------------------------------- */
typedef struct {
    double threshold;           /* -h (default 1e-6)          */
    int32_t threshold_cycling;  /* -i                          */
    int32_t et_type;            /* -t                          */
    double et_delta;            /* -a (default 1.0)            */
    int32_t coloring;           /* -c                          */
    int32_t vertex_order;       /* -d                          */
    int32_t max_colors;         /* -c/-d first argument (8)    */
    int32_t single_color_iter;  /* -c/-d second argument       */
    int32_t one_phase;          /* -p                          */
    int32_t flags;              /* VT_FLAG_*                   */
} vt_run_params;

#define VT_MAX_LEVELS 256
typedef struct {
    int32_t nlevels;
    double level_q[VT_MAX_LEVELS];
    int32_t level_iters[VT_MAX_LEVELS];
    int64_t level_nv[VT_MAX_LEVELS];
    int64_t level_ne[VT_MAX_LEVELS];
    int32_t ncalls;
    double call_q[VT_MAX_LEVELS + 1];
    int32_t call_iters[VT_MAX_LEVELS + 1];
    double final_q;
    int32_t total_iters;
    int64_t num_colors;
    double clustering_seconds;  /* wall time of the Louvain calls                  */
    double total_seconds;       /* wall time of the whole loop incl. rebuilds      */
} vt_run_result;

/* Runs the phase loop on `g` (not modified). */
int vt_run(vt_ctx* ctx, const vt_graph* g, const vt_run_params* p, vt_run_result* res);

/* ---- batched form (no reference counterpart: Vite runs one graph per job) -------------- */
/* Many independent graphs run the phase loop together as one device-resident disjoint
 * union: each Louvain iteration of every graph runs in the same launches (replayed from
 * captured HIP graphs), each graph leaving its phase on its own test, and all members are
 * coarsened together between phases.  Per graph, every call's return value and iteration
 * count and the final modularity equal vt_run on that graph alone.  Plain Louvain only
 * (threshold, threshold cycling, one phase); coloring and early termination are refused. */
typedef struct vt_batch vt_batch;
typedef struct {
    int32_t phases;   /* batched phases run (rounds of the phase loop)             */
    int32_t rounds;   /* batched iterations over all phases                        */
    double seconds;   /* wall time of the whole batched run                        */
    int32_t relists;  /* times the sweep lists dropped the vertices of finished graphs */
} vt_batch_stats;
/* Copies n device graphs into a batch; the graphs stay the caller's.  Each needs at least one
 * edge (with 2m = 0 modularity is undefined and the reference's loop never ends). */
int vt_batch_create(vt_ctx* ctx, int32_t n, vt_graph* const* graphs, vt_batch** out);
int vt_batch_info(const vt_batch* b, int32_t* n, int64_t* total_nv, int64_t* total_ne);
/* results: n entries, in the order the graphs were given.  stats may be NULL. */
int vt_batch_run(vt_ctx* ctx, vt_batch* b, const vt_run_params* p, vt_run_result* results, vt_batch_stats* stats);
/* vt_autotune on the batch's level-0 union, its vertices swept as one graph: picks the
 * bucket bounds and block cap for the batch's degree mix and keeps them on ctx. */
int vt_batch_autotune(vt_ctx* ctx, vt_batch* b, int32_t prep, int32_t reps, vt_tuning* out);
int vt_batch_free(vt_ctx* ctx, vt_batch* b);

/* ---- serving form (no reference counterpart: Vite runs one job per process) ------------ */
/* A request loop over several slots (worker threads, each with its own context and HIP
 * stream, so jobs of different slots overlap on the device).  A request generates a seeded
 * graph on the device and runs the phase loop on it.  Requests are admitted in arrival
 * order when a slot is free and their HBM reservation (vt_request_reservation) fits beside
 * those of the running jobs; a job cannot allocate beyond its reservation.  A result does
 * not depend on what else runs: every job runs on its own stream, and a batched job gives
 * each graph what it gets alone. */
#define VT_REQ_RGG 0  /* seeded RGG (vt_gen_rgg): param = extra edge percent         */
#define VT_REQ_BA 1   /* seeded Barabasi-Albert (vt_gen_ba): param = attachments m    */
typedef struct {
    int32_t kind;
    uint64_t seed;
    int64_t n;
    double param;
    vt_run_params run;
} vt_request;
typedef struct {
    int32_t status;           /* VT_OK or the failing call's code                          */
    int32_t slot;             /* slot that ran it                                          */
    double submit_s;          /* server clock, seconds since vt_server_create              */
    double start_s;           /* admitted to a slot                                        */
    double end_s;             /* finished                                                  */
    uint64_t reserved_bytes;  /* HBM reserved at admission                                 */
    uint64_t peak_bytes;      /* HBM high-water mark of the job                            */
    int64_t nv, ne;           /* size of the generated graph                               */
    int32_t batch;            /* requests run together in its job (1: alone)               */
} vt_request_info;
typedef struct vt_server vt_server;
/* Bytes the server reserves for a request (an upper bound on what its job allocates). */
uint64_t vt_request_reservation(const vt_request* req);
int vt_server_create(int device, uint64_t hbm_budget_bytes, int32_t slots, vt_server** out);
/* Small-request batching: a slot that admits a request of at most max_n vertices (plain
 * Louvain: no coloring, vertex order or early termination) also takes up to max_batch - 1
 * such requests with the same run parameters queued right behind it, as far as their
 * reservations fit, and runs them as one batched job (vt_batch_run); each result equals the
 * request's vt_run alone.  The batched job reserves twice its members' reservations (their
 * graphs and the union).  Default: max_batch 32, max_n 65536; max_batch 1 turns it off. */
int vt_server_set_batching(vt_server* s, int32_t max_batch, int64_t max_n);
/* Queues a request; VT_ERR_BUDGET if its reservation exceeds the whole budget.  Submit and
 * wait may be called from several host threads at once. */
int vt_server_submit(vt_server* s, const vt_request* req, int64_t* ticket);
/* Blocks until the request has finished; returns its status.  Each ticket is waited once:
 * a second wait on a ticket already being waited on returns VT_ERR_STATE, and one on a
 * ticket already collected VT_ERR_ARG. */
int vt_server_wait(vt_server* s, int64_t ticket, vt_run_result* res, vt_request_info* info);
/* Message of the calling thread's last failed vt_server_* call. */
const char* vt_server_last_error(const vt_server* s);
/* Finishes every queued request, then stops the slots. */
int vt_server_destroy(vt_server* s);

#ifdef __cplusplus
}
#endif
#endif
