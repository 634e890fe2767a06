/*
 * CPU oracle for Vite's Louvain hot path (world size 1).  TEST INFRASTRUCTURE ONLY:
 * tests/, the smoke check and the bench's CPU-baseline leg use it as the checker.
 *
 * Single-threaded C restatement.  Floating-point reductions are left folds in the
 * order one OpenMP thread of the reference performs them; build with
 * -ffp-contract=off so no FMA is formed (the reference build forms none).
 */
#include "vite_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------------ */
/* Graph utilities                                                                 */
/* ------------------------------------------------------------------------------ */

void or_graph_free(or_graph* g)
{
    if (!g) return;
    free(g->offs);
    free(g->edges);
    g->offs = NULL;
    g->edges = NULL;
    g->nv = g->ne = 0;
}

static int graph_alloc(or_graph* g, int64_t nv, int64_t ne)
{
    g->nv = nv;
    g->ne = ne;
    g->offs = (int64_t*)calloc((size_t)nv + 1, sizeof(int64_t));
    g->edges = (or_edge*)malloc(sizeof(or_edge) * (size_t)(ne > 0 ? ne : 1));
    return (g->offs && g->edges) ? 0 : -1;
}

int or_graph_copy(const or_graph* src, or_graph* dst)
{
    if (graph_alloc(dst, src->nv, src->ne)) return -1;
    memcpy(dst->offs, src->offs, sizeof(int64_t) * (size_t)(src->nv + 1));
    if (src->ne) memcpy(dst->edges, src->edges, sizeof(or_edge) * (size_t)src->ne);
    return 0;
}

/* File layout: nv:i64, ne:i64, offsets[nv+1]:i64, edges[ne]:{i64 tail, f64 weight}
 * (distgraph.cpp:149-150,181,220; writer distgraph.cpp:1007-1045). */
int or_read_bin(const char* path, or_graph* g)
{
    FILE* f = fopen(path, "rb");
    if (!f) return -1;
    int64_t hdr[2];
    if (fread(hdr, sizeof(int64_t), 2, f) != 2) { fclose(f); return -2; }
    if (hdr[0] < 0 || hdr[1] < 0) { fclose(f); return -3; }
    if (graph_alloc(g, hdr[0], hdr[1])) { fclose(f); return -4; }
    if (fread(g->offs, sizeof(int64_t), (size_t)hdr[0] + 1, f) != (size_t)hdr[0] + 1) { fclose(f); return -5; }
    if (hdr[1] && fread(g->edges, sizeof(or_edge), (size_t)hdr[1], f) != (size_t)hdr[1]) { fclose(f); return -6; }
    fclose(f);
    /* loadDistGraphMPIIO rebases offsets to start at 0 (distgraph.cpp:244-247). */
    const int64_t b = g->offs[0];
    for (int64_t i = 0; i <= g->nv; i++) g->offs[i] -= b;
    return 0;
}

int or_write_bin(const char* path, const or_graph* g)
{
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    int64_t hdr[2] = {g->nv, g->ne};
    int ok = fwrite(hdr, sizeof(int64_t), 2, f) == 2;
    ok = ok && fwrite(g->offs, sizeof(int64_t), (size_t)g->nv + 1, f) == (size_t)g->nv + 1;
    if (g->ne) ok = ok && fwrite(g->edges, sizeof(or_edge), (size_t)g->ne, f) == (size_t)g->ne;
    fclose(f);
    return ok ? 0 : -2;
}

uint64_t or_fnv_i64(const int64_t* v, int64_t n)
{
    uint64_t h = 1469598103934665603ull;
    for (int64_t i = 0; i < n; i++) {
        uint64_t u = (uint64_t)v[i];
        for (int b = 0; b < 8; b++) { h ^= (u >> (8 * b)) & 0xff; h *= 1099511628211ull; }
    }
    return h;
}

/* ------------------------------------------------------------------------------ */
/* Per-vertex community map: ClusterLocalMap + counter vector (louvain.cpp:2297-2298) */
/* ------------------------------------------------------------------------------ */

typedef struct {
    int64_t* hkeys;    /* open-addressing keys, -1 = empty      */
    int32_t* hslot;    /* slot index into comm/cnt               */
    uint32_t hcap;
    int64_t* comm;     /* slot -> community (insertion order)    */
    double* cnt;       /* slot -> counter                        */
    int32_t nslots;
} cmap;

static void cmap_init(cmap* m, int64_t maxdeg)
{
    uint32_t cap = 16;
    while (cap < 2 * (uint64_t)(maxdeg + 1)) cap <<= 1;
    m->hcap = cap;
    m->hkeys = (int64_t*)malloc(sizeof(int64_t) * cap);
    m->hslot = (int32_t*)malloc(sizeof(int32_t) * cap);
    m->comm = (int64_t*)malloc(sizeof(int64_t) * (size_t)(maxdeg + 2));
    m->cnt = (double*)malloc(sizeof(double) * (size_t)(maxdeg + 2));
    for (uint32_t i = 0; i < cap; i++) m->hkeys[i] = -1;
    m->nslots = 0;
}

static void cmap_free(cmap* m)
{
    free(m->hkeys); free(m->hslot); free(m->comm); free(m->cnt);
}

static inline uint32_t cmap_h(int64_t k, uint32_t cap)
{
    uint64_t z = (uint64_t)k * 0x9E3779B97F4A7C15ull;
    return (uint32_t)(z >> 32) & (cap - 1);
}

static void cmap_clear(cmap* m)
{
    for (int32_t s = 0; s < m->nslots; s++) {
        uint32_t h = cmap_h(m->comm[s], m->hcap);
        while (m->hkeys[h] != -1) { m->hkeys[h] = -1; h = (h + 1) & (m->hcap - 1); }
    }
    m->nslots = 0;
}

/* Returns the slot of `c`, or -1 when absent. */
static inline int32_t cmap_find(const cmap* m, int64_t c)
{
    uint32_t h = cmap_h(c, m->hcap);
    while (m->hkeys[h] != -1) {
        if (m->hkeys[h] == c) return m->hslot[h];
        h = (h + 1) & (m->hcap - 1);
    }
    return -1;
}

static inline void cmap_insert(cmap* m, int64_t c, double v)
{
    uint32_t h = cmap_h(c, m->hcap);
    while (m->hkeys[h] != -1) h = (h + 1) & (m->hcap - 1);
    m->hkeys[h] = c;
    m->hslot[h] = m->nslots;
    m->comm[m->nslots] = c;
    m->cnt[m->nslots] = v;
    m->nslots++;
}

static int64_t max_degree(const or_graph* g)
{
    int64_t md = 0;
    for (int64_t i = 0; i < g->nv; i++) {
        int64_t d = g->offs[i + 1] - g->offs[i];
        if (d > md) md = d;
    }
    return md;
}

/* ------------------------------------------------------------------------------ */
/* Louvain core                                                                    */
/* ------------------------------------------------------------------------------ */

/* distSumVertexDegree (louvain.cpp:2161-2186) and distCalcConstantForSecondTerm
 * (louvain.cpp:2188-2218): k_i is the left fold of row weights from 0.0; 2m is the left
 * fold of k_i over i (one thread's OpenMP reduction), the constant is 1/2m. */
double or_init_degrees(const or_graph* g, double* vdeg)
{
    for (int64_t i = 0; i < g->nv; i++) {
        double tw = 0.0;
        for (int64_t k = g->offs[i]; k < g->offs[i + 1]; k++) tw += g->edges[k].weight;
        vdeg[i] = tw;
    }
    double local = 0.0;
    for (int64_t i = 0; i < g->nv; i++) local += vdeg[i];
    double total = 0.0 + local;
    return 1.0 / total;
}

/* distGetMaxIndex (louvain.cpp:2220-2279).  Candidates are every map entry except the
 * vertex's own community; gain = 2*(e_iy - e_ix) - 2*k_i*(a_y - a_x)*constant evaluated
 * left to right; a candidate replaces the current best when its gain is larger, or equal,
 * non-zero and of smaller id.  Then the singleton rule (louvain.cpp:2275-2276). */
static int64_t get_max_index(const cmap* m, double selfLoop, const int64_t* csize,
                             const double* cdeg, double vDegree, int64_t currSize,
                             double currDegree, int64_t currComm, double constant)
{
    int64_t maxIndex = currComm;
    double curGain = 0.0, maxGain = 0.0;
    const double eix = m->cnt[0] - selfLoop;
    const double ax = currDegree - vDegree;
    int64_t maxSize = currSize;
    for (int32_t s = 0; s < m->nslots; s++) {
        const int64_t c = m->comm[s];
        if (c == currComm) continue;
        const double ay = cdeg[c];
        const int64_t size = csize[c];
        const double eiy = m->cnt[s];
        curGain = 2.0 * (eiy - eix) - 2.0 * vDegree * (ay - ax) * constant;
        if ((curGain > maxGain) || ((curGain == maxGain) && (curGain != 0.0) && (c < maxIndex))) {
            maxGain = curGain;
            maxIndex = c;
            maxSize = size;
        }
    }
    if ((maxSize == 1) && (currSize == 1) && (maxIndex > currComm)) maxIndex = currComm;
    return maxIndex;
}

/* distExecuteLouvainIteration (louvain.cpp:2281-2417) with distBuildLocalMapCounter
 * (louvain.cpp:2419-2466) inlined.  Deltas go to ddeg/dsize in call order (one thread's
 * sequence of `omp atomic` updates, louvain.cpp:2351-2358). */
static void exec_vertex(const or_graph* g, int64_t i, const int64_t* curr, int64_t* target,
                        const double* vdeg, const int64_t* csize, const double* cdeg,
                        int64_t* dsize, double* ddeg, double constant, double* cw, cmap* m)
{
    const int64_t cc = curr[i];
    const double ccDegree = cdeg[cc];
    const int64_t ccSize = csize[cc];
    const int64_t e0 = g->offs[i], e1 = g->offs[i + 1];
    int64_t localTarget;
    if (e0 != e1) {
        cmap_clear(m);
        cmap_insert(m, cc, 0.0);
        double selfLoop = 0.0;
        for (int64_t j = e0; j < e1; j++) {
            const int64_t tail = g->edges[j].tail;
            const double weight = g->edges[j].weight;
            if (tail == i) selfLoop += weight;
            const int64_t tcomm = curr[tail];
            const int32_t s = cmap_find(m, tcomm);
            if (s >= 0) m->cnt[s] += weight;
            else cmap_insert(m, tcomm, weight);
        }
        cw[i] += m->cnt[0];
        localTarget = get_max_index(m, selfLoop, csize, cdeg, vdeg[i], ccSize, ccDegree, cc, constant);
    } else {
        localTarget = cc;
    }
    if (localTarget != cc && localTarget != -1) {
        ddeg[localTarget] += vdeg[i];
        dsize[localTarget]++;
        ddeg[cc] -= vdeg[i];
        dsize[cc]--;
    }
    target[i] = localTarget;
}

void or_sweep(const or_graph* g, const int64_t* curr, const int64_t* csize, const double* cdeg,
              const double* vdeg, double constant, const int64_t* order, int64_t norder,
              int64_t* target, double* cw, int64_t* dsize, double* ddeg)
{
    cmap m;
    cmap_init(&m, max_degree(g));
    for (int64_t i = 0; i < g->nv; i++) { target[i] = curr[i]; cw[i] = 0.0; dsize[i] = 0; ddeg[i] = 0.0; }
    for (int64_t k = 0; k < norder; k++)
        exec_vertex(g, order ? order[k] : k, curr, target, vdeg, csize, cdeg, dsize, ddeg, constant, cw, &m);
    cmap_free(&m);
}

/* distComputeModularity (louvain.cpp:2468-2532): one thread's reductions are left folds
 * over i; ABS_MOD_PER_ITER takes the absolute value (louvain.cpp:2518-2524). */
double or_modularity(int64_t nv, const double* cw, const double* cdeg, double constant)
{
    double le_xx = 0.0, la2_x = 0.0;
    for (int64_t i = 0; i < nv; i++) {
        le_xx += cw[i];
        la2_x += cdeg[i] * cdeg[i];
    }
    const double e_xx = 0.0 + le_xx, a2_x = 0.0 + la2_x;
    return fabs((e_xx * constant) - (a2_x * constant * constant));
}

/* localCinfo += localCupdate over all communities (louvain.cpp:2534-2547). */
static void apply_deltas(int64_t nv, int64_t* csize, double* cdeg, int64_t* dsize, double* ddeg)
{
    for (int64_t i = 0; i < nv; i++) {
        csize[i] += dsize[i];
        cdeg[i] += ddeg[i];
        dsize[i] = 0;
        ddeg[i] = 0.0;
    }
}

/* The nine overloads (louvain.cpp:7-1938) as one routine:
 *   mode PLAIN  : plain (427-591), ET bool (7-205), ET prob (207-425)
 *   mode COLOR  : per-color delta application (593-1263); currComm is not refreshed
 *                 between colors
 *   mode VORDER : color order, one delta application per sweep (1270-1938)
 * numColor == 1 under COLOR/VORDER falls back to plain without ET (louvain.cpp:601-605). */
double or_louvain_phase(const or_graph* g, const or_phase_params* p, int64_t* cvect_out,
                        int* iters_out, or_trace* trace)
{
    int mode = p->mode, et = p->et_type;
    if (mode != OR_MODE_PLAIN && p->num_color == 1) { mode = OR_MODE_PLAIN; et = 0; }
    const int64_t nv = g->nv;
    const size_t n8 = (size_t)(nv > 0 ? nv : 1);

    int64_t* past = (int64_t*)malloc(n8 * sizeof(int64_t));
    int64_t* curr = (int64_t*)malloc(n8 * sizeof(int64_t));
    int64_t* target = (int64_t*)calloc(n8, sizeof(int64_t));   /* resize(nv): zeros */
    double* vdeg = (double*)malloc(n8 * sizeof(double));
    double* cw = (double*)calloc(n8, sizeof(double));
    double* frozen = (double*)calloc(n8, sizeof(double));
    int64_t* csize = (int64_t*)malloc(n8 * sizeof(int64_t));
    double* cdeg = (double*)malloc(n8 * sizeof(double));
    int64_t* dsize = (int64_t*)calloc(n8, sizeof(int64_t));
    double* ddeg = (double*)calloc(n8, sizeof(double));
    unsigned char* active = (unsigned char*)malloc(n8);
    double* p_curr = (double*)calloc(n8, sizeof(double));      /* value-initialised */
    double* p_prev = (double*)malloc(n8 * sizeof(double));
    int64_t* colorPtr = NULL;
    int64_t* colorIndex = NULL;

    /* distInitLouvain (louvain.cpp:2138-2159), distInitComm (louvain.cpp:2564-2581). */
    const double constant = or_init_degrees(g, vdeg);
    for (int64_t i = 0; i < nv; i++) {
        csize[i] = 1;
        cdeg[i] = vdeg[i];
        past[i] = curr[i] = i;
        active[i] = 1;
        p_prev[i] = 1.0;
    }

    /* Color CSR (louvain.cpp:647-677); one thread fills each color in vertex order. */
    const int64_t numColor = p->num_color;
    if (mode != OR_MODE_PLAIN) {
        colorPtr = (int64_t*)calloc((size_t)numColor + 1, sizeof(int64_t));
        int64_t* added = (int64_t*)calloc((size_t)numColor, sizeof(int64_t));
        colorIndex = (int64_t*)malloc(n8 * sizeof(int64_t));
        for (int64_t i = 0; i < nv; i++) {
            const int64_t s = (p->colors[i] < 0) ? (numColor - 1) : p->colors[i];
            colorPtr[s + 1]++;
        }
        for (int64_t c = 0; c < numColor; c++) colorPtr[c + 1] += colorPtr[c];
        for (int64_t i = 0; i < nv; i++) {
            const int64_t s = (p->colors[i] < 0) ? (numColor - 1) : p->colors[i];
            colorIndex[colorPtr[s] + added[s]++] = i;
        }
        free(added);
    }

    cmap m;
    cmap_init(&m, max_degree(g));

    double prevMod = p->lower, currMod = -1.0;
    int numIters = 0;
    const double one_minus_delta = 1.0 - p->et_delta;
    if (trace) trace->n = 0;

    for (;;) {
        numIters++;
        long vc_count = 0;
        int64_t moves = 0;
        for (int64_t i = 0; i < nv; i++) cw[i] = 0.0;

#define VISIT(v)                                                                           \
        do {                                                                               \
            const int64_t v_ = (v);                                                        \
            if (et && !active[v_]) { cw[v_] = frozen[v_]; vc_count++; }                    \
            else {                                                                         \
                exec_vertex(g, v_, curr, target, vdeg, csize, cdeg, dsize, ddeg, constant, cw, &m); \
                if (et) frozen[v_] = cw[v_];                                               \
                if (target[v_] != curr[v_]) moves++;                                       \
            }                                                                              \
        } while (0)

        if (mode == OR_MODE_PLAIN) {
            for (int64_t i = 0; i < nv; i++) VISIT(i);
        } else if (mode == OR_MODE_COLOR) {
            for (int64_t ci = 0; ci < numColor; ci++) {
                for (int64_t K = colorPtr[ci]; K < colorPtr[ci + 1]; K++) VISIT(colorIndex[K]);
                apply_deltas(nv, csize, cdeg, dsize, ddeg);
            }
        } else {
            for (int64_t ci = 0; ci < numColor; ci++)
                for (int64_t K = colorPtr[ci]; K < colorPtr[ci + 1]; K++) VISIT(colorIndex[K]);
        }
#undef VISIT

        /* ET types 3/4: the inactive *count* is compared with ET_CUTOFF (louvain.cpp:116-121). */
        if ((et == 3 || et == 4) && vc_count >= 90) {
            currMod = -1.0;
            if (trace && trace->n < trace->max_iters) { trace->q[trace->n] = NAN; trace->moves[trace->n] = moves; trace->n++; }
            break;
        }
        if (mode != OR_MODE_COLOR) apply_deltas(nv, csize, cdeg, dsize, ddeg);

        currMod = or_modularity(nv, cw, cdeg, constant);
        if (trace && trace->n < trace->max_iters) { trace->q[trace->n] = currMod; trace->moves[trace->n] = moves; trace->n++; }

        if ((currMod - prevMod) < p->thresh) break;
        prevMod = currMod;
        if (prevMod < p->lower) prevMod = p->lower;

        if (et == 0) {
            for (int64_t i = 0; i < nv; i++) {
                const int64_t tmp = past[i]; past[i] = curr[i]; curr[i] = target[i]; target[i] = tmp;
            }
        } else if (et == 1 || et == 3) {
            /* `targetComm[i] == currComm[i] == pastComm[i]` parses as
             * ((t == c) == p): a bool compared with a community id (louvain.cpp:174-175). */
            for (int64_t i = 0; i < nv; i++) {
                if (numIters > 2 && ((int64_t)(target[i] == curr[i]) == past[i])) active[i] = 0;
                else { const int64_t tmp = past[i]; past[i] = curr[i]; curr[i] = target[i]; target[i] = tmp; }
            }
        } else {
            /* Probabilistic freeze (louvain.cpp:380-400); whole vectors swap afterwards. */
            for (int64_t i = 0; i < nv; i++) {
                if (!active[i]) continue;
                if (numIters > 2 && curr[i] == past[i]) {
                    p_curr[i] = p_prev[i] * one_minus_delta;
                    if (p_curr[i] <= 0.02) active[i] = 0;
                }
                if (active[i]) { const int64_t tmp = past[i]; past[i] = curr[i]; curr[i] = target[i]; target[i] = tmp; }
            }
            double* t = p_curr; p_curr = p_prev; p_prev = t;
        }
    }

    for (int64_t i = 0; i < nv; i++) cvect_out[i] = past[i];
    *iters_out = numIters;

    cmap_free(&m);
    free(past); free(curr); free(target); free(vdeg); free(cw); free(frozen);
    free(csize); free(cdeg); free(dsize); free(ddeg); free(active); free(p_curr); free(p_prev);
    free(colorPtr); free(colorIndex);
    return prevMod;
}

/* ------------------------------------------------------------------------------ */
/* Coarsening                                                                      */
/* ------------------------------------------------------------------------------ */

typedef struct { int64_t c1, c2, seq; double w; } ctrip;

static int ctrip_cmp(const void* a, const void* b)
{
    const ctrip* x = (const ctrip*)a;
    const ctrip* y = (const ctrip*)b;
    if (x->c1 != y->c1) return x->c1 < y->c1 ? -1 : 1;
    if (x->c2 != y->c2) return x->c2 < y->c2 ? -1 : 1;
    return (x->seq < y->seq) ? -1 : (x->seq > y->seq);
}

/* distReNumber (rebuild.cpp:77-292): new id = rank of first appearance of cvect[i] as i
 * ascends (rebuild.cpp:121-135).  fill_newEdgesMap (rebuild.cpp:294-329): the weight of
 * coarse edge (C1,C2) is the left fold of fine weights over i in C1 ascending and each
 * row in CSR order, i.e. global edge order; the first contribution is stored as is.
 * send_newEdges (rebuild.cpp:331-478): rows sorted by column (std::map). */
int or_rebuild(const or_graph* g, const int64_t* cvect, or_graph* out)
{
    const int64_t nv = g->nv, ne = g->ne;
    int64_t* lookup = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nv > 0 ? nv : 1));
    for (int64_t i = 0; i < nv; i++) lookup[i] = -1;
    int64_t lnc = 0;
    for (int64_t i = 0; i < nv; i++) {
        const int64_t c = cvect[i];
        if (c < 0 || c >= nv) { free(lookup); return -1; }
        if (lookup[c] == -1) lookup[c] = lnc++;
    }
    ctrip* t = (ctrip*)malloc(sizeof(ctrip) * (size_t)(ne > 0 ? ne : 1));
    for (int64_t i = 0; i < nv; i++)
        for (int64_t k = g->offs[i]; k < g->offs[i + 1]; k++) {
            t[k].c1 = lookup[cvect[i]];
            t[k].c2 = lookup[cvect[g->edges[k].tail]];
            t[k].seq = k;
            t[k].w = g->edges[k].weight;
        }
    qsort(t, (size_t)ne, sizeof(ctrip), ctrip_cmp);
    int64_t nu = 0;
    for (int64_t k = 0; k < ne; k++)
        if (k == 0 || t[k].c1 != t[k - 1].c1 || t[k].c2 != t[k - 1].c2) nu++;
    if (graph_alloc(out, lnc, nu)) { free(t); free(lookup); return -2; }
    int64_t e = -1;
    for (int64_t k = 0; k < ne; k++) {
        if (k == 0 || t[k].c1 != t[k - 1].c1 || t[k].c2 != t[k - 1].c2) {
            e++;
            out->edges[e].tail = t[k].c2;
            out->edges[e].weight = t[k].w;
            out->offs[t[k].c1 + 1]++;
        } else {
            out->edges[e].weight += t[k].w;
        }
    }
    for (int64_t c = 0; c < lnc; c++) out->offs[c + 1] += out->offs[c];
    free(t);
    free(lookup);
    return 0;
}

/* ------------------------------------------------------------------------------ */
/* Coloring                                                                        */
/* ------------------------------------------------------------------------------ */

/* 32-bit integer hash with a seed xor (coloring.cpp:120-131): the six-step
 * add/xor/shift mix published by Robert Jenkins. */
uint32_t or_color_hash(uint32_t a, uint32_t seed)
{
    a ^= seed;
    a = (a + 0x7ed55d16u) + (a << 12);
    a = (a ^ 0xc761c23cu) + (a >> 19);
    a = (a + 0x165667b1u) + (a << 5);
    a = (a ^ 0xd3a2646cu) + (a << 9);
    a = (a + 0xfd7046c5u) + (a << 3);
    a = (a ^ 0xb55a4f09u) + (a >> 16);
    return a;
}

/* distColoringIteration (coloring.cpp:133-247) at one rank.  A vertex colored earlier in
 * the same round carries a color >= nextColor, so the skip test reads round-start state. */
static void color_iteration(const or_graph* g, int64_t* color, int64_t nHash, int64_t nextColor,
                            uint32_t seed, uint32_t* vHash)
{
    for (int64_t v = 0; v < g->nv; v++) {
        int64_t possible_colors = 2 * nHash;
        if (color[v] != -1) continue;
        int not_min = 0, not_max = 0;
        int can_color = 1;
        for (int64_t t = 0; t < nHash; t++) vHash[t] = or_color_hash((uint32_t)v, seed + 1043u * (uint32_t)t);
        for (int64_t k = g->offs[v]; k < g->offs[v + 1]; k++) {
            const int64_t tail = g->edges[k].tail;
            if (v == tail) continue;
            if (color[tail] != -1 && color[tail] < nextColor) continue;
            for (int64_t t = 0; t < nHash; t++) {
                const uint32_t jHash = or_color_hash((uint32_t)tail, seed + 1043u * (uint32_t)t);
                if (vHash[t] <= jHash && !(not_max & (0x1 << t))) { not_max |= (0x1 << t); possible_colors--; }
                if (vHash[t] >= jHash && !(not_min & (0x1 << t))) { not_min |= (0x1 << t); possible_colors--; }
            }
            if (possible_colors == 0) { can_color = 0; break; }
        }
        if (can_color) {
            const int64_t col_id = v % possible_colors;
            int64_t this_col_id = 0;
            for (int64_t t = 0; t < nHash; t++) {
                if (!(not_min & (0x1 << t)) && col_id == this_col_id) { color[v] = 2 * t + nextColor; break; }
                this_col_id += !(not_min & (0x1 << t));
                if (!(not_max & (0x1 << t)) && col_id == this_col_id) { color[v] = 2 * t + 1 + nextColor; break; }
                this_col_id += !(not_max & (0x1 << t));
            }
        }
    }
}

/* distColoringMultiHashMinMax (coloring.cpp:52-118). */
int64_t or_color(const or_graph* g, int64_t nHash, int target_percent, int single_iter,
                 int64_t* color)
{
    uint32_t seed = 1012;
    int64_t nextColor = 0, lastCount = 0, currentCount = 0;
    const int64_t tnv = g->nv;
    const int64_t targetCount = (tnv * target_percent) / 100;
    uint32_t* vHash = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(nHash > 0 ? nHash : 1));
    for (int64_t i = 0; i < tnv; i++) color[i] = -1;
    int finished = 0;
    while (!finished) {
        color_iteration(g, color, nHash, nextColor, seed, vHash);
        int64_t unassigned = 0;
        for (int64_t i = 0; i < tnv; i++) if (color[i] == -1) unassigned++;
        currentCount = tnv - unassigned;
        finished = single_iter || currentCount >= targetCount || lastCount == currentCount;
        nextColor += nHash * 2;
        seed = or_color_hash(seed, 0);
        lastCount = currentCount;
    }
    free(vHash);
    return nextColor;
}

/* ------------------------------------------------------------------------------ */
/* Phase loop (main.cpp:242-513)                                                   */
/* ------------------------------------------------------------------------------ */

int or_run(const or_graph* g0, const or_run_params* rp, or_run_result* res)
{
    memset(res, 0, sizeof(*res));
    or_graph dg;
    if (or_graph_copy(g0, &dg)) return -1;
    double currMod = -1.0, prevMod = -1.0, threshold = rp->threshold;
    int phase = 0, short_phase = 0, iters = 0, tot_iters = 0;
    int64_t* cvect = NULL;
    int64_t* colors = NULL;

    for (;;) {
        if (rp->threshold_cycling && !rp->one_phase) {
            if (short_phase > 12) short_phase = 0;
            if (short_phase >= 0 && short_phase <= 2) threshold = 1.0E-3;
            else if (short_phase >= 3 && short_phase <= 6) threshold = 1.0E-4;
            else if (short_phase >= 7 && short_phase <= 9) threshold = 1.0E-5;
            else threshold = rp->threshold;
        } else {
            threshold = rp->threshold;
        }

        free(cvect);
        cvect = (int64_t*)malloc(sizeof(int64_t) * (size_t)(dg.nv > 0 ? dg.nv : 1));
        or_phase_params pp;
        memset(&pp, 0, sizeof(pp));
        pp.lower = currMod;
        pp.thresh = threshold;
        pp.et_type = rp->et_type;
        pp.et_delta = rp->et_delta;
        if ((rp->coloring || rp->vertex_order) && phase == 0) {
            colors = (int64_t*)malloc(sizeof(int64_t) * (size_t)(dg.nv > 0 ? dg.nv : 1));
            const int64_t numColors = or_color(&dg, rp->max_colors / 2, 70, rp->single_color_iter, colors);
            res->num_colors = numColors;
            pp.mode = rp->coloring ? OR_MODE_COLOR : OR_MODE_VORDER;
            pp.num_color = numColors + 1;
            pp.colors = colors;
        }
        res->level_nv[res->nlevels < OR_MAX_LEVELS ? res->nlevels : OR_MAX_LEVELS - 1] = dg.nv;
        res->level_ne[res->nlevels < OR_MAX_LEVELS ? res->nlevels : OR_MAX_LEVELS - 1] = dg.ne;
        currMod = or_louvain_phase(&dg, &pp, cvect, &iters, NULL);
        if (res->ncalls <= OR_MAX_LEVELS) {
            res->call_q[res->ncalls] = currMod;
            res->call_iters[res->ncalls] = iters;
            res->call_cvect_fnv[res->ncalls] = or_fnv_i64(cvect, dg.nv);
            res->ncalls++;
        }
        free(colors);
        colors = NULL;
        tot_iters += iters;

        if ((currMod - prevMod) > threshold) {
            if (!rp->one_phase) {
                or_graph next;
                if (or_rebuild(&dg, cvect, &next)) { or_graph_free(&dg); free(cvect); return -2; }
                or_graph_free(&dg);
                dg = next;
            }
        } else {
            if (rp->threshold_cycling && !rp->one_phase && phase < 10) {
                memset(&pp, 0, sizeof(pp));
                pp.lower = currMod;
                pp.thresh = 1.0E-6;
                currMod = or_louvain_phase(&dg, &pp, cvect, &iters, NULL);
                if (res->ncalls <= OR_MAX_LEVELS) {
                    res->call_q[res->ncalls] = currMod;
                    res->call_iters[res->ncalls] = iters;
                    res->call_cvect_fnv[res->ncalls] = or_fnv_i64(cvect, dg.nv);
                    res->ncalls++;
                }
            }
            break;
        }
        if (res->nlevels < OR_MAX_LEVELS) {
            res->level_q[res->nlevels] = currMod;
            res->level_iters[res->nlevels] = tot_iters;
            res->nlevels++;
        }
        prevMod = currMod;
        currMod = -1.0;
        phase++;
        if (rp->one_phase) break;
        if (rp->threshold_cycling && !rp->one_phase) short_phase++;
        if (phase == 200 || tot_iters > 10000) break;
    }
    res->final_q = currMod;
    res->total_iters = tot_iters;
    free(cvect);
    or_graph_free(&dg);
    return 0;
}

/* ------------------------------------------------------------------------------ */
/* Generators                                                                      */
/* ------------------------------------------------------------------------------ */

static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

#define GOLDEN 0x9e3779b97f4a7c15ull

/* Counter-based stream: SplitMix64 output function applied to a keyed counter. */
uint64_t or_rng_u64(uint64_t seed, uint64_t stream, uint64_t k)
{
    const uint64_t key = mix64(seed + GOLDEN * (stream + 1));
    return mix64(key + GOLDEN * (k + 1));
}

static inline double u01(uint64_t z) { return (double)(z >> 11) * 0x1.0p-53; }

/* Radius for a mean degree of about 2 ln n: r = sqrt(2 ln n / (pi n)). */
double or_rgg_radius(int64_t n)
{
    if (n < 2) return 0.0;
    return sqrt(2.0 * log((double)n) / (M_PI * (double)n));
}

typedef struct { int64_t a, b; double w; int64_t t; } xedge;

static int has_edge(const or_graph* g, int64_t a, int64_t b)
{
    int64_t lo = g->offs[a], hi = g->offs[a + 1];
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (g->edges[mid].tail < b) lo = mid + 1; else hi = mid;
    }
    return lo < g->offs[a + 1] && g->edges[lo].tail == b;
}

typedef struct { int64_t key; int64_t k; } pkey;
static int pkey_cmp(const void* p, const void* q)
{
    const pkey* x = (const pkey*)p;
    const pkey* y = (const pkey*)q;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return (x->k < y->k) ? -1 : (x->k > y->k);
}

/* Merge undirected extra edges (a < b, weight w) into a CSR whose rows are sorted. */
static int merge_extra(or_graph* g, const xedge* xe, int64_t nx)
{
    const int64_t nv = g->nv;
    int64_t* xdeg = (int64_t*)calloc((size_t)nv + 1, sizeof(int64_t));
    for (int64_t k = 0; k < nx; k++) { xdeg[xe[k].a + 1]++; xdeg[xe[k].b + 1]++; }
    for (int64_t i = 0; i < nv; i++) xdeg[i + 1] += xdeg[i];
    or_edge* xr = (or_edge*)malloc(sizeof(or_edge) * (size_t)(2 * nx + 1));
    int64_t* fill = (int64_t*)calloc((size_t)nv, sizeof(int64_t));
    for (int64_t k = 0; k < nx; k++) {
        xr[xdeg[xe[k].a] + fill[xe[k].a]++] = (or_edge){xe[k].b, xe[k].w};
        xr[xdeg[xe[k].b] + fill[xe[k].b]++] = (or_edge){xe[k].a, xe[k].w};
    }
    /* sort each extra row by tail (insertion sort: rows are short) */
    for (int64_t i = 0; i < nv; i++)
        for (int64_t x = xdeg[i] + 1; x < xdeg[i + 1]; x++) {
            or_edge e = xr[x];
            int64_t y = x - 1;
            while (y >= xdeg[i] && xr[y].tail > e.tail) { xr[y + 1] = xr[y]; y--; }
            xr[y + 1] = e;
        }
    or_graph m;
    if (graph_alloc(&m, nv, g->ne + 2 * nx)) return -1;
    int64_t o = 0;
    for (int64_t i = 0; i < nv; i++) {
        m.offs[i] = o;
        int64_t p = g->offs[i], q = xdeg[i];
        while (p < g->offs[i + 1] || q < xdeg[i + 1]) {
            if (q >= xdeg[i + 1] || (p < g->offs[i + 1] && g->edges[p].tail < xr[q].tail)) m.edges[o++] = g->edges[p++];
            else m.edges[o++] = xr[q++];
        }
    }
    m.offs[nv] = o;
    free(xdeg); free(xr); free(fill);
    or_graph_free(g);
    *g = m;
    return 0;
}

/* Seeded RGG in [0,1]^2 with radius or_rgg_radius(n).
 *  - point k: x = u01(rng(seed,1,2k)), y = u01(rng(seed,1,2k+1));
 *  - cells of side 1/floor(1/r) >= r; vertex ids follow (cell id row-major, k), so the
 *    3x3 cell neighbourhood enumerates each row's tails in ascending order;
 *  - edge (u,v) iff sqrt(dx*dx + dy*dy) <= r (no FMA), weight = that distance;
 *  - extra edges: nx = (int64)(pct * m) / 100 (the rounding of distgraph.cpp:686);
 *    candidate t is (rng(seed,2,2t) % n, rng(seed,2,2t+1) % n) with weight
 *    u01(rng(seed,3,t)); self pairs, RGG edges and repeats of an earlier candidate are
 *    skipped; the first nx survivors in t order are kept. */
int or_gen_rgg(uint64_t seed, int64_t n, double extra_pct, or_graph* out)
{
    if (n < 1) return -1;
    const double r = or_rgg_radius(n);
    int64_t nc = (r > 0.0) ? (int64_t)floor(1.0 / r) : 1;
    if (nc < 1) nc = 1;
    if (nc > 32768) nc = 32768;
    double* px = (double*)malloc(sizeof(double) * (size_t)n);
    double* py = (double*)malloc(sizeof(double) * (size_t)n);
    pkey* pk = (pkey*)malloc(sizeof(pkey) * (size_t)n);
    for (int64_t k = 0; k < n; k++) {
        const double x = u01(or_rng_u64(seed, 1, 2 * (uint64_t)k));
        const double y = u01(or_rng_u64(seed, 1, 2 * (uint64_t)k + 1));
        int64_t cx = (int64_t)(x * (double)nc); if (cx >= nc) cx = nc - 1;
        int64_t cy = (int64_t)(y * (double)nc); if (cy >= nc) cy = nc - 1;
        pk[k].key = cy * nc + cx;
        pk[k].k = k;
        px[k] = x;
        py[k] = y;
    }
    qsort(pk, (size_t)n, sizeof(pkey), pkey_cmp);
    double* X = (double*)malloc(sizeof(double) * (size_t)n);
    double* Y = (double*)malloc(sizeof(double) * (size_t)n);
    const int64_t ncell = nc * nc;
    int64_t* cstart = (int64_t*)calloc((size_t)ncell + 1, sizeof(int64_t));
    for (int64_t v = 0; v < n; v++) {
        X[v] = px[pk[v].k];
        Y[v] = py[pk[v].k];
        cstart[pk[v].key + 1]++;
    }
    for (int64_t c = 0; c < ncell; c++) cstart[c + 1] += cstart[c];
    int64_t* cellof = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    for (int64_t v = 0; v < n; v++) cellof[v] = pk[v].key;
    free(px); free(py); free(pk);

    /* two passes: count, then fill */
    int64_t* offs = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    or_edge* edges = NULL;
    for (int pass = 0; pass < 2; pass++) {
        for (int64_t v = 0; v < n; v++) {
            const int64_t cy = cellof[v] / nc, cx = cellof[v] % nc;
            int64_t o = pass ? offs[v] : 0;
            for (int64_t yy = cy - 1; yy <= cy + 1; yy++) {
                if (yy < 0 || yy >= nc) continue;
                for (int64_t xx = cx - 1; xx <= cx + 1; xx++) {
                    if (xx < 0 || xx >= nc) continue;
                    const int64_t c = yy * nc + xx;
                    for (int64_t u = cstart[c]; u < cstart[c + 1]; u++) {
                        if (u == v) continue;
                        const double dx = X[v] - X[u];
                        const double dy = Y[v] - Y[u];
                        const double d = sqrt(dx * dx + dy * dy);
                        if (d <= r) {
                            if (pass) edges[o] = (or_edge){u, d};
                            o++;
                        }
                    }
                }
            }
            if (!pass) offs[v + 1] = o;
        }
        if (!pass) {
            for (int64_t v = 0; v < n; v++) offs[v + 1] += offs[v];
            edges = (or_edge*)malloc(sizeof(or_edge) * (size_t)(offs[n] > 0 ? offs[n] : 1));
        }
    }
    out->nv = n;
    out->ne = offs[n];
    out->offs = offs;
    out->edges = edges;
    free(X); free(Y); free(cstart); free(cellof);

    if (extra_pct > 0.0 && n > 1) {
        const int64_t m = out->ne / 2;
        const int64_t nx = ((int64_t)(extra_pct * (double)m)) / 100;
        xedge* acc = (xedge*)malloc(sizeof(xedge) * (size_t)(nx > 0 ? nx : 1));
        /* accepted pairs, open addressing (a = -1: empty slot) */
        int64_t cap = 2;
        while (cap < 2 * nx + 2) cap <<= 1;
        int64_t* ha = (int64_t*)malloc(sizeof(int64_t) * (size_t)cap);
        int64_t* hb = (int64_t*)malloc(sizeof(int64_t) * (size_t)cap);
        for (int64_t k = 0; k < cap; k++) ha[k] = -1;
        int64_t na = 0;
        uint64_t t = 0;
        while (na < nx) {
            const int64_t a0 = (int64_t)(or_rng_u64(seed, 2, 2 * t) % (uint64_t)n);
            const int64_t b0 = (int64_t)(or_rng_u64(seed, 2, 2 * t + 1) % (uint64_t)n);
            const double w = u01(or_rng_u64(seed, 3, t));
            const uint64_t tt = t++;
            if (a0 == b0) continue;
            const int64_t a = a0 < b0 ? a0 : b0, b = a0 < b0 ? b0 : a0;
            if (has_edge(out, a, b)) continue;
            uint64_t h = ((uint64_t)a * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)b * 0xC2B2AE3D27D4EB4Full);
            int64_t slot = (int64_t)((h ^ (h >> 29)) & (uint64_t)(cap - 1));
            int dup = 0;
            while (ha[slot] != -1) {
                if (ha[slot] == a && hb[slot] == b) { dup = 1; break; }
                slot = (slot + 1) & (cap - 1);
            }
            if (dup) continue;
            ha[slot] = a;
            hb[slot] = b;
            acc[na++] = (xedge){a, b, w, (int64_t)tt};
        }
        free(ha); free(hb);
        int rc = merge_extra(out, acc, na);
        free(acc);
        if (rc) return rc;
    }
    return 0;
}

typedef struct { int64_t a, b; } upair;
static int upair_cmp(const void* p, const void* q)
{
    const upair* x = (const upair*)p;
    const upair* y = (const upair*)q;
    if (x->a != y->a) return x->a < y->a ? -1 : 1;
    return (x->b < y->b) ? -1 : (x->b > y->b);
}

/* Barabasi-Albert by the Batagelj-Brandes edge-copy rule: edge e = v*m + j has source
 * v; its target is M[r] with r = rng(seed,4,e) % (2e+1) over the array
 * M[2e] = e/m, M[2e+1] = target(e).  Self-loops and repeated pairs are dropped; the
 * graph is undirected with unit weights. */
int or_gen_ba(uint64_t seed, int64_t n, int64_t m, or_graph* out)
{
    if (n < 1 || m < 1) return -1;
    const int64_t ne = n * m;
    int64_t* M = (int64_t*)malloc(sizeof(int64_t) * (size_t)(2 * ne));
    upair* pr = (upair*)malloc(sizeof(upair) * (size_t)(2 * ne));
    int64_t np = 0;
    for (int64_t e = 0; e < ne; e++) {
        M[2 * e] = e / m;
        const uint64_t r = or_rng_u64(seed, 4, (uint64_t)e) % (uint64_t)(2 * e + 1);
        M[2 * e + 1] = M[r];
        const int64_t a = M[2 * e], b = M[2 * e + 1];
        if (a == b) continue;
        pr[np++] = (upair){a, b};
        pr[np++] = (upair){b, a};
    }
    free(M);
    qsort(pr, (size_t)np, sizeof(upair), upair_cmp);
    int64_t nu = 0;
    for (int64_t k = 0; k < np; k++)
        if (k == 0 || pr[k].a != pr[k - 1].a || pr[k].b != pr[k - 1].b) pr[nu++] = pr[k];
    if (graph_alloc(out, n, nu)) { free(pr); return -2; }
    for (int64_t k = 0; k < nu; k++) {
        out->offs[pr[k].a + 1]++;
        out->edges[k] = (or_edge){pr[k].b, 1.0};
    }
    for (int64_t i = 0; i < n; i++) out->offs[i + 1] += out->offs[i];
    free(pr);
    return 0;
}

/* std::seed_seq::generate for one seed value and one output word ([rand.util.seedseq]),
 * as used by LCG::reseeder (utils.hpp:163-170). */
uint32_t or_seed_seq_1(uint32_t s0)
{
    const uint32_t n = 1, s = 1;
    uint32_t b[1] = {0x8b8b8b8bu};
    const uint32_t t = (n >= 623) ? 11 : (n >= 68) ? 7 : (n >= 39) ? 5 : (n >= 7) ? 3 : (n - 1) / 2;
    const uint32_t p = (n - t) / 2, q = p + t;
    const uint32_t mm = (s + 1 > n) ? s + 1 : n;
#define T_(x) ((x) ^ ((x) >> 27))
    for (uint32_t k = 0; k < mm; k++) {
        const uint32_t km1 = (k + n - 1) % n;
        uint32_t r1 = 1664525u * T_(b[k % n] ^ b[(k + p) % n] ^ b[km1]);
        uint32_t r2 = r1;
        if (k == 0) r2 += s;
        else if (k <= s) r2 += k % n + s0;
        else r2 += k % n;
        b[(k + p) % n] += r1;
        b[(k + q) % n] += r2;
        b[k % n] = r2;
    }
    for (uint32_t k = mm; k < mm + n; k++) {
        const uint32_t km1 = (k + n - 1) % n;
        uint32_t r3 = 1566083941u * T_(b[k % n] + b[(k + p) % n] + b[km1]);
        uint32_t r4 = r3 - k % n;
        b[(k + p) % n] ^= r3;
        b[(k + q) % n] ^= r4;
        b[k % n] = r4;
    }
#undef T_
    return b[0];
}

/* generateInMemGraph / generateRGG at one rank (distgraph.cpp:391-983):
 *   x0 = seed_seq{1}; x_{k+1} = 16807 x_k mod (2^31-1) (utils.hpp:254-295);
 *   drand = |x| * 1/(1 + (2^31-2)) = |x|/(2^31-1); X[i] = drand[i]; Y[j] = 0 + 1.0*drand[n+j] (utils.hpp:299-321);
 *   rn = (sqrt(ln n/(3.14159 n)) + sqrt(2.0736/n))/2 (distgraph.cpp:397-399);
 *   pair i<j is an edge iff sqrt(dx*dx+dy*dy) <= rn, weight the distance; rows sorted. */
int or_gen_vite_rgg(int64_t n, or_graph* out)
{
    if (n < 1) return -1;
    const double PI_V = 3.14159;
    const double rc = sqrt((double)log((double)n) / (double)(PI_V * (double)n));
    const double rt = sqrt((double)2.0736 / (double)n);
    const double rn = (rc + rt) / (double)2.0;
    int64_t* rnums = (int64_t*)malloc(sizeof(int64_t) * (size_t)(2 * n));
    double* drand = (double*)malloc(sizeof(double) * (size_t)(2 * n));
    rnums[0] = (int64_t)or_seed_seq_1(1);
    for (int64_t i = 1; i < 2 * n; i++) rnums[i] = (rnums[i - 1] * 16807 + 0) % 2147483647;
    const double mult = 1.0 / (double)(1.0 + (double)(2147483647 - 1));
    for (int64_t i = 0; i < 2 * n; i++) drand[i] = (double)((double)fabs((double)rnums[i]) * mult);
    double* X = drand;
    double* Y = (double*)malloc(sizeof(double) * (size_t)n);
    const double range = 1.0 / (double)1;
    for (int64_t j = 0; j < n; j++) Y[j] = 0.0 + (double)(range * drand[n + j]);
    int64_t* deg = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = i + 1; j < n; j++) {
            const double dx = X[i] - X[j], dy = Y[i] - Y[j];
            if (sqrt(dx * dx + dy * dy) <= rn) { deg[i + 1]++; deg[j + 1]++; }
        }
    for (int64_t i = 0; i < n; i++) deg[i + 1] += deg[i];
    if (graph_alloc(out, n, deg[n])) return -2;
    memcpy(out->offs, deg, sizeof(int64_t) * (size_t)(n + 1));
    int64_t* fill = (int64_t*)calloc((size_t)n, sizeof(int64_t));
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = i + 1; j < n; j++) {
            const double dx = X[i] - X[j], dy = Y[i] - Y[j];
            const double ed = sqrt(dx * dx + dy * dy);
            if (ed <= rn) {
                out->edges[deg[i] + fill[i]++] = (or_edge){j, ed};
                out->edges[deg[j] + fill[j]++] = (or_edge){i, ed};
            }
        }
    /* rows are appended in ascending partner order except that, for row j, partners
     * i < j arrive before partners > j; both runs ascend, so the row is sorted. */
    free(fill); free(deg); free(rnums); free(drand); free(Y);
    return 0;
}

/* ------------------------------------------------------------------------------ */
/* Ground-truth comparison                                                          */
/* ------------------------------------------------------------------------------ */

/* Members of each community id in [0, nc) grouped in vertex order (the counting layout of
 * compare.cpp:97-140 at one thread); sizes[c] is community c's member count. */
static void group_members(int64_t n, const int64_t* c, int64_t nc, int64_t* ptr, int64_t* idx, int64_t* sizes)
{
    int64_t* added = (int64_t*)calloc((size_t)nc, sizeof(int64_t));
    for (int64_t i = 0; i <= nc; i++) ptr[i] = 0;
    for (int64_t i = 0; i < n; i++) ptr[c[i] + 1]++;
    for (int64_t i = 0; i < nc; i++) sizes[i] = ptr[i + 1];
    for (int64_t i = 0; i < nc; i++) ptr[i + 1] += ptr[i];
    for (int64_t i = 0; i < n; i++) idx[ptr[c[i]] + added[c[i]]++] = i;
    free(added);
}

static int cmp_i64(const void* a, const void* b)
{
    const int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return x < y ? -1 : x > y;
}

/* compute_gini_coeff (compare.cpp:310-337): sizes sorted ascending, then the one-thread left
 * folds of (i+1)*size[i] and size[i]; numColors is an int there. */
static double gini_coeff(int64_t* sizes, int numColors)
{
    qsort(sizes, (size_t)numColors, sizeof(int64_t), cmp_i64);
    double numFunc = 0.0, denFunc = 0.0;
    for (int64_t i = 0; i < numColors; i++) {
        numFunc = numFunc + (double)((i + 1) * sizes[i]);
        denFunc = denFunc + (double)sizes[i];
    }
    return ((2 * numFunc) / (numColors * denFunc)) - ((double)(numColors + 1) / (double)numColors);
}

/* compare_communities (compare.cpp:58-307): C1 = truth, C2 = found.  For every pair of
 * members of a C1 community, same-same if C2 also joins them, else same-diff; for every
 * pair of members of a C2 community, diff-same if C1 separates them.  Then precision,
 * recall and F-score as at compare.cpp:266-272. */
int or_compare(int64_t n, const int64_t* truth, const int64_t* found, or_compare_stats* out)
{
    if (n < 1) return -1;
    int64_t nc1 = -1, nc2 = -1;
    for (int64_t i = 0; i < n; i++) {
        if (truth[i] < 0 || found[i] < 0) return -1;
        if (truth[i] > nc1) nc1 = truth[i];
        if (found[i] > nc2) nc2 = found[i];
    }
    nc1++;
    nc2++;
    int64_t* ptr1 = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nc1 + 1));
    int64_t* idx1 = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    int64_t* size1 = (int64_t*)malloc(sizeof(int64_t) * (size_t)nc1);
    int64_t* ptr2 = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nc2 + 1));
    int64_t* idx2 = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    int64_t* size2 = (int64_t*)malloc(sizeof(int64_t) * (size_t)nc2);
    group_members(n, truth, nc1, ptr1, idx1, size1);
    group_members(n, found, nc2, ptr2, idx2, size2);
    int64_t ss = 0, sd = 0, ds = 0, agree = 0;
    for (int64_t c = 0; c < nc1; c++)
        for (int64_t i = ptr1[c]; i < ptr1[c + 1]; i++)
            for (int64_t j = i + 1; j < ptr1[c + 1]; j++) {
                if (found[idx1[i]] == found[idx1[j]]) ss++;
                else sd++;
            }
    for (int64_t c = 0; c < nc2; c++)
        for (int64_t i = ptr2[c]; i < ptr2[c + 1]; i++)
            for (int64_t j = i + 1; j < ptr2[c + 1]; j++) {
                if (truth[idx2[i]] == truth[idx2[j]]) agree++;
                else ds++;
            }
    out->n = n;
    out->nc_truth = nc1;
    out->nc_found = nc2;
    out->same_same = ss;
    out->same_diff = sd;
    out->diff_same = ds;
    out->precision = (double)ss / (double)(ss + ds);
    out->recall = (double)ss / (double)(ss + sd);
    out->f_score = 2.0 * ((out->precision * out->recall) / (out->precision + out->recall));
    out->gini_truth = gini_coeff(size1, (int)nc1);
    out->gini_found = gini_coeff(size2, (int)nc2);
    free(ptr1); free(idx1); free(size1); free(ptr2); free(idx2); free(size2);
    return agree == ss ? 0 : -2;  /* the reference asserts the two counts agree */
}
