// Batched form: many independent graphs run the whole phase loop of main.cpp:242-513 as one
// device-resident disjoint union.  Every Louvain phase of every graph runs in the same
// launches (batch_phase), each graph stopping on its own test; between phases the union is
// coarsened once (every member's partition at the same time) and reassembled from the
// members that go on.  Per graph, every call's return value, iteration count and the final
// modularity equal what vt_run gives on that graph alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>
#include <numeric>
#include <vector>

#include "vt_internal.hpp"

struct vt_batch {
    int32_t ng = 0;
    std::vector<int64_t> voff, eoff;  // level-0 vertex / edge offsets of each graph in the union
    vt_graph* u = nullptr;            // level-0 union
    vt::DBuf<int32_t> gid;            // graph of each union vertex
};

namespace vt {

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Largest g in [0, n) with off[g] <= x (off ascending; empty ranges are skipped over).
__device__ __forceinline__ int32_t segment_of(const int64_t* __restrict__ off, int32_t n, int64_t x)
{
    int32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (off[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Destination vertex i of piece g copies its source row start, rebased to the piece's edge
// offset in the union; gid records the piece.
__global__ void k_assemble_offs(int32_t ng, const int64_t* __restrict__ dvoff, const int64_t* __restrict__ deoff,
                                const int64_t* const* __restrict__ soffs, const int64_t* __restrict__ sv0,
                                const int64_t* __restrict__ se0, int64_t nv, int64_t* __restrict__ offs,
                                int32_t* __restrict__ gid)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t g = segment_of(dvoff, ng, i);
        const int64_t local = i - dvoff[g];
        offs[i] = soffs[g][sv0[g] + local] - se0[g] + deoff[g];
        gid[i] = g;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) offs[nv] = deoff[ng];
}

// Destination edge k of piece g copies its source record with the tail moved from the
// piece's source vertex range to its union range.
__global__ void k_assemble_edges(int32_t ng, const int64_t* __restrict__ dvoff, const int64_t* __restrict__ deoff,
                                 const DEdge* const* __restrict__ sedges, const int64_t* __restrict__ sv0,
                                 const int64_t* __restrict__ se0, int64_t ne, DEdge* __restrict__ edges)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < ne; k += (int64_t)gridDim.x * blockDim.x) {
        const int32_t g = segment_of(deoff, ng, k);
        DEdge e = sedges[g][se0[g] + (k - deoff[g])];
        e.tail = e.tail - sv0[g] + dvoff[g];
        edges[k] = e;
    }
}

__global__ void k_fill_i64(int64_t* __restrict__ p, int64_t n, int64_t v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_first_member(int64_t nv, const int64_t* __restrict__ part, unsigned long long* __restrict__ first)
{
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x)
        atomicMin(&first[part[v]], (unsigned long long)v);
}

// Communities per graph: the coarse graph numbers communities by first appearance over
// the union, so graph g's coarse vertices are a contiguous range of this length.
__global__ void k_count_communities(int64_t nv, const int64_t* __restrict__ part,
                                    const unsigned long long* __restrict__ first, const int32_t* __restrict__ gid,
                                    int32_t* __restrict__ cnt)
{
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x)
        if (first[part[v]] == (unsigned long long)v) atomicAdd(&cnt[gid[v]], 1);
}

__global__ void k_gather_i64(const int64_t* __restrict__ src, const int64_t* __restrict__ idx, int32_t n,
                             int64_t* __restrict__ out)
{
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = src[idx[i]];
}

struct Piece {
    const int64_t* offs;
    const DEdge* edges;
    int64_t v0, e0, nv, ne;
};

// Builds a union graph from pieces in order; returns it with its gid array and the host
// vertex / edge offsets of each piece.
vt_graph* assemble(vt_ctx* ctx, const std::vector<Piece>& pcs, DBuf<int32_t>& gid, std::vector<int64_t>& voff,
                   std::vector<int64_t>& eoff)
{
    const int32_t ng = (int32_t)pcs.size();
    voff.assign((size_t)ng + 1, 0);
    eoff.assign((size_t)ng + 1, 0);
    std::vector<const int64_t*> so(ng);
    std::vector<const DEdge*> se(ng);
    std::vector<int64_t> sv0(ng), se0(ng);
    for (int32_t g = 0; g < ng; g++) {
        voff[g + 1] = voff[g] + pcs[g].nv;
        eoff[g + 1] = eoff[g] + pcs[g].ne;
        so[g] = pcs[g].offs;
        se[g] = pcs[g].edges;
        sv0[g] = pcs[g].v0;
        se0[g] = pcs[g].e0;
    }
    const int64_t nv = voff[ng], ne = eoff[ng];
    if (nv >= (int64_t(1) << 31) - 1) fail(ctx, VT_ERR_LIMIT, "batch: %lld vertices in all exceed 2^31 - 2", (long long)nv);
    DBuf<int64_t> dvoff(ctx, ng + 1), deoff(ctx, ng + 1), dsv0(ctx, ng), dse0(ctx, ng);
    DBuf<const int64_t*> dso(ctx, ng);
    DBuf<const DEdge*> dse(ctx, ng);
    h2d(ctx, dvoff.p, voff.data(), 8 * voff.size());
    h2d(ctx, deoff.p, eoff.data(), 8 * eoff.size());
    h2d(ctx, dsv0.p, sv0.data(), 8 * (size_t)ng);
    h2d(ctx, dse0.p, se0.data(), 8 * (size_t)ng);
    h2d(ctx, dso.p, so.data(), sizeof(void*) * (size_t)ng);
    h2d(ctx, dse.p, se.data(), sizeof(void*) * (size_t)ng);
    vt_graph* u = graph_new(ctx, nv, ne);
    try {
        gid = DBuf<int32_t>(ctx, (size_t)std::max<int64_t>(nv, 1));
        VT_LAUNCH(ctx, "assemble_offs", k_assemble_offs, grid_for(std::max<int64_t>(nv, 1), 256, 8192), 256, 0, ng,
                  (const int64_t*)dvoff.p, (const int64_t*)deoff.p, (const int64_t* const*)dso.p,
                  (const int64_t*)dsv0.p, (const int64_t*)dse0.p, nv, u->offs, gid.p);
        if (ne)
            VT_LAUNCH(ctx, "assemble_edges", k_assemble_edges, grid_for(ne, 256, 16384), 256, 0, ng,
                      (const int64_t*)dvoff.p, (const int64_t*)deoff.p, (const DEdge* const*)dse.p,
                      (const int64_t*)dsv0.p, (const int64_t*)dse0.p, ne, u->edges);
        graph_finalize(ctx, u);
    } catch (...) {
        graph_delete(ctx, u);
        throw;
    }
    return u;
}

enum { kRun = 0, kFinalPass = 1, kDone = 2 };

struct GraphLoop {
    int state = kRun;
    double currMod = -1.0, prevMod = -1.0;
    int phase = 0, short_phase = 0, tot_iters = 0;
};

void record_call(vt_run_result& r, double q, int32_t iters)
{
    if (r.ncalls <= VT_MAX_LEVELS) {
        r.call_q[r.ncalls] = q;
        r.call_iters[r.ncalls] = iters;
        r.ncalls++;
    }
}

// The loop of run_loop (vt_driver.hip) for every graph of the batch, one batched phase per
// round.  A graph whose loop would run its extra 1e-6 pass (threshold cycling, early exit)
// takes part in the next round with its current graph instead of the coarsened one.
void batch_run(vt_ctx* ctx, vt_batch* b, const vt_run_params* rp, vt_run_result* res, int* phases_out,
               int* rounds_out, int* relists_out)
{
    if (rp->coloring || rp->vertex_order || rp->et_type != 0)
        fail(ctx, VT_ERR_LIMIT, "vt_batch_run: the batched form runs plain Louvain (no coloring, no early termination)");
    if (rp->one_phase && rp->threshold_cycling)
        fail(ctx, VT_ERR_ARG, "vt_batch_run: one-phase and threshold cycling are exclusive");
    const int32_t ng = b->ng;
    for (int32_t g = 0; g < ng; g++) std::memset(&res[g], 0, sizeof(vt_run_result));
    std::vector<GraphLoop> gl((size_t)ng);
    const bool use_graph = !(rp->flags & VT_FLAG_NO_GRAPH);
    const int sum_flags = rp->flags & VT_FLAG_SEQUENTIAL_SUMS;
    const double t_start = now_s();

    std::vector<int32_t> mem((size_t)ng);
    std::iota(mem.begin(), mem.end(), 0);
    vt_graph* cur = b->u;
    bool cur_owned = false;
    DBuf<int32_t> cur_gid_owned;
    const int32_t* cur_gid = b->gid.p;
    std::vector<int64_t> voff = b->voff, eoff = b->eoff;
    int phases = 0, rounds_total = 0, relists_total = 0;
    auto drop_cur = [&]() {
        if (cur_owned) graph_delete(ctx, cur);
        else if (cur->part) {
            dfree(ctx, cur->part, 8 * (size_t)std::max<int64_t>(cur->nv, 1));
            cur->part = nullptr;
        }
        cur_owned = false;
    };
    try {
        while (!mem.empty()) {
            const int32_t m = (int32_t)mem.size();
            std::vector<double> lower(m), thresh(m);
            for (int32_t j = 0; j < m; j++) {
                GraphLoop& s = gl[mem[j]];
                vt_run_result& r = res[mem[j]];
                if (s.state == kFinalPass) {
                    lower[j] = s.currMod;
                    thresh[j] = 1.0E-6;
                    continue;
                }
                double threshold = rp->threshold;
                if (rp->threshold_cycling && !rp->one_phase) {
                    if (s.short_phase > 12) s.short_phase = 0;
                    if (s.short_phase <= 2) threshold = 1.0E-3;
                    else if (s.short_phase <= 6) threshold = 1.0E-4;
                    else if (s.short_phase <= 9) threshold = 1.0E-5;
                }
                lower[j] = s.currMod;
                thresh[j] = threshold;
                const int lvl = std::min(r.nlevels, VT_MAX_LEVELS - 1);
                r.level_nv[lvl] = voff[j + 1] - voff[j];
                r.level_ne[lvl] = eoff[j + 1] - eoff[j];
            }
            SegSums seg;
            seg_sums_plan(ctx, voff, seg);
            std::vector<double> q;
            std::vector<int32_t> it;
            int rounds = 0;
            int relists = 0;
            batch_phase(ctx, cur, cur_gid, seg, lower, thresh, sum_flags, use_graph, q, it, &rounds, &relists);
            relists_total += relists;
            phases++;
            rounds_total += rounds;

            std::vector<int32_t> next_mem;
            std::vector<char> coarse;  // per next member: from the coarsened union (1) or the current one (0)
            std::vector<int32_t> from;  // its index among the current members
            for (int32_t j = 0; j < m; j++) {
                const int32_t g = mem[j];
                GraphLoop& s = gl[g];
                vt_run_result& r = res[g];
                if (s.state == kFinalPass) {
                    s.currMod = q[j];
                    record_call(r, q[j], it[j]);
                    s.state = kDone;
                    continue;
                }
                s.currMod = q[j];
                record_call(r, q[j], it[j]);
                s.tot_iters += it[j];
                bool go_on = false;
                if ((s.currMod - s.prevMod) > thresh[j]) {
                    go_on = !rp->one_phase;
                } else {
                    if (rp->threshold_cycling && !rp->one_phase && s.phase < 10) {
                        s.state = kFinalPass;
                        next_mem.push_back(g);
                        coarse.push_back(0);
                        from.push_back(j);
                    } else {
                        s.state = kDone;
                    }
                    continue;
                }
                if (r.nlevels < VT_MAX_LEVELS) {
                    r.level_q[r.nlevels] = s.currMod;
                    r.level_iters[r.nlevels] = s.tot_iters;
                    r.nlevels++;
                }
                s.prevMod = s.currMod;
                s.currMod = -1.0;
                s.phase++;
                if (rp->one_phase) go_on = false;
                if (rp->threshold_cycling && !rp->one_phase) s.short_phase++;
                if (s.phase == 200 || s.tot_iters > 10000) go_on = false;  // main.cpp:504-512
                if (go_on) {
                    next_mem.push_back(g);
                    coarse.push_back(1);
                    from.push_back(j);
                } else {
                    s.state = kDone;
                }
            }
            if (next_mem.empty()) {
                drop_cur();
                mem.clear();
                break;
            }
            // Coarsen every member at once, then keep the members that go on.
            vt_graph* cg = nullptr;
            std::vector<int64_t> cvoff, ceoff;
            if (std::find(coarse.begin(), coarse.end(), 1) != coarse.end()) {
                const int64_t nv = cur->nv;
                DBuf<unsigned long long> first(ctx, (size_t)std::max<int64_t>(nv, 1));
                DBuf<int32_t> cnt(ctx, m);
                VT_LAUNCH(ctx, "fill_i64", k_fill_i64, grid_for(nv, 256, 8192), 256, 0, (int64_t*)first.p, nv,
                          (int64_t)LLONG_MAX);
                dzero(ctx, cnt.p, 4 * (size_t)m);
                VT_LAUNCH(ctx, "first_member", k_first_member, grid_for(nv, 256, 8192), 256, 0, nv,
                          (const int64_t*)cur->part, first.p);
                VT_LAUNCH(ctx, "count_communities", k_count_communities, grid_for(nv, 256, 8192), 256, 0, nv,
                          (const int64_t*)cur->part, (const unsigned long long*)first.p, cur_gid, cnt.p);
                std::vector<int32_t> hc(m);
                d2h(ctx, hc.data(), cnt.p, 4 * (size_t)m);
                cg = rebuild(ctx, cur, cur->part, false);
                cvoff.assign((size_t)m + 1, 0);
                for (int32_t j = 0; j < m; j++) cvoff[j + 1] = cvoff[j] + hc[j];
                DBuf<int64_t> di(ctx, m + 1), dout(ctx, m + 1);
                h2d(ctx, di.p, cvoff.data(), 8 * (size_t)(m + 1));
                VT_LAUNCH(ctx, "gather_offs", k_gather_i64, grid_for(m + 1, 256, 64), 256, 0, (const int64_t*)cg->offs,
                          (const int64_t*)di.p, m + 1, dout.p);
                ceoff.assign((size_t)m + 1, 0);
                d2h(ctx, ceoff.data(), dout.p, 8 * (size_t)(m + 1));
            }
            std::vector<Piece> pcs;
            for (size_t k = 0; k < next_mem.size(); k++) {
                const int32_t j = from[k];
                if (coarse[k])
                    pcs.push_back(Piece{cg->offs, cg->edges, cvoff[j], ceoff[j], cvoff[j + 1] - cvoff[j],
                                        ceoff[j + 1] - ceoff[j]});
                else
                    pcs.push_back(Piece{cur->offs, cur->edges, voff[j], eoff[j], voff[j + 1] - voff[j],
                                        eoff[j + 1] - eoff[j]});
            }
            DBuf<int32_t> ngid;
            std::vector<int64_t> nvoff, neoff;
            vt_graph* nu = nullptr;
            try {
                nu = assemble(ctx, pcs, ngid, nvoff, neoff);
            } catch (...) {
                if (cg) graph_delete(ctx, cg);
                throw;
            }
            if (cg) graph_delete(ctx, cg);
            drop_cur();
            cur = nu;
            cur_owned = true;
            cur_gid_owned = std::move(ngid);
            cur_gid = cur_gid_owned.p;
            voff = std::move(nvoff);
            eoff = std::move(neoff);
            mem = std::move(next_mem);
        }
    } catch (...) {
        drop_cur();
        throw;
    }
    const double secs = now_s() - t_start;
    for (int32_t g = 0; g < ng; g++) {
        res[g].final_q = gl[g].currMod;
        res[g].total_iters = gl[g].tot_iters;
        res[g].clustering_seconds = secs;
        res[g].total_seconds = secs;
    }
    if (phases_out) *phases_out = phases;
    if (rounds_out) *rounds_out = rounds_total;
    if (relists_out) *relists_out = relists_total;
}

}  // namespace

}  // namespace vt

using namespace vt;

extern "C" {

int vt_batch_create(vt_ctx* ctx, int32_t n, vt_graph* const* graphs, vt_batch** out)
{
    if (!ctx || !out || n < 1 || !graphs) return VT_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        VT_HIP(ctx, hipSetDevice(ctx->device));
        std::vector<Piece> pcs;
        for (int32_t g = 0; g < n; g++) {
            if (!graphs[g]) fail(ctx, VT_ERR_ARG, "vt_batch_create: graph %d is null", g);
            // 2m = 0 leaves modularity undefined (the reference's loop would not end)
            if (graphs[g]->nv < 1 || graphs[g]->ne < 1)
                fail(ctx, VT_ERR_ARG, "vt_batch_create: graph %d has no edges", g);
            pcs.push_back(Piece{graphs[g]->offs, graphs[g]->edges, 0, 0, graphs[g]->nv, graphs[g]->ne});
        }
        vt_batch* b = new vt_batch();
        try {
            b->ng = n;
            b->u = assemble(ctx, pcs, b->gid, b->voff, b->eoff);
            sync(ctx);
        } catch (...) {
            if (b->u) graph_delete(ctx, b->u);
            delete b;
            throw;
        }
        *out = b;
    });
}

int vt_batch_info(const vt_batch* b, int32_t* n, int64_t* total_nv, int64_t* total_ne)
{
    if (!b) return VT_ERR_ARG;
    if (n) *n = b->ng;
    if (total_nv) *total_nv = b->u->nv;
    if (total_ne) *total_ne = b->u->ne;
    return VT_OK;
}

int vt_batch_run(vt_ctx* ctx, vt_batch* b, const vt_run_params* p, vt_run_result* results, vt_batch_stats* stats)
{
    if (!ctx || !b || !p || !results) return VT_ERR_ARG;
    return guarded(ctx, [&] {
        VT_HIP(ctx, hipSetDevice(ctx->device));
        const double t0 = now_s();
        int phases = 0, rounds = 0, relists = 0;
        batch_run(ctx, b, p, results, &phases, &rounds, &relists);
        sync(ctx);
        if (stats) {
            stats->phases = phases;
            stats->rounds = rounds;
            stats->seconds = now_s() - t0;
            stats->relists = relists;
        }
    });
}

int vt_batch_autotune(vt_ctx* ctx, vt_batch* b, int32_t prep, int32_t reps, vt_tuning* out)
{
    if (!ctx || !b) return VT_ERR_ARG;
    // The launch shapes depend on the degree mix, which the level-0 union carries: its sweeps
    // run as one graph (one 1/2m) touch the same rows and tables as the batched sweeps, and
    // no setting changes a result.
    return vt_autotune(ctx, b->u, prep, reps, out);
}

int vt_batch_free(vt_ctx* ctx, vt_batch* b)
{
    if (!ctx) return VT_ERR_ARG;
    return guarded(ctx, [&] {
        if (!b) return;
        VT_HIP(ctx, hipSetDevice(ctx->device));
        sync(ctx);
        b->gid.reset();
        graph_delete(ctx, b->u);
        delete b;
    });
}

}  // extern "C"
