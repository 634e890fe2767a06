// The phase loop of main.cpp:242-513 on device-resident graphs: threshold cycling,
// coloring / vertex ordering in phase 0, early termination, rebuild-or-exit, the final
// 1e-6 pass and the phase / iteration caps.  Every phase calls the same device phase as
// vt_louvain_phase; every rebuild the same coarsening as vt_rebuild.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <vector>

#include "vt_internal.hpp"

namespace vt {

double run_phase(vt_ctx* ctx, vt_graph* g, const vt_phase_params* p, int64_t* cvect_out, int32_t* iters_out);
int64_t color_graph(vt_ctx* ctx, const vt_graph* g, int64_t nHash, int target_percent, bool single, int64_t* d_colors);

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void run_loop(vt_ctx* ctx, const vt_graph* g0, const vt_run_params* rp, vt_run_result* res)
{
    std::memset(res, 0, sizeof(*res));
    if (rp->et_type < 0 || rp->et_type > 4) fail(ctx, VT_ERR_ARG, "vt_run: et_type must be 0..4");
    if ((rp->et_type == 2 || rp->et_type == 4) && (rp->et_delta > 1.0 || rp->et_delta < 0.0))
        fail(ctx, VT_ERR_ARG, "vt_run: et_delta must be in [0, 1] (main.cpp:720-723)");
    if (rp->coloring && rp->vertex_order) fail(ctx, VT_ERR_ARG, "vt_run: coloring and vertex ordering are exclusive");
    if (rp->one_phase && rp->threshold_cycling) fail(ctx, VT_ERR_ARG, "vt_run: one-phase and threshold cycling are exclusive");

    // The working graph: the caller's at phase 0, then each rebuilt level (owned here).
    vt_graph* dg = const_cast<vt_graph*>(g0);
    vt_graph* owned = nullptr;
    int64_t* saved_part = g0->part;  // restore the caller graph's partition afterwards
    const_cast<vt_graph*>(g0)->part = nullptr;

    double currMod = -1.0, prevMod = -1.0, threshold = rp->threshold;
    int phase = 0, short_phase = 0, tot_iters = 0;
    int32_t iters = 0;
    const double t_start = now_s();
    double ctime = 0.0;
    auto cleanup = [&]() {
        if (owned) graph_delete(ctx, owned);
        owned = nullptr;
        vt_graph* g = const_cast<vt_graph*>(g0);
        if (g->part) dfree(ctx, g->part, 8 * (size_t)std::max<int64_t>(g->nv, 1));
        g->part = saved_part;
    };
    try {
        for (;;) {
            if (rp->threshold_cycling && !rp->one_phase) {
                if (short_phase > 12) short_phase = 0;
                if (short_phase <= 2) threshold = 1.0E-3;
                else if (short_phase <= 6) threshold = 1.0E-4;
                else if (short_phase <= 9) threshold = 1.0E-5;
                else threshold = rp->threshold;
            } else {
                threshold = rp->threshold;
            }
            vt_phase_params pp;
            std::memset(&pp, 0, sizeof(pp));
            pp.lower = currMod;
            pp.thresh = threshold;
            pp.et_type = rp->et_type;
            pp.et_delta = rp->et_delta;
            pp.flags = rp->flags;
            std::vector<int64_t> colors;
            const double t1 = now_s();
            if ((rp->coloring || rp->vertex_order) && phase == 0) {
                DBuf<int64_t> dc(ctx, (size_t)std::max<int64_t>(dg->nv, 1));
                const int64_t numColors = color_graph(ctx, dg, rp->max_colors / 2, 70, rp->single_color_iter != 0, dc.p);
                colors.resize(dg->nv);
                d2h(ctx, colors.data(), dc.p, 8 * (size_t)dg->nv);
                res->num_colors = numColors;
                pp.mode = rp->coloring ? VT_MODE_COLOR : VT_MODE_VORDER;
                pp.num_color = numColors + 1;
                pp.colors = colors.data();
            }
            const int lvl = std::min(res->nlevels, VT_MAX_LEVELS - 1);
            res->level_nv[lvl] = dg->nv;
            res->level_ne[lvl] = dg->ne;
            currMod = run_phase(ctx, dg, &pp, nullptr, &iters);
            if (res->ncalls <= VT_MAX_LEVELS) {
                res->call_q[res->ncalls] = currMod;
                res->call_iters[res->ncalls] = iters;
                res->ncalls++;
            }
            const double t0 = now_s();
            ctime += t0 - t1;
            tot_iters += iters;

            if ((currMod - prevMod) > threshold) {
                if (!rp->one_phase) {
                    vt_graph* next = rebuild(ctx, dg, dg->part, false);
                    if (owned) graph_delete(ctx, owned);
                    owned = next;
                    dg = next;
                }
            } else {
                // main.cpp:451-461: with threshold cycling and an early exit, one more plain
                // phase at 1e-6 sets the final modularity.
                if (rp->threshold_cycling && !rp->one_phase && phase < 10) {
                    std::memset(&pp, 0, sizeof(pp));
                    pp.lower = currMod;
                    pp.thresh = 1.0E-6;
                    pp.flags = rp->flags;
                    const double tx = now_s();
                    currMod = run_phase(ctx, dg, &pp, nullptr, &iters);
                    ctime += now_s() - tx;
                    if (res->ncalls <= VT_MAX_LEVELS) {
                        res->call_q[res->ncalls] = currMod;
                        res->call_iters[res->ncalls] = iters;
                        res->ncalls++;
                    }
                }
                break;
            }
            if (res->nlevels < VT_MAX_LEVELS) {
                res->level_q[res->nlevels] = currMod;
                res->level_iters[res->nlevels] = tot_iters;
                res->nlevels++;
            }
            prevMod = currMod;
            currMod = -1.0;
            phase++;
            if (rp->one_phase) break;
            if (rp->threshold_cycling && !rp->one_phase) short_phase++;
            if (phase == 200 || tot_iters > 10000) break;  // main.cpp:504-512
        }
    } catch (...) {
        cleanup();
        throw;
    }
    res->final_q = currMod;
    res->total_iters = tot_iters;
    res->clustering_seconds = ctime;
    res->total_seconds = now_s() - t_start;
    cleanup();
}

}  // namespace vt

using namespace vt;

extern "C" int vt_run(vt_ctx* ctx, const vt_graph* g, const vt_run_params* p, vt_run_result* res)
{
    if (!ctx || !g || !p || !res) return VT_ERR_ARG;
    return guarded(ctx, [&] {
        VT_HIP(ctx, hipSetDevice(ctx->device));
        run_loop(ctx, g, p, res);
        sync(ctx);
    });
}
