// Distance-1 coloring: distColoringMultiHashMinMax (coloring.cpp:52-118) at world size 1.
//
// Each round, every uncolored vertex evaluates nHash seeded hashes against the hashes of
// its neighbours that were uncolored at the start of the round (a neighbour colored in
// this round carries a color >= nextColor, which the reference's skip test ignores, so the
// round reads only round-start state, coloring.cpp:173-181).  A vertex that is the strict
// or tied minimum / maximum for some hash takes a color from the remaining min/max slots
// (coloring.cpp:220-242).  One lane per vertex, or a wave per vertex for rows of degree
// kColorLongRow and more; rounds repeat until the reference's stop rule fires (single
// round, >= target_pct colored, or no progress).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "vt_internal.hpp"

namespace vt {

constexpr int kMaxHash = 32;

// The seeded 32-bit mix of coloring.cpp:120-131 (six add/xor/shift steps published by
// Robert Jenkins).
__host__ __device__ inline uint32_t color_hash(uint32_t a, uint32_t seed)
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

__global__ void k_color_init(int64_t* colors, int64_t nv)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x)
        colors[i] = -1;
}

// Reads the round-start colors from `color` and writes the round's result to `next`.
// The color a vertex takes from the min / max slots its neighbours left (coloring.cpp:220-242),
// or -1.  The slots are the same whatever order the neighbours were seen in, so this is
// also the sequential loop's answer (its early exit only skips neighbours once no slot is
// left).
__device__ __forceinline__ int64_t choose_color(int64_t v, int64_t nHash, int64_t nextColor, unsigned not_min,
                                                unsigned not_max)
{
    const int64_t possible = 2 * nHash - __popc(not_min) - __popc(not_max);
    if (possible <= 0) return -1;
    const int64_t col_id = v % possible;
    int64_t this_col_id = 0;
    for (int64_t t = 0; t < nHash; t++) {
        if (!(not_min & (1u << t)) && col_id == this_col_id) return 2 * t + nextColor;
        this_col_id += !(not_min & (1u << t));
        if (!(not_max & (1u << t)) && col_id == this_col_id) return 2 * t + 1 + nextColor;
        this_col_id += !(not_max & (1u << t));
    }
    return -1;
}

// Rows of at least this degree are colored by a wave each (k_color_round_long): one thread
// walking a hub's 36 K neighbours held a round for 3 ms on BA n = 2^22.
constexpr int64_t kColorLongRow = 128;

__global__ void k_color_round(const int64_t* __restrict__ offs, const DEdge* __restrict__ edges, int64_t nv,
                              const int64_t* __restrict__ color, int64_t* __restrict__ next, int64_t nHash,
                              int64_t nextColor, uint32_t seed, unsigned long long* __restrict__ uncolored,
                              int64_t long_min)
{
    unsigned long long left = 0;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x) {
        if (color[v] != -1) { next[v] = color[v]; continue; }
        if (offs[v + 1] - offs[v] >= long_min) continue;  // k_color_round_long
        uint32_t vHash[kMaxHash];
        int64_t possible = 2 * nHash;
        unsigned not_min = 0, not_max = 0;
        bool can_color = true;
        for (int64_t t = 0; t < nHash; t++) vHash[t] = color_hash((uint32_t)v, seed + 1043u * (uint32_t)t);
        for (int64_t k = offs[v]; k < offs[v + 1]; k++) {
            const int64_t tail = edges[k].tail;
            if (tail == v) continue;
            const int64_t ct = color[tail];
            if (ct != -1 && ct < nextColor) continue;
            for (int64_t t = 0; t < nHash; t++) {
                const uint32_t jHash = color_hash((uint32_t)tail, seed + 1043u * (uint32_t)t);
                if (vHash[t] <= jHash && !(not_max & (1u << t))) { not_max |= (1u << t); possible--; }
                if (vHash[t] >= jHash && !(not_min & (1u << t))) { not_min |= (1u << t); possible--; }
            }
            if (possible == 0) { can_color = false; break; }
        }
        int64_t chosen = -1;
        if (can_color) {
            const int64_t col_id = v % possible;
            int64_t this_col_id = 0;
            for (int64_t t = 0; t < nHash; t++) {
                if (!(not_min & (1u << t)) && col_id == this_col_id) { chosen = 2 * t + nextColor; break; }
                this_col_id += !(not_min & (1u << t));
                if (!(not_max & (1u << t)) && col_id == this_col_id) { chosen = 2 * t + 1 + nextColor; break; }
                this_col_id += !(not_max & (1u << t));
            }
        }
        next[v] = chosen;
        if (chosen < 0) left++;
    }
    for (int o = 32; o > 0; o >>= 1) left += __shfl_xor(left, o);
    if ((threadIdx.x & 63) == 0 && left) atomicAdd(uncolored, left);
}

// Uncolored rows of degree >= long_min: a wave per row, each lane testing every 64th
// neighbour; the lanes' min / max masks are OR-ed across the wave (the sets do not depend
// on order), then choose_color as for the short rows.
__global__ void __launch_bounds__(256) k_color_round_long(const int64_t* __restrict__ offs,
                                                          const DEdge* __restrict__ edges, int64_t nv,
                                                          const int64_t* __restrict__ color, int64_t* __restrict__ next,
                                                          int64_t nHash, int64_t nextColor, uint32_t seed,
                                                          unsigned long long* __restrict__ uncolored, int64_t long_min)
{
    const int lane = threadIdx.x & 63;
    unsigned long long left = 0;
    for (int64_t v = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); v < nv; v += (int64_t)gridDim.x * 4) {
        const int64_t e0 = offs[v], e1 = offs[v + 1];
        if (e1 - e0 < long_min || color[v] != -1) continue;  // wave-uniform
        uint32_t vHash[kMaxHash];
        for (int64_t t = 0; t < nHash; t++) vHash[t] = color_hash((uint32_t)v, seed + 1043u * (uint32_t)t);
        unsigned not_min = 0, not_max = 0;
        const unsigned all = nHash >= 32 ? ~0u : ((1u << nHash) - 1);
        for (int64_t k = e0 + lane; k < e1; k += 64) {
            const int64_t tail = edges[k].tail;
            if (tail == v) continue;
            const int64_t ct = color[tail];
            if (ct != -1 && ct < nextColor) continue;
            for (int64_t t = 0; t < nHash; t++) {
                const uint32_t jHash = color_hash((uint32_t)tail, seed + 1043u * (uint32_t)t);
                if (vHash[t] <= jHash) not_max |= 1u << t;
                if (vHash[t] >= jHash) not_min |= 1u << t;
            }
            if (not_min == all && not_max == all) break;  // no slot left in this lane's view
        }
        for (int o = 32; o > 0; o >>= 1) {
            not_min |= (unsigned)__shfl_xor((int)not_min, o);
            not_max |= (unsigned)__shfl_xor((int)not_max, o);
        }
        const int64_t chosen = choose_color(v, nHash, nextColor, not_min, not_max);
        if (lane == 0) {
            next[v] = chosen;
            if (chosen < 0) left++;
        }
    }
    if (lane == 0 && left) atomicAdd(uncolored, left);
}

int64_t color_graph(vt_ctx* ctx, const vt_graph* g, int64_t nHash, int target_percent, bool single, int64_t* d_colors)
{
    if (nHash < 1 || nHash > kMaxHash)
        fail(ctx, VT_ERR_ARG, "vt_color: nhash must be in [1, %d] (got %lld)", kMaxHash, (long long)nHash);
    const int64_t tnv = g->nv;
    uint32_t seed = 1012;
    int64_t nextColor = 0, lastCount = 0, currentCount = 0;
    const int64_t targetCount = (tnv * target_percent) / 100;
    DBuf<unsigned long long> cnt(ctx, 2);
    DBuf<int64_t> other(ctx, (size_t)std::max<int64_t>(tnv, 1));
    int64_t* cur = d_colors;
    int64_t* nxt = other.p;
    if (tnv) VT_LAUNCH(ctx, "color_init", k_color_init, grid_for(tnv, 256, 8192), 256, 0, cur, tnv);
    bool finished = false;
    while (!finished) {
        dzero(ctx, cnt.p, 8);
        if (tnv) {
            const int64_t long_min = g->max_degree >= kColorLongRow ? kColorLongRow : INT64_MAX;
            VT_LAUNCH(ctx, "color_round", k_color_round, grid_for(tnv, 256, 16384), 256, 0, g->offs, g->edges, tnv,
                      (const int64_t*)cur, nxt, nHash, nextColor, seed, cnt.p, long_min);
            if (long_min != INT64_MAX)
                VT_LAUNCH(ctx, "color_round_long", k_color_round_long, grid_for(tnv, 4, 16384), 256, 0, g->offs,
                          g->edges, tnv, (const int64_t*)cur, nxt, nHash, nextColor, seed, cnt.p, long_min);
        }
        std::swap(cur, nxt);
        unsigned long long unassigned = 0;
        d2h(ctx, &unassigned, cnt.p, 8);
        currentCount = tnv - (int64_t)unassigned;
        finished = single || currentCount >= targetCount || lastCount == currentCount;
        nextColor += nHash * 2;
        seed = color_hash(seed, 0);
        lastCount = currentCount;
    }
    if (cur != d_colors) d2d(ctx, d_colors, cur, 8 * (size_t)tnv);
    sync(ctx);
    return nextColor;
}

}  // namespace vt

using namespace vt;

extern "C" int vt_color(vt_ctx* ctx, const vt_graph* g, int64_t nhash, int32_t target_pct, int32_t single_iter,
                        int64_t* colors_out, int64_t* next_color_out)
{
    if (!ctx || !g) return VT_ERR_ARG;
    return guarded(ctx, [&] {
        VT_HIP(ctx, hipSetDevice(ctx->device));
        DBuf<int64_t> c(ctx, (size_t)std::max<int64_t>(g->nv, 1));
        const int64_t nc = color_graph(ctx, g, nhash, target_pct, single_iter != 0, c.p);
        if (colors_out) d2h(ctx, colors_out, c.p, 8 * (size_t)g->nv);
        if (next_color_out) *next_color_out = nc;
        sync(ctx);
    });
}
