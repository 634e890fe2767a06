// Coarsening: distbuildNextLevelGraph (rebuild.cpp:480-504) at world size 1.
//
//  1. renumber: new id of community c = rank of its first appearance in cvect as the
//     vertex index ascends (distReNumber, rebuild.cpp:121-135) -- atomicMin of the first
//     vertex per community, a flag scan over vertices, one gather;
//  2. one (C1,C2) key per fine edge with the edge's weight, radix-sorted stably, so equal
//     keys keep global edge order = "vertices of C1 ascending, each row in CSR order"
//     (fill_newEdgesMap, rebuild.cpp:294-329);
//  3. every key run is folded left to right by one thread, the first contribution stored
//     as is; runs come out sorted by C2 inside C1 (std::map order, rebuild.cpp:469-476).
//     Self-loops are kept.
#include <hip/hip_runtime.h>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <vector>

#include "vt_internal.hpp"

namespace vt {

__global__ void k_first_vertex(const int64_t* __restrict__ cvect, int64_t nv, unsigned long long* __restrict__ first,
                               unsigned long long* __restrict__ bad)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = cvect[i];
        if (c < 0 || c >= nv) { atomicAdd(bad, 1ull); continue; }
        atomicMin(&first[c], (unsigned long long)i);
    }
}

__global__ void k_fill_u64(unsigned long long* p, int64_t n, unsigned long long v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_first_flags(const int64_t* __restrict__ cvect, int64_t nv, const unsigned long long* __restrict__ first,
                              int64_t* __restrict__ flag)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x)
        flag[i] = (first[cvect[i]] == (unsigned long long)i) ? 1 : 0;
}

// newid[i] = rank of vertex i's community.
__global__ void k_vertex_newid(const int64_t* __restrict__ cvect, int64_t nv, const unsigned long long* __restrict__ first,
                               const int64_t* __restrict__ rank, uint32_t* __restrict__ newid)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x)
        newid[i] = (uint32_t)rank[first[cvect[i]]];
}

__global__ void k_total(const int64_t* __restrict__ flag, const int64_t* __restrict__ scan, int64_t n, int64_t* out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = n ? scan[n - 1] + flag[n - 1] : 0;
}

// One key per fine edge: (C1 << bits) | C2, value = weight.  A block takes 256 consecutive
// rows, caches their starts and new ids in LDS, and walks their (contiguous) edges with
// consecutive threads on consecutive edges; each thread finds its edge's row by binary
// search over the cached starts.
constexpr int kRowsPerBlock = 256;

__global__ void __launch_bounds__(kRowsPerBlock) k_edge_keys(const int64_t* __restrict__ offs,
                                                             const DEdge* __restrict__ edges, int64_t nv,
                                                             const uint32_t* __restrict__ newid, int bits,
                                                             unsigned long long* __restrict__ keys,
                                                             double* __restrict__ vals)
{
    __shared__ int64_t so[kRowsPerBlock + 1];
    __shared__ uint32_t sid[kRowsPerBlock];
    for (int64_t r0 = (int64_t)blockIdx.x * kRowsPerBlock; r0 < nv; r0 += (int64_t)gridDim.x * kRowsPerBlock) {
        const int nr = (int)(nv - r0 < kRowsPerBlock ? nv - r0 : kRowsPerBlock);
        if ((int)threadIdx.x < nr) {
            so[threadIdx.x] = offs[r0 + threadIdx.x];
            sid[threadIdx.x] = newid[r0 + threadIdx.x];
        }
        if (threadIdx.x == 0) so[nr] = offs[r0 + nr];
        __syncthreads();
        const int64_t e1 = so[nr];
        for (int64_t k = so[0] + threadIdx.x; k < e1; k += kRowsPerBlock) {
            int lo = 0, hi = nr - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (so[mid] <= k) lo = mid;
                else hi = mid - 1;
            }
            const DEdge e = edges[k];
            keys[k] = ((unsigned long long)sid[lo] << bits) | newid[e.tail];
            vals[k] = e.weight;
        }
        __syncthreads();
    }
}

__global__ void k_run_heads(const unsigned long long* __restrict__ keys, int64_t n, int64_t* __restrict__ head)
{
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
        head[j] = (j == 0 || keys[j] != keys[j - 1]) ? 1 : 0;
}

// starts[r] = first position of run r (pos = exclusive scan of the head flags).
__global__ void k_run_starts(const int64_t* __restrict__ head, const int64_t* __restrict__ pos, int64_t n, int64_t nu,
                             int64_t* __restrict__ starts)
{
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
        if (head[j]) starts[pos[j]] = j;
    if (blockIdx.x == 0 && threadIdx.x == 0) starts[nu] = n;
}

// Run r: left fold of its weights in sorted (= fine edge) order, the first contribution as
// is; the bounds are known up front, so the loads run ahead of the dependent adds.
__global__ void k_run_fold(const unsigned long long* __restrict__ keys, const double* __restrict__ vals,
                           const int64_t* __restrict__ starts, int64_t nu, int bits, DEdge* __restrict__ out,
                           unsigned long long* __restrict__ rowcount)
{
    const unsigned long long lomask = (1ull << bits) - 1;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nu; r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = starts[r], e = starts[r + 1];
        const unsigned long long key = keys[b];
        double acc = vals[b];
        int64_t k = b + 1;
        for (; k + 4 <= e; k += 4) {
            const double x0 = vals[k], x1 = vals[k + 1], x2 = vals[k + 2], x3 = vals[k + 3];
            acc = acc + x0;
            acc = acc + x1;
            acc = acc + x2;
            acc = acc + x3;
        }
        for (; k < e; k++) acc = acc + vals[k];
        DEdge o;
        o.tail = (int64_t)(key & lomask);
        o.weight = acc;
        out[r] = o;
        atomicAdd(&rowcount[key >> bits], 1ull);
    }
}

vt_graph* rebuild(vt_ctx* ctx, const vt_graph* g, const int64_t* d_cvect)
{
    const int64_t nv = g->nv, ne = g->ne;
    if (ne >= (int64_t(1) << 32)) fail(ctx, VT_ERR_LIMIT, "vt_rebuild: %lld edges; this build supports < 2^32", (long long)ne);
    const size_t n1 = (size_t)std::max<int64_t>(nv, 1);
    DBuf<unsigned long long> first(ctx, n1), bad(ctx, 2);
    DBuf<int64_t> flag(ctx, n1), rank(ctx, n1), tot(ctx, 2);
    DBuf<uint32_t> newid(ctx, n1);
    dzero(ctx, bad.p, 16);
    VT_LAUNCH(ctx, "rb_fill", k_fill_u64, grid_for(nv, 256, 8192), 256, 0, first.p, nv, ~0ull);
    VT_LAUNCH(ctx, "rb_first", k_first_vertex, grid_for(nv, 256, 8192), 256, 0, d_cvect, nv, first.p, bad.p);
    unsigned long long nbad = 0;
    d2h(ctx, &nbad, bad.p, 8);
    if (nbad) fail(ctx, VT_ERR_ARG, "vt_rebuild: %llu community ids outside [0, nv)", nbad);
    VT_LAUNCH(ctx, "rb_flags", k_first_flags, grid_for(nv, 256, 8192), 256, 0, d_cvect, nv,
              (const unsigned long long*)first.p, flag.p);
    exclusive_scan_i64(ctx, flag.p, rank.p, nv);
    VT_LAUNCH(ctx, "rb_total", k_total, 1, 64, 0, (const int64_t*)flag.p, (const int64_t*)rank.p, nv, tot.p);
    int64_t nc = 0;
    d2h(ctx, &nc, tot.p, 8);
    VT_LAUNCH(ctx, "rb_newid", k_vertex_newid, grid_for(nv, 256, 8192), 256, 0, d_cvect, nv,
              (const unsigned long long*)first.p, (const int64_t*)rank.p, newid.p);
    first.reset();
    flag.reset();
    rank.reset();

    const int bits = (int)bits_for(nc > 1 ? nc - 1 : 1);
    const size_t e1 = (size_t)std::max<int64_t>(ne, 1);
    DBuf<unsigned long long> k0(ctx, e1), k1(ctx, e1);
    DBuf<double> v0(ctx, e1), v1(ctx, e1);
    if (nv)
        VT_LAUNCH(ctx, "rb_keys", k_edge_keys, grid_for(nv, kRowsPerBlock, 16384), kRowsPerBlock, 0, g->offs, g->edges,
                  nv, (const uint32_t*)newid.p, bits, k0.p, v0.p);
    const unsigned long long* ks = k0.p;
    const double* vs = v0.p;
    if (ne > 1) {
        rocprim::double_buffer<unsigned long long> kb(k0.p, k1.p);
        rocprim::double_buffer<double> vb(v0.p, v1.p);
        size_t tb = 0;
        VT_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tb, kb, vb, (size_t)ne, 0, 2 * bits, ctx->stream));
        void* tmp = scratch(ctx, "sort_tmp", tb);
        {
            TimedScope ts(ctx, "rb_sort");
            VT_HIP(ctx, rocprim::radix_sort_pairs(tmp, tb, kb, vb, (size_t)ne, 0, 2 * bits, ctx->stream));
        }
        ks = kb.current();
        vs = vb.current();
    }
    DBuf<int64_t> head(ctx, e1), pos(ctx, e1);
    int64_t nu = 0;
    if (ne) {
        VT_LAUNCH(ctx, "rb_heads", k_run_heads, grid_for(ne, 256, 16384), 256, 0, ks, ne, head.p);
        exclusive_scan_i64(ctx, head.p, pos.p, ne);
        VT_LAUNCH(ctx, "rb_total", k_total, 1, 64, 0, (const int64_t*)head.p, (const int64_t*)pos.p, ne, tot.p);
        d2h(ctx, &nu, tot.p, 8);
    }
    vt_graph* out = graph_new(ctx, nc, nu);
    try {
        DBuf<unsigned long long> rc(ctx, (size_t)nc + 1);
        dzero(ctx, rc.p, 8 * ((size_t)nc + 1));
        if (ne) {
            DBuf<int64_t> starts(ctx, (size_t)nu + 1);
            VT_LAUNCH(ctx, "rb_starts", k_run_starts, grid_for(ne, 256, 16384), 256, 0, (const int64_t*)head.p,
                      (const int64_t*)pos.p, ne, nu, starts.p);
            VT_LAUNCH(ctx, "rb_fold", k_run_fold, grid_for(nu, 256, 16384), 256, 0, ks, vs, (const int64_t*)starts.p, nu,
                      bits, out->edges, rc.p);
        }
        exclusive_scan_i64(ctx, (const int64_t*)rc.p, out->offs, nc + 1);
        graph_finalize(ctx, out);
    } catch (...) {
        graph_delete(ctx, out);
        throw;
    }
    return out;
}

}  // namespace vt

using namespace vt;

extern "C" int vt_rebuild(vt_ctx* ctx, const vt_graph* g, const int64_t* cvect, vt_graph** out)
{
    if (!ctx || !g || !out) return VT_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        VT_HIP(ctx, hipSetDevice(ctx->device));
        DBuf<int64_t> dc;
        const int64_t* src = g->part;
        if (cvect) {
            dc = DBuf<int64_t>(ctx, (size_t)std::max<int64_t>(g->nv, 1));
            h2d(ctx, dc.p, cvect, 8 * (size_t)g->nv);
            src = dc.p;
        } else if (!src) {
            fail(ctx, VT_ERR_STATE, "vt_rebuild: no partition given and none left by vt_louvain_phase");
        }
        *out = rebuild(ctx, g, src);
        sync(ctx);
    });
}
