// Coarsening: distbuildNextLevelGraph (rebuild.cpp:480-504) at world size 1.
//
//  1. renumber: new id of community c = rank of its first appearance in cvect as the
//     vertex index ascends (distReNumber, rebuild.cpp:121-135) -- atomicMin of the first
//     vertex per community, a flag scan over vertices, one gather;
//  2. coarse edge (C1, C2) weight = left fold over the vertices of C1 ascending, each row in
//     CSR order (fill_newEdgesMap, rebuild.cpp:294-329), the first contribution as is:
//     - internal edges (C2 == C1, most of them after a phase): one thread per community
//       walks its members (a stable sort of vertices by new id) and folds them in place;
//     - external edges: compacted in global edge order, radix-sorted stably by (C1, C2),
//       and every run folded left to right by one thread;
//  3. rows sorted by C2 (std::map order, rebuild.cpp:469-476), the self-loop entry placed at
//     its column among the external runs.  Self-loops are kept.
// Graphs with 2^31 or more edges take the older path that sorts every edge by (C1, C2).
#include <hip/hip_runtime.h>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <climits>
#include <vector>

#include "vt_internal.hpp"

namespace vt {

#ifndef VT_SPLIT_REBUILD
#define VT_SPLIT_REBUILD 1
#endif
constexpr bool kSplitRebuild = VT_SPLIT_REBUILD;

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

// ---- internal / external split -------------------------------------------------------------
// Most fine edges end inside their own community.  Their coarse self-loop weight is the
// left fold of the community's internal edge weights over its members in vertex order,
// each row in CSR order: one thread per community walks its members directly, with no sort.
// Only the external edges (C1 != C2) go through the stable (C1, C2) radix sort.  The
// self-loop entry is then placed at its sorted column in the coarse row.

__global__ void k_count_members(const uint32_t* __restrict__ newid, int64_t nv, int64_t* __restrict__ cnt)
{
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x)
        atomicAdd((unsigned long long*)&cnt[newid[v]], 1ull);
}

__global__ void k_iota_u32(uint32_t* __restrict__ p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)i;
}

__global__ void k_member_degrees(const uint32_t* __restrict__ mem, int64_t nv, const int64_t* __restrict__ offs,
                                 int64_t* __restrict__ mdeg)
{
    for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m <= nv; m += (int64_t)gridDim.x * blockDim.x)
        mdeg[m] = m < nv ? offs[mem[m] + 1] - offs[mem[m]] : 0;
}

// One wave per community walks its members' rows as one concatenated row (moff: start of
// each member's row in it), 64 entries at a time with coalesced loads.  Each lane stages
// its weight, or -0.0 when the edge leaves the community, and lane 0 adds the staged values
// in order onto a sum that starts at -0.0.  -0.0 + w == w and x + (-0.0) == x for every
// double, so the result is the left fold of the internal weights with the first one as is.
constexpr int kFoldWaves = 4;

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void __launch_bounds__(64 * kFoldWaves) k_internal_fold(
    int64_t nc, const int64_t* __restrict__ cstart, const uint32_t* __restrict__ mem, const int64_t* __restrict__ moff,
    const int64_t* __restrict__ offs, const DEdge* __restrict__ edges, const uint32_t* __restrict__ newid,
    double* __restrict__ self_w, unsigned char* __restrict__ has_self)
{
    __shared__ double st[kFoldWaves][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int64_t c = (int64_t)blockIdx.x * kFoldWaves + w; c < nc; c += (int64_t)gridDim.x * kFoldWaves) {
        const int64_t m0 = cstart[c], m1 = cstart[c + 1], nm = m1 - m0;
        const int64_t s0 = moff[m0], s1 = moff[m1];
        // up to 64 members: lane i holds member i's start in the concatenated row and its
        // row start in the CSR, and lanes find their member with shuffles
        const bool small = nm <= 64;
        int64_t my_s = INT64_MAX, my_row = 0;
        if (small && lane < nm) {
            my_s = moff[m0 + lane];
            my_row = offs[mem[m0 + lane]];
        }
        double acc = -0.0;
        bool any = false;
        for (int64_t base = s0; base < s1; base += 64) {
            const int64_t p = base + lane;
            double x = -0.0;
            bool in = false;
            int64_t k = 0;
            if (small) {
                int lo = 0;
#pragma unroll
                for (int step = 32; step > 0; step >>= 1) {
                    const int cand = lo + step;
                    const int64_t cs = __shfl(my_s, cand < 64 ? cand : 63);
                    if (cand < nm && cs <= p) lo = cand;
                }
                k = __shfl(my_row, lo) + (p - __shfl(my_s, lo));
            } else if (p < s1) {
                int64_t lo = m0, hi = m1 - 1;  // the member whose row holds position p
                while (lo < hi) {
                    const int64_t mid = (lo + hi + 1) >> 1;
                    if (moff[mid] <= p) lo = mid;
                    else hi = mid - 1;
                }
                k = offs[mem[lo]] + (p - moff[lo]);
            }
            if (p < s1) {
                const DEdge e = edges[k];
                in = newid[e.tail] == (uint32_t)c;
                if (in) x = e.weight;
            }
            any = any || (__ballot(in) != 0);
            st[w][lane] = x;
            wave_lds_sync();
            if (lane == 0) {
                const int n = (int)(s1 - base < 64 ? s1 - base : 64);
                for (int i = 0; i < n; i++) acc = acc + st[w][i];
            }
            wave_lds_sync();
        }
        if (lane == 0) {
            self_w[c] = acc;
            has_self[c] = any ? 1 : 0;
        }
    }
}

// Edge-parallel over blocks of 256 rows (row starts and new ids cached in LDS): pass 0 flags
// the external edges, pass 1 writes their (C1 << bits | C2, weight) at their compacted
// positions, keeping global edge order.
template <int PASS>
__global__ void __launch_bounds__(kRowsPerBlock) k_ext_edges(const int64_t* __restrict__ offs,
                                                             const DEdge* __restrict__ edges, int64_t nv,
                                                             const uint32_t* __restrict__ newid, int bits,
                                                             int32_t* __restrict__ flag, const int32_t* __restrict__ pos,
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
            const uint32_t c2 = newid[e.tail];
            if constexpr (PASS == 0) {
                flag[k] = c2 != sid[lo] ? 1 : 0;
            } else {
                if (c2 != sid[lo]) {
                    keys[pos[k]] = ((unsigned long long)sid[lo] << bits) | c2;
                    vals[pos[k]] = e.weight;
                }
            }
        }
        __syncthreads();
    }
}

__global__ void k_ext_runs(const unsigned long long* __restrict__ keys, const double* __restrict__ vals,
                           const int64_t* __restrict__ starts, int64_t nu, int bits, uint32_t* __restrict__ rc1,
                           uint32_t* __restrict__ rc2, double* __restrict__ rw)
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
        rc1[r] = (uint32_t)(key >> bits);
        rc2[r] = (uint32_t)(key & lomask);
        rw[r] = acc;
    }
}

// First position in a[lo, hi) whose value is >= x (a ascending there).
__device__ __forceinline__ int64_t lower_bound_u32(const uint32_t* __restrict__ a, int64_t lo, int64_t hi, uint32_t x)
{
    while (lo < hi) {
        const int64_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Per community: its first external run, how many it has, and how many sort before the
// self-loop column (runs are sorted by C1, then C2) -- binary searches, no atomics.
__global__ void k_comm_runs(int64_t nc, const uint32_t* __restrict__ rc1, const uint32_t* __restrict__ rc2, int64_t nu,
                            int64_t* __restrict__ ext_first, int64_t* __restrict__ ext_cnt, int64_t* __restrict__ lt_cnt)
{
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t f = lower_bound_u32(rc1, 0, nu, (uint32_t)c);
        const int64_t l = lower_bound_u32(rc1, f, nu, (uint32_t)c + 1);
        ext_first[c] = f;
        ext_cnt[c] = l - f;
        lt_cnt[c] = lower_bound_u32(rc2, f, l, (uint32_t)c) - f;
    }
}

__global__ void k_row_counts(int64_t nc, const int64_t* __restrict__ ext_cnt, const unsigned char* __restrict__ has_self,
                             int64_t* __restrict__ rc)
{
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c <= nc; c += (int64_t)gridDim.x * blockDim.x)
        rc[c] = c < nc ? ext_cnt[c] + has_self[c] : 0;
}

// Run r of community C1 (its runs are contiguous, in C2 order) goes to its rank among them,
// one further when the self-loop entry sorts before it; the self-loop entry goes after the
// runs with C2 < C1.
__global__ void k_place(const uint32_t* __restrict__ rc1, const uint32_t* __restrict__ rc2, const double* __restrict__ rw,
                        int64_t nu, const int64_t* __restrict__ out_offs, const int64_t* __restrict__ ext_first,
                        const unsigned char* __restrict__ has_self, DEdge* __restrict__ out)
{
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nu; r += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t c1 = rc1[r], c2 = rc2[r];
        const int64_t p = out_offs[c1] + (r - ext_first[c1]) + ((has_self[c1] && c2 > c1) ? 1 : 0);
        DEdge o;
        o.tail = (int64_t)c2;
        o.weight = rw[r];
        out[p] = o;
    }
}

__global__ void k_place_self(int64_t nc, const int64_t* __restrict__ out_offs, const int64_t* __restrict__ lt_cnt,
                             const unsigned char* __restrict__ has_self, const double* __restrict__ self_w,
                             DEdge* __restrict__ out)
{
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
        if (!has_self[c]) continue;
        DEdge o;
        o.tail = c;
        o.weight = self_w[c];
        out[out_offs[c] + lt_cnt[c]] = o;
    }
}

__global__ void k_total_i32(const int32_t* __restrict__ flag, const int32_t* __restrict__ scan, int64_t n, int64_t* out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = n ? (int64_t)scan[n - 1] + flag[n - 1] : 0;
}

// The rebuild with the internal / external split (needs ne < 2^31 for 32-bit positions).
static vt_graph* rebuild_split(vt_ctx* ctx, const vt_graph* g, const uint32_t* newid, int64_t nc)
{
    const int64_t nv = g->nv, ne = g->ne;
    const int bits = (int)bits_for(nc > 1 ? nc - 1 : 1);
    const size_t c1n = (size_t)nc + 1, n1 = (size_t)std::max<int64_t>(nv, 1), e1 = (size_t)std::max<int64_t>(ne, 1);
    DBuf<int64_t> tot(ctx, 2);

    // members of each community, ascending vertex order (stable sort by new id)
    DBuf<int64_t> ccnt(ctx, c1n), cstart(ctx, c1n);
    dzero(ctx, ccnt.p, 8 * c1n);
    VT_LAUNCH(ctx, "rb_members", k_count_members, grid_for(nv, 256, 8192), 256, 0, newid, nv, ccnt.p);
    exclusive_scan_i64(ctx, ccnt.p, cstart.p, nc + 1);
    ccnt.reset();
    DBuf<uint32_t> mk(ctx, n1), mv(ctx, n1), mk2(ctx, n1), mv2(ctx, n1);
    d2d(ctx, mk.p, newid, 4 * (size_t)nv);
    VT_LAUNCH(ctx, "rb_iota", k_iota_u32, grid_for(nv, 256, 8192), 256, 0, mv.p, nv);
    bool malt = false;
    sort_pairs_u32(ctx, mk.p, mv.p, nv, bits, mk2.p, mv2.p, &malt);
    const uint32_t* members = malt ? mv2.p : mv.p;

    // self-loop weight of every community from its internal edges
    DBuf<double> self_w(ctx, c1n);
    DBuf<unsigned char> has_self(ctx, c1n);
    {
        DBuf<int64_t> mdeg(ctx, n1 + 1), moff(ctx, n1 + 1);
        VT_LAUNCH(ctx, "rb_mdeg", k_member_degrees, grid_for(nv + 1, 256, 8192), 256, 0, members, nv,
                  (const int64_t*)g->offs, mdeg.p);
        exclusive_scan_i64(ctx, mdeg.p, moff.p, nv + 1);
        VT_LAUNCH(ctx, "rb_internal", k_internal_fold, grid_for(nc, kFoldWaves, 65536), 64 * kFoldWaves, 0, nc,
                  (const int64_t*)cstart.p, members, (const int64_t*)moff.p, (const int64_t*)g->offs,
                  (const DEdge*)g->edges, newid, self_w.p, has_self.p);
    }
    mk.reset();
    mv.reset();
    mk2.reset();
    mv2.reset();
    cstart.reset();

    // external edges, compacted in global edge order, then sorted stably by (C1, C2)
    int64_t next = 0;
    DBuf<unsigned long long> k0, k1;
    DBuf<double> v0, v1;
    {
        DBuf<int32_t> flag(ctx, e1), pos(ctx, e1);
        if (ne) {
            VT_LAUNCH(ctx, "rb_ext_flags", k_ext_edges<0>, grid_for(nv, kRowsPerBlock, 16384), kRowsPerBlock, 0, g->offs,
                      g->edges, nv, newid, bits, flag.p, (const int32_t*)nullptr, (unsigned long long*)nullptr,
                      (double*)nullptr);
            exclusive_scan_i32(ctx, flag.p, pos.p, ne);
            VT_LAUNCH(ctx, "rb_total", k_total_i32, 1, 64, 0, (const int32_t*)flag.p, (const int32_t*)pos.p, ne, tot.p);
            d2h(ctx, &next, tot.p, 8);
        }
        const size_t x1 = (size_t)std::max<int64_t>(next, 1);
        k0 = DBuf<unsigned long long>(ctx, x1);
        k1 = DBuf<unsigned long long>(ctx, x1);
        v0 = DBuf<double>(ctx, x1);
        v1 = DBuf<double>(ctx, x1);
        if (next)
            VT_LAUNCH(ctx, "rb_ext_keys", k_ext_edges<1>, grid_for(nv, kRowsPerBlock, 16384), kRowsPerBlock, 0, g->offs,
                      g->edges, nv, newid, bits, (int32_t*)nullptr, (const int32_t*)pos.p, k0.p, v0.p);
    }
    const unsigned long long* ks = k0.p;
    const double* vs = v0.p;
    if (next > 1) {
        rocprim::double_buffer<unsigned long long> kb(k0.p, k1.p);
        rocprim::double_buffer<double> vb(v0.p, v1.p);
        size_t tb = 0;
        VT_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tb, kb, vb, (size_t)next, 0, 2 * bits, ctx->stream));
        void* tmp = scratch(ctx, "sort_tmp", tb);
        {
            TimedScope ts(ctx, "rb_sort");
            VT_HIP(ctx, rocprim::radix_sort_pairs(tmp, tb, kb, vb, (size_t)next, 0, 2 * bits, ctx->stream));
        }
        ks = kb.current();
        vs = vb.current();
    }

    // runs of equal (C1, C2) folded left to right
    int64_t nu = 0;
    DBuf<int64_t> starts;
    {
        const size_t x1 = (size_t)std::max<int64_t>(next, 1);
        DBuf<int64_t> head(ctx, x1), hpos(ctx, x1);
        if (next) {
            VT_LAUNCH(ctx, "rb_heads", k_run_heads, grid_for(next, 256, 16384), 256, 0, ks, next, head.p);
            exclusive_scan_i64(ctx, head.p, hpos.p, next);
            VT_LAUNCH(ctx, "rb_total", k_total, 1, 64, 0, (const int64_t*)head.p, (const int64_t*)hpos.p, next, tot.p);
            d2h(ctx, &nu, tot.p, 8);
        }
        starts = DBuf<int64_t>(ctx, (size_t)nu + 1);
        if (next)
            VT_LAUNCH(ctx, "rb_starts", k_run_starts, grid_for(next, 256, 16384), 256, 0, (const int64_t*)head.p,
                      (const int64_t*)hpos.p, next, nu, starts.p);
    }
    DBuf<uint32_t> rc1(ctx, (size_t)std::max<int64_t>(nu, 1)), rc2(ctx, (size_t)std::max<int64_t>(nu, 1));
    DBuf<double> rw(ctx, (size_t)std::max<int64_t>(nu, 1));
    if (nu)
        VT_LAUNCH(ctx, "rb_fold", k_ext_runs, grid_for(nu, 256, 16384), 256, 0, ks, vs, (const int64_t*)starts.p, nu, bits,
                  rc1.p, rc2.p, rw.p);

    // coarse rows: the runs of C1 in C2 order with the self-loop entry at its column
    DBuf<int64_t> ext_first(ctx, c1n), ext_cnt(ctx, c1n), lt_cnt(ctx, c1n), rc(ctx, c1n), out_offs(ctx, c1n);
    VT_LAUNCH(ctx, "rb_comm_runs", k_comm_runs, grid_for(nc, 256, 16384), 256, 0, nc, (const uint32_t*)rc1.p,
              (const uint32_t*)rc2.p, nu, ext_first.p, ext_cnt.p, lt_cnt.p);
    VT_LAUNCH(ctx, "rb_rowcount", k_row_counts, grid_for(nc + 1, 256, 8192), 256, 0, nc, (const int64_t*)ext_cnt.p,
              (const unsigned char*)has_self.p, rc.p);
    exclusive_scan_i64(ctx, rc.p, out_offs.p, nc + 1);
    int64_t total = 0;
    d2h(ctx, &total, out_offs.p + nc, 8);
    vt_graph* out = graph_new(ctx, nc, total);
    try {
        d2d(ctx, out->offs, out_offs.p, 8 * c1n);
        if (nu)
            VT_LAUNCH(ctx, "rb_place", k_place, grid_for(nu, 256, 16384), 256, 0, (const uint32_t*)rc1.p,
                      (const uint32_t*)rc2.p, (const double*)rw.p, nu, (const int64_t*)out_offs.p,
                      (const int64_t*)ext_first.p, (const unsigned char*)has_self.p, out->edges);
        VT_LAUNCH(ctx, "rb_place_self", k_place_self, grid_for(nc, 256, 8192), 256, 0, nc, (const int64_t*)out_offs.p,
                  (const int64_t*)lt_cnt.p, (const unsigned char*)has_self.p, (const double*)self_w.p, out->edges);
        graph_finalize(ctx, out);
    } catch (...) {
        graph_delete(ctx, out);
        throw;
    }
    return out;
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
    if (kSplitRebuild && ne < (int64_t(1) << 31)) return rebuild_split(ctx, g, newid.p, nc);

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
