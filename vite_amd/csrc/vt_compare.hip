// Ground-truth comparison of two partitions (compare_communities, compare.cpp:58-307, and
// compute_gini_coeff, compare.cpp:310-337).
//
// The reference counts, for every pair of vertices inside one community of C1, whether C2
// also puts them together (same-same) or not (same-diff), and for every pair inside one
// community of C2 whether C1 separates them (diff-same), by walking all pairs: O(sum of
// squared community sizes).  The counts are integers, so any order gives the same values:
// here every vertex adds, with one atomic increment each, the number of vertices before it
// that share its C1 community, its C2 community, and its (C1, C2) pair, which sums to
// C(n,2) over each group.  The pair groups live in an open-addressed HBM table of 64-bit
// keys (C1 * nC2 + C2) at most half full.  Then
//   same-same = sum over (a,b) of C(n_ab, 2),
//   same-diff = sum over a of C(n_a, 2) - same-same,
//   diff-same = sum over b of C(n_b, 2) - same-same,
// exactly the reference's counts.  Precision, recall, F-score and the Gini coefficients
// (a left fold over the sorted community sizes, zero-size ids included) are formed on the
// host with the reference's expressions and order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "vt_internal.hpp"

namespace vt {
namespace {

constexpr unsigned long long kEmptyKey = ~0ull;
constexpr int kCmpBlock = 256;

// Per-block minimum and maximum of both partitions.
__global__ void __launch_bounds__(kCmpBlock) k_cmp_range(const int64_t* __restrict__ c1, const int64_t* __restrict__ c2,
                                                         int64_t n, int64_t* __restrict__ part)
{
    __shared__ int64_t s[4][kCmpBlock];
    int64_t mn1 = INT64_MAX, mx1 = INT64_MIN, mn2 = INT64_MAX, mx2 = INT64_MIN;
    for (int64_t i = blockIdx.x * (int64_t)kCmpBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kCmpBlock) {
        const int64_t a = c1[i], b = c2[i];
        mn1 = a < mn1 ? a : mn1;
        mx1 = a > mx1 ? a : mx1;
        mn2 = b < mn2 ? b : mn2;
        mx2 = b > mx2 ? b : mx2;
    }
    s[0][threadIdx.x] = mn1;
    s[1][threadIdx.x] = mx1;
    s[2][threadIdx.x] = mn2;
    s[3][threadIdx.x] = mx2;
    __syncthreads();
    for (int o = kCmpBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            const int t = threadIdx.x;
            s[0][t] = s[0][t + o] < s[0][t] ? s[0][t + o] : s[0][t];
            s[1][t] = s[1][t + o] > s[1][t] ? s[1][t + o] : s[1][t];
            s[2][t] = s[2][t + o] < s[2][t] ? s[2][t + o] : s[2][t];
            s[3][t] = s[3][t + o] > s[3][t] ? s[3][t + o] : s[3][t];
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) part[4 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void k_cmp_fill(unsigned long long* __restrict__ p, int64_t n, unsigned long long v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Every vertex: increments of its C1 size, its C2 size and its (C1, C2) pair count; the
// old values summed give the three pair totals (sums[0..2]).
__global__ void __launch_bounds__(kCmpBlock) k_cmp_count(const int64_t* __restrict__ c1, const int64_t* __restrict__ c2,
                                                         int64_t n, uint64_t nc2, unsigned long long* __restrict__ size1,
                                                         unsigned long long* __restrict__ size2,
                                                         unsigned long long* __restrict__ keys,
                                                         unsigned long long* __restrict__ cnt, int logcap,
                                                         unsigned long long* __restrict__ sums)
{
    __shared__ unsigned long long s[3][kCmpBlock];
    unsigned long long p1 = 0, p2 = 0, p12 = 0;
    const unsigned long long mask = (1ull << logcap) - 1;
    for (int64_t i = blockIdx.x * (int64_t)kCmpBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kCmpBlock) {
        const uint64_t a = (uint64_t)c1[i], b = (uint64_t)c2[i];
        p1 += atomicAdd(&size1[a], 1ull);
        p2 += atomicAdd(&size2[b], 1ull);
        const unsigned long long key = a * nc2 + b;
        unsigned long long h = (key * 0x9E3779B97F4A7C15ull) >> (64 - logcap);
        for (;;) {  // ends: at most n distinct keys in a table of at least 2n slots
            const unsigned long long old = atomicCAS(&keys[h], kEmptyKey, key);
            if (old == kEmptyKey || old == key) break;
            h = (h + 1) & mask;
        }
        p12 += atomicAdd(&cnt[h], 1ull);
    }
    s[0][threadIdx.x] = p1;
    s[1][threadIdx.x] = p2;
    s[2][threadIdx.x] = p12;
    __syncthreads();
    for (int o = kCmpBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int k = 0; k < 3; k++) s[k][threadIdx.x] += s[k][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 3) atomicAdd(&sums[threadIdx.x], s[threadIdx.x][0]);
}

// compute_gini_coeff (compare.cpp:310-337) on sizes already sorted ascending: the one-thread
// left folds of (i+1)*size[i] and size[i], then the reference's closing expression.
double gini_sorted(const std::vector<uint64_t>& sorted)
{
    const int numColors = (int)sorted.size();
    double numFunc = 0.0, denFunc = 0.0;
    for (int64_t i = 0; i < numColors; i++) {
        numFunc = numFunc + (double)((i + 1) * (int64_t)sorted[i]);
        denFunc = denFunc + (double)(int64_t)sorted[i];
    }
    return ((2 * numFunc) / (numColors * denFunc)) - ((double)(numColors + 1) / (double)numColors);
}

// Community sizes sorted ascending on the device, then read back.
std::vector<uint64_t> sorted_sizes(vt_ctx* ctx, unsigned long long* d_size, int64_t nc, int64_t n)
{
    std::vector<uint64_t> h((size_t)nc);
    DBuf<uint64_t> kalt(ctx, (size_t)nc);
    DBuf<uint32_t> v(ctx, (size_t)nc), valt(ctx, (size_t)nc);
    bool alt = false;
    sort_pairs_u64(ctx, (uint64_t*)d_size, v.p, nc, (int)bits_for(n), kalt.p, valt.p, &alt);
    d2h(ctx, h.data(), alt ? (const void*)kalt.p : (const void*)d_size, 8 * (size_t)nc);
    return h;
}

}  // namespace
}  // namespace vt

using namespace vt;

extern "C" int vt_compare_communities(vt_ctx* ctx, int64_t n, const int64_t* truth, const int64_t* found,
                                      vt_compare_stats* out)
{
    if (!ctx || !truth || !found || !out) return VT_ERR_ARG;
    return guarded(ctx, [&] {
        if (n < 1) fail(ctx, VT_ERR_ARG, "vt_compare_communities: n = %lld, need at least one vertex", (long long)n);
        VT_HIP(ctx, hipSetDevice(ctx->device));
        DBuf<int64_t> c1(ctx, (size_t)n), c2(ctx, (size_t)n);
        h2d(ctx, c1.p, truth, 8 * (size_t)n);
        h2d(ctx, c2.p, found, 8 * (size_t)n);

        // community id ranges
        const unsigned nb = grid_for(n, kCmpBlock, 1024);
        DBuf<int64_t> part(ctx, 4 * (size_t)nb);
        VT_LAUNCH(ctx, "cmp_range", k_cmp_range, nb, kCmpBlock, 0, (const int64_t*)c1.p, (const int64_t*)c2.p, n,
                  part.p);
        std::vector<int64_t> hp(4 * (size_t)nb);
        d2h(ctx, hp.data(), part.p, 8 * hp.size());
        int64_t mn1 = INT64_MAX, mx1 = INT64_MIN, mn2 = INT64_MAX, mx2 = INT64_MIN;
        for (unsigned b = 0; b < nb; b++) {
            mn1 = std::min(mn1, hp[4 * b]);
            mx1 = std::max(mx1, hp[4 * b + 1]);
            mn2 = std::min(mn2, hp[4 * b + 2]);
            mx2 = std::max(mx2, hp[4 * b + 3]);
        }
        if (mn1 < 0 || mn2 < 0)
            fail(ctx, VT_ERR_ARG, "vt_compare_communities: community ids must be >= 0 (smallest: %lld in truth, %lld in found)",
                 (long long)mn1, (long long)mn2);
        // the reference sizes its arrays by the largest id and passes the count to the Gini
        // coefficient as an int
        if (mx1 >= INT32_MAX || mx2 >= INT32_MAX)
            fail(ctx, VT_ERR_LIMIT, "vt_compare_communities: community ids must stay below 2^31 - 1");
        const int64_t nc1 = mx1 + 1, nc2 = mx2 + 1;

        int logcap = 1;
        while ((1ll << logcap) < 2 * n) logcap++;
        const int64_t cap = 1ll << logcap;
        DBuf<unsigned long long> size1(ctx, (size_t)nc1), size2(ctx, (size_t)nc2), keys(ctx, (size_t)cap),
            cnt(ctx, (size_t)cap), sums(ctx, 3);
        dzero(ctx, size1.p, 8 * (size_t)nc1);
        dzero(ctx, size2.p, 8 * (size_t)nc2);
        dzero(ctx, cnt.p, 8 * (size_t)cap);
        dzero(ctx, sums.p, 24);
        VT_LAUNCH(ctx, "cmp_fill", k_cmp_fill, grid_for(cap, 256, 8192), 256, 0, keys.p, cap, kEmptyKey);
        VT_LAUNCH(ctx, "cmp_count", k_cmp_count, grid_for(n, kCmpBlock, 8192), kCmpBlock, 0, (const int64_t*)c1.p,
                  (const int64_t*)c2.p, n, (uint64_t)nc2, size1.p, size2.p, keys.p, cnt.p, logcap, sums.p);
        unsigned long long s[3];
        d2h(ctx, s, sums.p, sizeof(s));
        const std::vector<uint64_t> z1 = sorted_sizes(ctx, size1.p, nc1, n);
        const std::vector<uint64_t> z2 = sorted_sizes(ctx, size2.p, nc2, n);
        sync(ctx);

        vt_compare_stats r;
        std::memset(&r, 0, sizeof(r));
        r.n = n;
        r.nc_truth = nc1;
        r.nc_found = nc2;
        r.same_same = (int64_t)s[2];
        r.same_diff = (int64_t)(s[0] - s[2]);
        r.diff_same = (int64_t)(s[1] - s[2]);
        // compare.cpp:267-272, term by term
        r.precision = (double)r.same_same / (double)(r.same_same + r.diff_same);
        r.recall = (double)r.same_same / (double)(r.same_same + r.same_diff);
        r.f_score = 2.0 * ((r.precision * r.recall) / (r.precision + r.recall));
        r.gini_truth = gini_sorted(z1);
        r.gini_found = gini_sorted(z2);
        *out = r;
    });
}
