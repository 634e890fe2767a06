// Device graph generators.
//
// Random geometric graph (the workload graphs): n points from a counter-based stream in
// [0,1]^2, radius r = sqrt(2 ln n / (pi n)), cells of side 1/floor(1/r) >= r.  Vertices are
// numbered in (cell row-major, point index) order, so walking the 3x3 cell neighbourhood
// of a vertex yields its row already sorted by tail.  Edge iff sqrt(dx*dx + dy*dy) <= r
// (no FMA, IEEE sqrt), weight = that distance; both directions see the same operands up to
// sign, so the weight is symmetric.  Optional extra edges: nx = (int64)(pct*m)/100 (the
// rounding of distgraph.cpp:686), candidate t = (rng % n, rng % n) with weight U(0,1),
// skipping self pairs, RGG edges and repeats of an earlier candidate; the first nx
// survivors in t order are kept.
//
// Barabasi-Albert: the Batagelj-Brandes edge-copy rule with the recursion resolved per
// edge in parallel: edge e = v*m + j points at M[r], r = rng % (2e+1), over the implicit
// array M[2e] = e/m, M[2e+1] = target(e).  Self-loops and repeated pairs are dropped,
// weights are 1.
//
// Both are bit-identical to the CPU restatement used as the test oracle.
#include <hip/hip_runtime.h>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "vt_internal.hpp"

namespace vt {

__host__ __device__ inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

constexpr uint64_t kGolden = 0x9e3779b97f4a7c15ull;

__host__ __device__ inline uint64_t rng_key(uint64_t seed, uint64_t stream) { return mix64(seed + kGolden * (stream + 1)); }
__host__ __device__ inline uint64_t rng_at(uint64_t key, uint64_t k) { return mix64(key + kGolden * (k + 1)); }
__host__ __device__ inline double u01(uint64_t z) { return (double)(z >> 11) * 0x1.0p-53; }

// ---- RGG ----------------------------------------------------------------------------------

__global__ void k_rgg_points(uint64_t key, int64_t n, int64_t nc, uint32_t* __restrict__ cell, uint32_t* __restrict__ idx)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const double x = u01(rng_at(key, 2 * (uint64_t)k));
        const double y = u01(rng_at(key, 2 * (uint64_t)k + 1));
        int64_t cx = (int64_t)(x * (double)nc); if (cx >= nc) cx = nc - 1;
        int64_t cy = (int64_t)(y * (double)nc); if (cy >= nc) cy = nc - 1;
        cell[k] = (uint32_t)(cy * nc + cx);
        idx[k] = (uint32_t)k;
    }
}

__global__ void k_rgg_gather(uint64_t key, int64_t n, const uint32_t* __restrict__ perm, double2* __restrict__ xy,
                             unsigned long long* __restrict__ cellcount, const uint32_t* __restrict__ cell)
{
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = perm[v];
        xy[v] = make_double2(u01(rng_at(key, 2 * k)), u01(rng_at(key, 2 * k + 1)));
        atomicAdd(&cellcount[cell[v]], 1ull);
    }
}

// pass 0 counts, pass 1 writes; the walk order is the same in both.
template <int PASS>
__global__ void k_rgg_edges(int64_t n, int64_t nc, double r, const double2* __restrict__ xy,
                            const uint32_t* __restrict__ cell, const int64_t* __restrict__ cstart,
                            int64_t* __restrict__ deg_or_offs, DEdge* __restrict__ edges)
{
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const double2 p = xy[v];
        const int64_t c = cell[v];
        const int64_t cy = c / nc, cx = c % nc;
        int64_t o = PASS ? deg_or_offs[v] : 0;
        for (int64_t yy = cy - 1; yy <= cy + 1; yy++) {
            if (yy < 0 || yy >= nc) continue;
            for (int64_t xx = cx - 1; xx <= cx + 1; xx++) {
                if (xx < 0 || xx >= nc) continue;
                const int64_t cc = yy * nc + xx;
                for (int64_t u = cstart[cc]; u < cstart[cc + 1]; u++) {
                    if (u == v) continue;
                    const double2 q = xy[u];
                    const double dx = p.x - q.x;
                    const double dy = p.y - q.y;
                    const double d = sqrt(dx * dx + dy * dy);
                    if (d <= r) {
                        if (PASS) { DEdge e; e.tail = u; e.weight = d; edges[o] = e; }
                        o++;
                    }
                }
            }
        }
        if (!PASS) deg_or_offs[v] = o;
    }
}

__device__ inline bool csr_has(const int64_t* offs, const DEdge* edges, int64_t a, int64_t b)
{
    int64_t lo = offs[a], hi = offs[a + 1];
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (edges[mid].tail < b) lo = mid + 1; else hi = mid;
    }
    return lo < offs[a + 1] && edges[lo].tail == b;
}

// Candidate t -> key (a << bits | b) or ~0 when it cannot survive.
__global__ void k_extra_cand(uint64_t kpair, int64_t n, int bits, int64_t T, const int64_t* __restrict__ offs,
                             const DEdge* __restrict__ edges, unsigned long long* __restrict__ keys,
                             uint32_t* __restrict__ tval)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < T; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t a0 = (int64_t)(rng_at(kpair, 2 * (uint64_t)t) % (uint64_t)n);
        const int64_t b0 = (int64_t)(rng_at(kpair, 2 * (uint64_t)t + 1) % (uint64_t)n);
        unsigned long long key = ~0ull;
        if (a0 != b0) {
            const int64_t a = a0 < b0 ? a0 : b0, b = a0 < b0 ? b0 : a0;
            if (!csr_has(offs, edges, a, b)) key = ((unsigned long long)a << bits) | (unsigned long long)b;
        }
        keys[t] = key;
        tval[t] = (uint32_t)t;
    }
}

// Marks the first (lowest t) candidate of every distinct valid pair.
__global__ void k_extra_first(const unsigned long long* __restrict__ keys, const uint32_t* __restrict__ tval, int64_t T,
                              int64_t* __restrict__ surv)
{
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < T; j += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long k = keys[j];
        surv[tval[j]] = (k != ~0ull && (j == 0 || keys[j - 1] != k)) ? 1 : 0;
    }
}

// Survivors with rank < nx, emitted as two directed keys each, value = weight index t.
__global__ void k_extra_emit(uint64_t kpair, uint64_t kw, int64_t n, int bits, int64_t T, int64_t nx,
                             const int64_t* __restrict__ surv, const int64_t* __restrict__ rank,
                             unsigned long long* __restrict__ dkeys, double* __restrict__ dw)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < T; t += (int64_t)gridDim.x * blockDim.x) {
        if (!surv[t] || rank[t] >= nx) continue;
        const int64_t a0 = (int64_t)(rng_at(kpair, 2 * (uint64_t)t) % (uint64_t)n);
        const int64_t b0 = (int64_t)(rng_at(kpair, 2 * (uint64_t)t + 1) % (uint64_t)n);
        const int64_t a = a0 < b0 ? a0 : b0, b = a0 < b0 ? b0 : a0;
        const double w = u01(rng_at(kw, (uint64_t)t));
        const int64_t q = 2 * rank[t];
        dkeys[q] = ((unsigned long long)a << bits) | (unsigned long long)b;
        dw[q] = w;
        dkeys[q + 1] = ((unsigned long long)b << bits) | (unsigned long long)a;
        dw[q + 1] = w;
    }
}

__global__ void k_key_src_count(const unsigned long long* __restrict__ keys, int64_t n, int bits,
                                unsigned long long* __restrict__ cnt)
{
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&cnt[keys[j] >> bits], 1ull);
}

__global__ void k_add_i64(const int64_t* a, const unsigned long long* b, int64_t n, int64_t* out)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = a[i] + (int64_t)b[i];
}

// Merges each vertex's sorted RGG row with its sorted extra row.
__global__ void k_merge_rows(int64_t n, const int64_t* __restrict__ ro, const DEdge* __restrict__ re,
                             const int64_t* __restrict__ xo, const unsigned long long* __restrict__ xk,
                             const double* __restrict__ xw, int bits, const int64_t* __restrict__ oo,
                             DEdge* __restrict__ out)
{
    const unsigned long long lomask = (1ull << bits) - 1;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        int64_t p = ro[v], pe = ro[v + 1], q = xo[v], qe = xo[v + 1], o = oo[v];
        while (p < pe || q < qe) {
            if (q >= qe || (p < pe && re[p].tail < (int64_t)(xk[q] & lomask))) out[o++] = re[p++];
            else { DEdge e; e.tail = (int64_t)(xk[q] & lomask); e.weight = xw[q]; out[o++] = e; q++; }
        }
    }
}

static void sort_u64_u32(vt_ctx* ctx, unsigned long long* k, uint32_t* v, unsigned long long* k2, uint32_t* v2, int64_t n,
                         int bits, const unsigned long long** ko, const uint32_t** vo)
{
    *ko = k;
    *vo = v;
    if (n <= 1) return;
    rocprim::double_buffer<unsigned long long> kb(k, k2);
    rocprim::double_buffer<uint32_t> vb(v, v2);
    size_t tb = 0;
    VT_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tb, kb, vb, (size_t)n, 0, bits, ctx->stream));
    void* tmp = scratch(ctx, "sort_tmp", tb);
    VT_HIP(ctx, rocprim::radix_sort_pairs(tmp, tb, kb, vb, (size_t)n, 0, bits, ctx->stream));
    *ko = kb.current();
    *vo = vb.current();
}

static void sort_u64_f64(vt_ctx* ctx, unsigned long long* k, double* v, unsigned long long* k2, double* v2, int64_t n,
                         int bits, const unsigned long long** ko, const double** vo)
{
    *ko = k;
    *vo = v;
    if (n <= 1) return;
    rocprim::double_buffer<unsigned long long> kb(k, k2);
    rocprim::double_buffer<double> vb(v, v2);
    size_t tb = 0;
    VT_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tb, kb, vb, (size_t)n, 0, bits, ctx->stream));
    void* tmp = scratch(ctx, "sort_tmp", tb);
    VT_HIP(ctx, rocprim::radix_sort_pairs(tmp, tb, kb, vb, (size_t)n, 0, bits, ctx->stream));
    *ko = kb.current();
    *vo = vb.current();
}

double rgg_radius(int64_t n)
{
    if (n < 2) return 0.0;
    return std::sqrt(2.0 * std::log((double)n) / (M_PI * (double)n));
}

vt_graph* gen_rgg(vt_ctx* ctx, uint64_t seed, int64_t n, double extra_pct)
{
    if (n < 1 || n >= (int64_t(1) << 31) - 1) fail(ctx, VT_ERR_ARG, "vt_gen_rgg: n must be in [1, 2^31-1)");
    if (extra_pct < 0.0 || extra_pct > 100.0) fail(ctx, VT_ERR_ARG, "vt_gen_rgg: extra_pct must be in [0, 100]");
    const double r = rgg_radius(n);
    int64_t nc = (r > 0.0) ? (int64_t)std::floor(1.0 / r) : 1;
    nc = std::max<int64_t>(1, std::min<int64_t>(nc, 32768));
    const int64_t ncell = nc * nc;
    const uint64_t kpts = rng_key(seed, 1);
    const size_t n1 = (size_t)n;

    DBuf<uint32_t> c0(ctx, n1), c1(ctx, n1), i0(ctx, n1), i1(ctx, n1);
    VT_LAUNCH(ctx, "rgg_points", k_rgg_points, grid_for(n, 256, 16384), 256, 0, kpts, n, nc, c0.p, i0.p);
    bool alt = false;
    sort_pairs_u32(ctx, c0.p, i0.p, n, (int)bits_for(ncell), c1.p, i1.p, &alt);
    const uint32_t* cell = alt ? c1.p : c0.p;
    const uint32_t* perm = alt ? i1.p : i0.p;
    DBuf<double2> xy(ctx, n1);
    DBuf<unsigned long long> ccount(ctx, (size_t)ncell + 1);
    DBuf<int64_t> cstart(ctx, (size_t)ncell + 1);
    dzero(ctx, ccount.p, 8 * ((size_t)ncell + 1));
    VT_LAUNCH(ctx, "rgg_gather", k_rgg_gather, grid_for(n, 256, 16384), 256, 0, kpts, n, perm, xy.p, ccount.p, cell);
    exclusive_scan_i64(ctx, (const int64_t*)ccount.p, cstart.p, ncell + 1);
    ccount.reset();

    DBuf<int64_t> deg(ctx, n1 + 1), offs(ctx, n1 + 1), tot(ctx, 2);
    dzero(ctx, deg.p, 8 * (n1 + 1));
    VT_LAUNCH(ctx, "rgg_count", k_rgg_edges<0>, grid_for(n, 256, 16384), 256, 0, n, nc, r, (const double2*)xy.p, cell,
              (const int64_t*)cstart.p, deg.p, (DEdge*)nullptr);
    exclusive_scan_i64(ctx, deg.p, offs.p, n + 1);
    int64_t ne = 0;
    d2h(ctx, &ne, offs.p + n, 8);

    vt_graph* g = graph_new(ctx, n, ne);
    try {
        d2d(ctx, g->offs, offs.p, 8 * (n1 + 1));
        VT_LAUNCH(ctx, "rgg_fill", k_rgg_edges<1>, grid_for(n, 256, 16384), 256, 0, n, nc, r, (const double2*)xy.p, cell,
                  (const int64_t*)cstart.p, g->offs, g->edges);
    } catch (...) {
        graph_delete(ctx, g);
        throw;
    }
    xy.reset();
    c0.reset(); c1.reset(); i0.reset(); i1.reset();
    cstart.reset();

    if (extra_pct > 0.0 && n > 1) {
        const int64_t m = ne / 2;
        const int64_t nx = ((int64_t)(extra_pct * (double)m)) / 100;
        if (nx > 0) {
            const int bits = (int)bits_for(n - 1);
            const uint64_t kpair = rng_key(seed, 2), kw = rng_key(seed, 3);
            int64_t T = nx + nx / 4 + 1024;
            int64_t have = 0;
            DBuf<int64_t> surv, rank;
            for (;;) {
                if (T >= (int64_t(1) << 32)) fail(ctx, VT_ERR_LIMIT, "vt_gen_rgg: too many extra-edge candidates");
                DBuf<unsigned long long> k0(ctx, (size_t)T), k1(ctx, (size_t)T);
                DBuf<uint32_t> t0(ctx, (size_t)T), t1(ctx, (size_t)T);
                surv = DBuf<int64_t>(ctx, (size_t)T);
                rank = DBuf<int64_t>(ctx, (size_t)T);
                VT_LAUNCH(ctx, "rgg_extra_cand", k_extra_cand, grid_for(T, 256, 16384), 256, 0, kpair, n, bits, T, g->offs,
                          g->edges, k0.p, t0.p);
                const unsigned long long* ks;
                const uint32_t* ts;
                sort_u64_u32(ctx, k0.p, t0.p, k1.p, t1.p, T, 2 * bits, &ks, &ts);
                VT_LAUNCH(ctx, "rgg_extra_first", k_extra_first, grid_for(T, 256, 16384), 256, 0, ks, ts, T, surv.p);
                exclusive_scan_i64(ctx, surv.p, rank.p, T);
                int64_t last[2];
                d2h(ctx, &last[0], rank.p + (T - 1), 8);
                d2h(ctx, &last[1], surv.p + (T - 1), 8);
                have = last[0] + last[1];
                if (have >= nx) break;
                T *= 2;
            }
            const int64_t nd = 2 * nx;
            DBuf<unsigned long long> dk0(ctx, (size_t)nd), dk1(ctx, (size_t)nd);
            DBuf<double> dw0(ctx, (size_t)nd), dw1(ctx, (size_t)nd);
            VT_LAUNCH(ctx, "rgg_extra_emit", k_extra_emit, grid_for(T, 256, 16384), 256, 0, kpair, kw, n, bits, T, nx,
                      (const int64_t*)surv.p, (const int64_t*)rank.p, dk0.p, dw0.p);
            const unsigned long long* xk;
            const double* xw;
            sort_u64_f64(ctx, dk0.p, dw0.p, dk1.p, dw1.p, nd, 2 * bits, &xk, &xw);
            DBuf<unsigned long long> xc(ctx, n1 + 1);
            DBuf<int64_t> xo(ctx, n1 + 1), deg2(ctx, n1 + 1), oo(ctx, n1 + 1);
            dzero(ctx, xc.p, 8 * (n1 + 1));
            VT_LAUNCH(ctx, "rgg_extra_count", k_key_src_count, grid_for(nd, 256, 16384), 256, 0, xk, nd, bits, xc.p);
            exclusive_scan_i64(ctx, (const int64_t*)xc.p, xo.p, n + 1);
            // row degrees = RGG degree + extra degree
            DBuf<int64_t> rdeg(ctx, n1 + 1);
            VT_LAUNCH(ctx, "rgg_deg", k_add_i64, grid_for(n + 1, 256, 16384), 256, 0, (const int64_t*)deg.p,
                      (const unsigned long long*)xc.p, n + 1, rdeg.p);
            exclusive_scan_i64(ctx, rdeg.p, oo.p, n + 1);
            vt_graph* g2 = graph_new(ctx, n, ne + nd);
            try {
                d2d(ctx, g2->offs, oo.p, 8 * (n1 + 1));
                VT_LAUNCH(ctx, "rgg_merge", k_merge_rows, grid_for(n, 256, 16384), 256, 0, n, (const int64_t*)g->offs,
                          (const DEdge*)g->edges, (const int64_t*)xo.p, xk, xw, bits, (const int64_t*)g2->offs, g2->edges);
                sync(ctx);
            } catch (...) {
                graph_delete(ctx, g2);
                graph_delete(ctx, g);
                throw;
            }
            graph_delete(ctx, g);
            g = g2;
        }
    }
    graph_finalize(ctx, g);
    return g;
}

// ---- BA -----------------------------------------------------------------------------------

__global__ void k_ba_pairs(uint64_t key, int64_t n, int64_t m, int bits, unsigned long long* __restrict__ keys)
{
    const int64_t ne = n * m;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t src = e / m;
        uint64_t x = (uint64_t)e;
        uint64_t r = rng_at(key, x) % (2 * x + 1);
        while (r & 1) {
            x = (r - 1) / 2;
            r = rng_at(key, x) % (2 * x + 1);
        }
        const int64_t tgt = (int64_t)(r / 2) / m;
        if (src == tgt) {
            keys[2 * e] = ~0ull;
            keys[2 * e + 1] = ~0ull;
        } else {
            keys[2 * e] = ((unsigned long long)src << bits) | (unsigned long long)tgt;
            keys[2 * e + 1] = ((unsigned long long)tgt << bits) | (unsigned long long)src;
        }
    }
}

__global__ void k_unique_heads(const unsigned long long* __restrict__ keys, int64_t n, int64_t* __restrict__ head)
{
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
        head[j] = (keys[j] != ~0ull && (j == 0 || keys[j - 1] != keys[j])) ? 1 : 0;
}

__global__ void k_unique_write(const unsigned long long* __restrict__ keys, const int64_t* __restrict__ head,
                               const int64_t* __restrict__ pos, int64_t n, int bits, DEdge* __restrict__ out,
                               unsigned long long* __restrict__ cnt)
{
    const unsigned long long lomask = (1ull << bits) - 1;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        if (!head[j]) continue;
        DEdge e;
        e.tail = (int64_t)(keys[j] & lomask);
        e.weight = 1.0;
        out[pos[j]] = e;
        atomicAdd(&cnt[keys[j] >> bits], 1ull);
    }
}

vt_graph* gen_ba(vt_ctx* ctx, uint64_t seed, int64_t n, int64_t m)
{
    if (n < 1 || m < 1 || n >= (int64_t(1) << 31) - 1) fail(ctx, VT_ERR_ARG, "vt_gen_ba: need n >= 1, m >= 1");
    if (n * m >= (int64_t(1) << 33)) fail(ctx, VT_ERR_LIMIT, "vt_gen_ba: n*m too large");
    const int bits = (int)bits_for(n > 1 ? n - 1 : 1);
    const int64_t nk = 2 * n * m;
    DBuf<unsigned long long> k0(ctx, (size_t)nk), k1(ctx, (size_t)nk);
    VT_LAUNCH(ctx, "ba_pairs", k_ba_pairs, grid_for(n * m, 256, 16384), 256, 0, rng_key(seed, 4), n, m, bits, k0.p);
    const unsigned long long* ks = k0.p;
    if (nk > 1) {
        rocprim::double_buffer<unsigned long long> kb(k0.p, k1.p);
        size_t tb = 0;
        // all 64 bits: the "dropped" marker ~0 sorts last
        VT_HIP(ctx, rocprim::radix_sort_keys(nullptr, tb, kb, (size_t)nk, 0, 64, ctx->stream));
        void* tmp = scratch(ctx, "sort_tmp", tb);
        VT_HIP(ctx, rocprim::radix_sort_keys(tmp, tb, kb, (size_t)nk, 0, 64, ctx->stream));
        ks = kb.current();
    }
    DBuf<int64_t> head(ctx, (size_t)nk), pos(ctx, (size_t)nk), tot(ctx, 2);
    VT_LAUNCH(ctx, "ba_heads", k_unique_heads, grid_for(nk, 256, 16384), 256, 0, ks, nk, head.p);
    exclusive_scan_i64(ctx, head.p, pos.p, nk);
    int64_t last[2];
    d2h(ctx, &last[0], pos.p + (nk - 1), 8);
    d2h(ctx, &last[1], head.p + (nk - 1), 8);
    const int64_t nu = last[0] + last[1];
    vt_graph* g = graph_new(ctx, n, nu);
    try {
        DBuf<unsigned long long> cnt(ctx, (size_t)n + 1);
        dzero(ctx, cnt.p, 8 * ((size_t)n + 1));
        VT_LAUNCH(ctx, "ba_write", k_unique_write, grid_for(nk, 256, 16384), 256, 0, ks, (const int64_t*)head.p,
                  (const int64_t*)pos.p, nk, bits, g->edges, cnt.p);
        exclusive_scan_i64(ctx, (const int64_t*)cnt.p, g->offs, n + 1);
        graph_finalize(ctx, g);
    } catch (...) {
        graph_delete(ctx, g);
        throw;
    }
    return g;
}

// ---- Vite's LCG RGG (generateRGG at one rank) ----------------------------------------------

// x0 * 16807^i mod (2^31 - 1): the i-th Lehmer value, by binary exponentiation (products of
// two values below 2^31 fit in 64 bits).
__device__ inline uint64_t lehmer_at(uint64_t x0, uint64_t i)
{
    const uint64_t p = 2147483647ull;
    uint64_t base = 16807ull, acc = 1ull;
    while (i) {
        if (i & 1) acc = (acc * base) % p;
        base = (base * base) % p;
        i >>= 1;
    }
    return (x0 * acc) % p;
}

// drand[i] = |x_i| * mult (utils.hpp:296-297); X = drand[0..n), Y[j] = 0 + 1.0 * drand[n + j]
// (utils.hpp:319-320 at one rank).
__global__ void k_vite_points(uint64_t x0, int64_t n, double mult, int64_t nc, double* __restrict__ X,
                              double* __restrict__ Y, uint32_t* __restrict__ cell, uint32_t* __restrict__ idx)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double x = (double)((double)fabs((double)lehmer_at(x0, (uint64_t)i)) * mult);
        const double y = 0.0 + (double)(1.0 * (double)((double)fabs((double)lehmer_at(x0, (uint64_t)(n + i))) * mult));
        X[i] = x;
        Y[i] = y;
        int64_t cx = (int64_t)(x * (double)nc); if (cx >= nc) cx = nc - 1;
        int64_t cy = (int64_t)(y * (double)nc); if (cy >= nc) cy = nc - 1;
        cell[i] = (uint32_t)(cy * nc + cx);
        idx[i] = (uint32_t)i;
    }
}

// Pair (i, j) is an edge iff sqrt(dx*dx + dy*dy) <= rn with dx = X[lo] - X[hi]
// (distgraph.cpp:536-541 computes it once per pair with i < j).  Pass 1 writes the row and
// sorts it by tail (distgraph.cpp:916-924).
template <int PASS>
__global__ void k_vite_edges(int64_t n, int64_t nc, double rn, const double* __restrict__ X, const double* __restrict__ Y,
                             const uint32_t* __restrict__ cell_of, const uint32_t* __restrict__ sorted_ids,
                             const int64_t* __restrict__ cstart, int64_t* __restrict__ deg_or_offs,
                             DEdge* __restrict__ edges)
{
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = cell_of[v];
        const int64_t cy = c / nc, cx = c % nc;
        const int64_t o0 = PASS ? deg_or_offs[v] : 0;
        int64_t o = o0;
        for (int64_t yy = cy - 1; yy <= cy + 1; yy++) {
            if (yy < 0 || yy >= nc) continue;
            for (int64_t xx = cx - 1; xx <= cx + 1; xx++) {
                if (xx < 0 || xx >= nc) continue;
                const int64_t cc = yy * nc + xx;
                for (int64_t k = cstart[cc]; k < cstart[cc + 1]; k++) {
                    const int64_t u = sorted_ids[k];
                    if (u == v) continue;
                    const int64_t lo = u < v ? u : v, hi = u < v ? v : u;
                    const double dx = X[lo] - X[hi];
                    const double dy = Y[lo] - Y[hi];
                    const double ed = sqrt(dx * dx + dy * dy);
                    if (ed <= rn) {
                        if (PASS) { DEdge e; e.tail = u; e.weight = ed; edges[o] = e; }
                        o++;
                    }
                }
            }
        }
        if (!PASS) deg_or_offs[v] = o;
        else {
            for (int64_t a = o0 + 1; a < o; a++) {
                const DEdge e = edges[a];
                int64_t b = a - 1;
                while (b >= o0 && edges[b].tail > e.tail) { edges[b + 1] = edges[b]; b--; }
                edges[b + 1] = e;
            }
        }
    }
}

__global__ void k_cell_count(const uint32_t* __restrict__ cell, int64_t n, unsigned long long* __restrict__ cnt)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&cnt[cell[i]], 1ull);
}

vt_graph* gen_vite_rgg(vt_ctx* ctx, int64_t n, uint32_t x0)
{
    if (n < 1 || n >= (int64_t(1) << 30)) fail(ctx, VT_ERR_ARG, "vt_gen_vite_rgg: n must be in [1, 2^30)");
    const double PI_V = 3.14159;  // distgraph.hpp:67
    const double rc = std::sqrt((double)std::log((double)n) / (double)(PI_V * (double)n));
    const double rt = std::sqrt((double)2.0736 / (double)n);
    const double rn = (rc + rt) / (double)2.0;
    const double mult = 1.0 / (double)(1.0 + (double)(2147483647 - 1));
    int64_t nc = (rn > 0.0) ? (int64_t)std::floor(1.0 / rn) : 1;
    nc = std::max<int64_t>(1, std::min<int64_t>(nc, 32768));
    const int64_t ncell = nc * nc;
    const size_t n1 = (size_t)n;
    DBuf<double> X(ctx, n1), Y(ctx, n1);
    DBuf<uint32_t> c0(ctx, n1), c1(ctx, n1), i0(ctx, n1), i1(ctx, n1);
    VT_LAUNCH(ctx, "vrgg_points", k_vite_points, grid_for(n, 256, 16384), 256, 0, (uint64_t)x0, n, mult, nc, X.p, Y.p,
              c0.p, i0.p);
    DBuf<uint32_t> cell_of(ctx, n1);
    d2d(ctx, cell_of.p, c0.p, 4 * n1);
    bool alt = false;
    sort_pairs_u32(ctx, c0.p, i0.p, n, (int)bits_for(ncell), c1.p, i1.p, &alt);
    const uint32_t* sorted_ids = alt ? i1.p : i0.p;
    DBuf<unsigned long long> ccount(ctx, (size_t)ncell + 1);
    DBuf<int64_t> cstart(ctx, (size_t)ncell + 1);
    dzero(ctx, ccount.p, 8 * ((size_t)ncell + 1));
    VT_LAUNCH(ctx, "vrgg_cells", k_cell_count, grid_for(n, 256, 16384), 256, 0, (const uint32_t*)cell_of.p, n, ccount.p);
    exclusive_scan_i64(ctx, (const int64_t*)ccount.p, cstart.p, ncell + 1);
    DBuf<int64_t> deg(ctx, n1 + 1), offs(ctx, n1 + 1);
    dzero(ctx, deg.p, 8 * (n1 + 1));
    VT_LAUNCH(ctx, "vrgg_count", k_vite_edges<0>, grid_for(n, 256, 16384), 256, 0, n, nc, rn, (const double*)X.p,
              (const double*)Y.p, (const uint32_t*)cell_of.p, sorted_ids, (const int64_t*)cstart.p, deg.p, (DEdge*)nullptr);
    exclusive_scan_i64(ctx, deg.p, offs.p, n + 1);
    int64_t ne = 0;
    d2h(ctx, &ne, offs.p + n, 8);
    vt_graph* g = graph_new(ctx, n, ne);
    try {
        d2d(ctx, g->offs, offs.p, 8 * (n1 + 1));
        VT_LAUNCH(ctx, "vrgg_fill", k_vite_edges<1>, grid_for(n, 256, 16384), 256, 0, n, nc, rn, (const double*)X.p,
                  (const double*)Y.p, (const uint32_t*)cell_of.p, sorted_ids, (const int64_t*)cstart.p, g->offs, g->edges);
        graph_finalize(ctx, g);
    } catch (...) {
        graph_delete(ctx, g);
        throw;
    }
    return g;
}

}  // namespace vt

using namespace vt;

extern "C" {

int vt_gen_vite_rgg(vt_ctx* ctx, int64_t n, uint32_t x0, vt_graph** out)
{
    if (!ctx || !out) return VT_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        VT_HIP(ctx, hipSetDevice(ctx->device));
        *out = gen_vite_rgg(ctx, n, x0);
        sync(ctx);
    });
}

int vt_gen_rgg(vt_ctx* ctx, uint64_t seed, int64_t n, double extra_pct, vt_graph** out)
{
    if (!ctx || !out) return VT_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        VT_HIP(ctx, hipSetDevice(ctx->device));
        *out = gen_rgg(ctx, seed, n, extra_pct);
        sync(ctx);
    });
}

int vt_gen_ba(vt_ctx* ctx, uint64_t seed, int64_t n, int64_t m, vt_graph** out)
{
    if (!ctx || !out) return VT_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        VT_HIP(ctx, hipSetDevice(ctx->device));
        *out = gen_ba(ctx, seed, n, m);
        sync(ctx);
    });
}

}  // extern "C"
