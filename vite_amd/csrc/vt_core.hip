// Context, memory, timing, graph transfer and device primitives for vite_amd.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cstdarg>
#include <mutex>
#include <cstring>
#include <vector>

#include "vt_internal.hpp"

namespace vt {

void fail(vt_ctx* ctx, int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    throw Fail{code};
}

void hip_check(vt_ctx* ctx, hipError_t e, const char* what, const char* file, int line)
{
    if (e != hipSuccess) fail(ctx, VT_ERR_HIP, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
}

// Device memory: stream-ordered blocks (hipMallocAsync on the context stream) behind a
// per-context cache.  A freed block is kept and handed to a later request of up to 4x less
// (all work of a context is ordered on its one stream, so the reuse is safe); freeing a
// large block back to the driver costs a quarter of a millisecond, and the phase loop frees
// and reallocates its buffers level after level.  The cache is trimmed when the budget or
// the device runs short, and released with the scratch buffers.
void trim_cache(vt_ctx* ctx)
{
    for (auto& kv : ctx->cache) (void)hipFreeAsync(kv.second, ctx->stream);
    ctx->cache.clear();
    ctx->cached = 0;
}

void* dalloc(vt_ctx* ctx, size_t bytes)
{
    if (ctx->budget && ctx->in_use + ctx->cached + bytes > ctx->budget) trim_cache(ctx);
    if (ctx->budget && ctx->in_use + bytes > ctx->budget)
        fail(ctx, VT_ERR_BUDGET, "allocation of %zu bytes exceeds the HBM budget (%llu in use of %llu)",
             bytes, (unsigned long long)ctx->in_use, (unsigned long long)ctx->budget);
    void* p = nullptr;
    size_t size = bytes;
    auto it = ctx->cache.lower_bound(bytes);
    if (it != ctx->cache.end() && it->first <= std::max<size_t>(4 * bytes, bytes + (1u << 20)) &&
        (!ctx->budget || ctx->in_use + it->first <= ctx->budget)) {
        size = it->first;
        p = it->second;
        ctx->cache.erase(it);
        ctx->cached -= size;
    } else {
        hipError_t e = hipMallocAsync(&p, bytes, ctx->stream);
        if (e != hipSuccess && !ctx->cache.empty()) {
            trim_cache(ctx);
            (void)hipStreamSynchronize(ctx->stream);
            e = hipMallocAsync(&p, bytes, ctx->stream);
        }
        if (e != hipSuccess) fail(ctx, VT_ERR_HIP, "hipMallocAsync(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    ctx->blocks[p] = size;
    ctx->in_use += size;
    if (ctx->in_use + ctx->cached > ctx->peak) ctx->peak = ctx->in_use + ctx->cached;
    return p;
}

void dfree(vt_ctx* ctx, void* p, size_t)
{
    if (!p) return;
    auto it = ctx->blocks.find(p);
    if (it == ctx->blocks.end()) return;
    const size_t size = it->second;
    ctx->blocks.erase(it);
    ctx->in_use -= std::min<uint64_t>(ctx->in_use, size);
    ctx->cache.emplace(size, p);
    ctx->cached += size;
}

void* scratch(vt_ctx* ctx, const std::string& key, size_t bytes)
{
    auto it = ctx->scratch.find(key);
    if (it != ctx->scratch.end()) {
        if (it->second.second >= bytes) return it->second.first;
        sync(ctx);
        dfree(ctx, it->second.first, it->second.second);
        ctx->scratch.erase(it);
    }
    size_t want = bytes < 256 ? 256 : bytes;
    void* p = dalloc(ctx, want);
    ctx->scratch[key] = {p, want};
    return p;
}

void release_scratch(vt_ctx* ctx)
{
    for (auto& kv : ctx->scratch) dfree(ctx, kv.second.first, kv.second.second);
    ctx->scratch.clear();
    trim_cache(ctx);
}

TimedScope::TimedScope(vt_ctx* c, const char* n) : ctx(c), name(n)
{
    if (!ctx->profiling) return;
    auto take = [&]() {
        hipEvent_t e;
        if (!ctx->event_pool.empty()) { e = ctx->event_pool.back(); ctx->event_pool.pop_back(); }
        else VT_HIP(ctx, hipEventCreate(&e));
        return e;
    };
    e0 = take();
    e1 = take();
    VT_HIP(ctx, hipEventRecord(e0, ctx->stream));
}

TimedScope::~TimedScope()
{
    if (!e0) return;
    (void)hipEventRecord(e1, ctx->stream);
    ctx->pending.push_back({name, e0, e1});
}

void flush_timings(vt_ctx* ctx)
{
    if (ctx->pending.empty()) return;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& p : ctx->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.e0, p.e1) == hipSuccess) {
            auto& s = ctx->stats[p.name];
            s.count++;
            s.total_ms += ms;
        }
        ctx->event_pool.push_back(p.e0);
        ctx->event_pool.push_back(p.e1);
    }
    ctx->pending.clear();
}

void sync(vt_ctx* ctx)
{
    VT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->stage_off = 0;
    flush_timings(ctx);
}

// Copies up to kStageMax bytes go through the context's pinned staging ring (kStageCap
// bytes): the source is copied there at once, so the caller may reuse it on return, and the
// device copy is queued without waiting for the stream.  Larger copies (graph uploads) copy
// from the caller's memory directly.
constexpr size_t kStageCap = size_t(16) << 20;
constexpr size_t kStageMax = size_t(4) << 20;

void h2d(vt_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    if (!bytes) return;
    if (bytes > kStageMax) {
        VT_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        return;
    }
    if (!ctx->stage) {
        VT_HIP(ctx, hipHostMalloc((void**)&ctx->stage, kStageCap, hipHostMallocDefault));
        ctx->stage_off = 0;
    }
    const size_t need = (bytes + 255) & ~size_t(255);
    if (ctx->stage_off + need > kStageCap) {
        // the ring is full: its earlier copies must have run before it is overwritten
        VT_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ctx->stage_off = 0;
    }
    unsigned char* slot = ctx->stage + ctx->stage_off;
    std::memcpy(slot, src, bytes);
    VT_HIP(ctx, hipMemcpyAsync(dst, slot, bytes, hipMemcpyHostToDevice, ctx->stream));
    ctx->stage_off += need;
}
// Reads of up to kStageMax bytes (the per-iteration scalars, counts) land in the staging ring
// past its queued copies, as a copy to pinned memory, and are copied out after the wait.
void d2h(vt_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    if (bytes && bytes <= kStageMax) {
        if (!ctx->stage) {
            VT_HIP(ctx, hipHostMalloc((void**)&ctx->stage, kStageCap, hipHostMallocDefault));
            ctx->stage_off = 0;
        }
        if (ctx->stage_off + bytes > kStageCap) {
            VT_HIP(ctx, hipStreamSynchronize(ctx->stream));
            ctx->stage_off = 0;
        }
        unsigned char* slot = ctx->stage + ctx->stage_off;
        VT_HIP(ctx, hipMemcpyAsync(slot, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        VT_HIP(ctx, hipStreamSynchronize(ctx->stream));
        std::memcpy(dst, slot, bytes);
    } else {
        if (bytes) VT_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        VT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    ctx->stage_off = 0;  // the stream has drained: every staged copy has run
}
void d2d(vt_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    if (bytes) VT_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
}

__global__ void k_fill_zero(uint64_t* p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}
__global__ void k_fill_zero_bytes(unsigned char* p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Zero fill with a kernel (no memset node, so the sequence stays graph-capturable).
void dzero(vt_ctx* ctx, void* dst, size_t bytes)
{
    if (!bytes) return;
    if ((bytes % 8) == 0 && ((uintptr_t)dst % 8) == 0)
        hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(bytes / 8, 256, 4096)), dim3(256), 0, ctx->stream,
                           (uint64_t*)dst, (int64_t)(bytes / 8));
    else
        hipLaunchKernelGGL(k_fill_zero_bytes, dim3(grid_for(bytes, 256, 4096)), dim3(256), 0, ctx->stream,
                           (unsigned char*)dst, (int64_t)bytes);
    VT_HIP(ctx, hipGetLastError());
}

int64_t bits_for(int64_t maxval)
{
    int64_t b = 1;
    while (b < 63 && (int64_t(1) << b) <= maxval) b++;
    return b;
}

// ---- deterministic sums --------------------------------------------------------------

constexpr int kSumBlock = 256;
constexpr int kSumGrid = 1024;

// Each block owns the contiguous range [b*chunk, min(n,(b+1)*chunk)); thread t folds
// elements t, t+256, ... of it in order; the block then halves in LDS.  The grid and
// chunk depend only on n, so the result is a fixed function of the inputs.
__global__ void __launch_bounds__(kSumBlock) k_sum2_partial(const double* __restrict__ a,
                                                            const double* __restrict__ b, int64_t n,
                                                            int64_t chunk, double2* __restrict__ part)
{
    __shared__ double sa[kSumBlock], sb[kSumBlock];
    const int64_t lo = (int64_t)blockIdx.x * chunk;
    const int64_t hi = (lo + chunk < n) ? lo + chunk : n;
    double xa = 0.0, xb = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kSumBlock) {
        xa = xa + a[i];
        if (b) { const double v = b[i]; xb = xb + v * v; }
    }
    sa[threadIdx.x] = xa;
    sb[threadIdx.x] = xb;
    __syncthreads();
    for (int s = kSumBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sa[threadIdx.x] = sa[threadIdx.x] + sa[threadIdx.x + s];
            sb[threadIdx.x] = sb[threadIdx.x] + sb[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = make_double2(sa[0], sb[0]);
}

__global__ void __launch_bounds__(kSumBlock) k_sum2_final(const double2* __restrict__ part, int np,
                                                          double* __restrict__ out)
{
    __shared__ double sa[kSumBlock], sb[kSumBlock];
    double xa = 0.0, xb = 0.0;
    for (int i = threadIdx.x; i < np; i += kSumBlock) { xa = xa + part[i].x; xb = xb + part[i].y; }
    sa[threadIdx.x] = xa;
    sb[threadIdx.x] = xb;
    __syncthreads();
    for (int s = kSumBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sa[threadIdx.x] = sa[threadIdx.x] + sa[threadIdx.x + s];
            sb[threadIdx.x] = sb[threadIdx.x] + sb[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = sa[0]; out[1] = sb[0]; }
}

// The blocking of an n-element tree sum: at most kSumGrid blocks, four strides of 256
// elements per block at least.  A function of n alone.
static void sum2_plan(int64_t n, int64_t* grid_out, int64_t* chunk_out)
{
    int64_t grid = (n + 4 * kSumBlock - 1) / (4 * kSumBlock);
    if (grid < 1) grid = 1;
    if (grid > kSumGrid) grid = kSumGrid;
    const int64_t chunk = (n + grid - 1) / grid;
    *grid_out = grid;
    *chunk_out = chunk > 0 ? chunk : 1;
}

void tree_sum2(vt_ctx* ctx, const double* a, const double* b, int64_t n, double* d_out2)
{
    int64_t grid = 0, chunk = 0;
    sum2_plan(n, &grid, &chunk);
    double2* part = (double2*)scratch(ctx, "sum2_part", sizeof(double2) * kSumGrid);
    VT_LAUNCH(ctx, "sum2_partial", k_sum2_partial, (unsigned)grid, kSumBlock, 0, a, b, n, chunk, part);
    VT_LAUNCH(ctx, "sum2_final", k_sum2_final, 1, kSumBlock, 0, (const double2*)part, (int)grid, d_out2);
}

// Left folds from 0.0 in index order: one OpenMP thread's reduction order.
__global__ void k_seq_sum2(const double* __restrict__ a, const double* __restrict__ b, int64_t n,
                           double* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double xa = 0.0, xb = 0.0;
    for (int64_t i = 0; i < n; i++) {
        xa = xa + a[i];
        if (b) { const double v = b[i]; xb = xb + v * v; }
    }
    out[0] = 0.0 + xa;
    out[1] = 0.0 + xb;
}

void seq_sum2(vt_ctx* ctx, const double* a, const double* b, int64_t n, double* d_out2)
{
    VT_LAUNCH(ctx, "seq_sum2", k_seq_sum2, 1, 64, 0, a, b, n, d_out2);
}

// ---- per-graph sums over a batch of disjoint graphs ----------------------------------------
// A graph of n vertices at offset voff[g] is reduced with exactly the blocking and trees
// tree_sum2 uses on an n-vertex array (same grid, chunk, strided thread folds and block
// tree; same final block), so its sums in a batch are bit-identical to its sums alone.

__global__ void __launch_bounds__(kSumBlock) k_seg_sum2_partial(const double* __restrict__ a,
                                                                const double* __restrict__ b,
                                                                const int64_t* __restrict__ voff,
                                                                const int32_t* __restrict__ bgraph,
                                                                const int32_t* __restrict__ bsub,
                                                                const int64_t* __restrict__ chunk,
                                                                const unsigned char* __restrict__ gact,
                                                                double2* __restrict__ part)
{
    __shared__ double sa[kSumBlock], sb[kSumBlock];
    const int32_t g = bgraph[blockIdx.x];
    if (gact && !gact[g]) return;  // a graph that has left the loop: its sums are not read
    const int64_t base = voff[g], n = voff[g + 1] - base, ch = chunk[g];
    const int64_t lo = (int64_t)bsub[blockIdx.x] * ch;
    const int64_t hi = (lo + ch < n) ? lo + ch : n;
    double xa = 0.0, xb = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kSumBlock) {
        xa = xa + a[base + i];
        if (b) { const double v = b[base + i]; xb = xb + v * v; }
    }
    sa[threadIdx.x] = xa;
    sb[threadIdx.x] = xb;
    __syncthreads();
    for (int s = kSumBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sa[threadIdx.x] = sa[threadIdx.x] + sa[threadIdx.x + s];
            sb[threadIdx.x] = sb[threadIdx.x] + sb[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = make_double2(sa[0], sb[0]);
}

__global__ void __launch_bounds__(kSumBlock) k_seg_sum2_final(const double2* __restrict__ part,
                                                              const int64_t* __restrict__ pbase,
                                                              const unsigned char* __restrict__ gact,
                                                              double* __restrict__ out)
{
    __shared__ double sa[kSumBlock], sb[kSumBlock];
    const int32_t g = blockIdx.x;
    if (gact && !gact[g]) return;
    const double2* p = part + pbase[g];
    const int np = (int)(pbase[g + 1] - pbase[g]);
    double xa = 0.0, xb = 0.0;
    for (int i = threadIdx.x; i < np; i += kSumBlock) { xa = xa + p[i].x; xb = xb + p[i].y; }
    sa[threadIdx.x] = xa;
    sb[threadIdx.x] = xb;
    __syncthreads();
    for (int s = kSumBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sa[threadIdx.x] = sa[threadIdx.x] + sa[threadIdx.x + s];
            sb[threadIdx.x] = sb[threadIdx.x] + sb[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[2 * g] = sa[0]; out[2 * g + 1] = sb[0]; }
}

// Left folds per graph, as k_seq_sum2 forms them for one graph.
__global__ void k_seg_seq_sum2(const double* __restrict__ a, const double* __restrict__ b,
                               const int64_t* __restrict__ voff, int32_t ng, const unsigned char* __restrict__ gact,
                               double* __restrict__ out)
{
    for (int32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += gridDim.x * blockDim.x) {
        if (gact && !gact[g]) continue;
        double xa = 0.0, xb = 0.0;
        for (int64_t i = voff[g]; i < voff[g + 1]; i++) {
            xa = xa + a[i];
            if (b) { const double v = b[i]; xb = xb + v * v; }
        }
        out[2 * g] = 0.0 + xa;
        out[2 * g + 1] = 0.0 + xb;
    }
}

void seg_sums_plan(vt_ctx* ctx, const std::vector<int64_t>& voff, SegSums& s)
{
    const int32_t ng = (int32_t)voff.size() - 1;
    std::vector<int32_t> bgraph, bsub;
    std::vector<int64_t> chunk(ng), pbase(ng + 1, 0);
    for (int32_t g = 0; g < ng; g++) {
        const int64_t n = voff[g + 1] - voff[g];
        int64_t grid = (n + 4 * kSumBlock - 1) / (4 * kSumBlock);
        if (grid < 1) grid = 1;
        if (grid > kSumGrid) grid = kSumGrid;
        const int64_t ch = (n + grid - 1) / grid;
        chunk[g] = ch > 0 ? ch : 1;
        for (int64_t b = 0; b < grid; b++) {
            bgraph.push_back(g);
            bsub.push_back((int32_t)b);
        }
        pbase[g + 1] = pbase[g] + grid;
    }
    s.ng = ng;
    s.nblocks = (int64_t)bgraph.size();
    s.voff = DBuf<int64_t>(ctx, voff.size());
    s.bgraph = DBuf<int32_t>(ctx, bgraph.size());
    s.bsub = DBuf<int32_t>(ctx, bsub.size());
    s.chunk = DBuf<int64_t>(ctx, chunk.size());
    s.pbase = DBuf<int64_t>(ctx, pbase.size());
    s.part = DBuf<double2>(ctx, bgraph.size());
    h2d(ctx, s.voff.p, voff.data(), 8 * voff.size());
    h2d(ctx, s.bgraph.p, bgraph.data(), 4 * bgraph.size());
    h2d(ctx, s.bsub.p, bsub.data(), 4 * bsub.size());
    h2d(ctx, s.chunk.p, chunk.data(), 8 * chunk.size());
    h2d(ctx, s.pbase.p, pbase.data(), 8 * pbase.size());
}

void seg_tree_sum2(vt_ctx* ctx, const SegSums& s, const double* a, const double* b, double* d_out,
                   const unsigned char* gact)
{
    if (s.ng == 0) return;
    VT_LAUNCH(ctx, "seg_sum2_partial", k_seg_sum2_partial, (unsigned)s.nblocks, kSumBlock, 0, a, b,
              (const int64_t*)s.voff.p, (const int32_t*)s.bgraph.p, (const int32_t*)s.bsub.p,
              (const int64_t*)s.chunk.p, gact, s.part.p);
    VT_LAUNCH(ctx, "seg_sum2_final", k_seg_sum2_final, (unsigned)s.ng, kSumBlock, 0, (const double2*)s.part.p,
              (const int64_t*)s.pbase.p, gact, d_out);
}

void seg_seq_sum2(vt_ctx* ctx, const SegSums& s, const double* a, const double* b, double* d_out,
                  const unsigned char* gact)
{
    if (s.ng == 0) return;
    VT_LAUNCH(ctx, "seg_seq_sum2", k_seg_seq_sum2, grid_for(s.ng, 64, 4096), 64, 0, a, b, (const int64_t*)s.voff.p,
              s.ng, gact, d_out);
}

// ---- scans and sorts (rocPRIM) --------------------------------------------------------

void exclusive_scan_i64(vt_ctx* ctx, const int64_t* in, int64_t* out, int64_t n)
{
    if (n <= 0) return;
    size_t tb = 0;
    VT_HIP(ctx, rocprim::exclusive_scan(nullptr, tb, in, out, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(),
                                        ctx->stream));
    void* tmp = scratch(ctx, "scan_tmp", tb);
    TimedScope ts(ctx, "scan_i64");
    VT_HIP(ctx, rocprim::exclusive_scan(tmp, tb, in, out, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(),
                                        ctx->stream));
}

void exclusive_scan_i32(vt_ctx* ctx, const int32_t* in, int32_t* out, int64_t n)
{
    if (n <= 0) return;
    size_t tb = 0;
    VT_HIP(ctx, rocprim::exclusive_scan(nullptr, tb, in, out, (int32_t)0, (size_t)n, rocprim::plus<int32_t>(),
                                        ctx->stream));
    void* tmp = scratch(ctx, "scan_tmp", tb);
    TimedScope ts(ctx, "scan_i32");
    VT_HIP(ctx, rocprim::exclusive_scan(tmp, tb, in, out, (int32_t)0, (size_t)n, rocprim::plus<int32_t>(),
                                        ctx->stream));
}

void sort_pairs_u32(vt_ctx* ctx, uint32_t* keys, uint32_t* vals, int64_t n, int bits, uint32_t* keys_alt,
                    uint32_t* vals_alt, bool* result_in_alt)
{
    *result_in_alt = false;
    if (n <= 1) return;
    rocprim::double_buffer<uint32_t> kb(keys, keys_alt);
    rocprim::double_buffer<uint32_t> vb(vals, vals_alt);
    size_t tb = 0;
    VT_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tb, kb, vb, (size_t)n, 0, bits, ctx->stream));
    void* tmp = scratch(ctx, "sort_tmp", tb);
    TimedScope ts(ctx, "sort_u32");
    VT_HIP(ctx, rocprim::radix_sort_pairs(tmp, tb, kb, vb, (size_t)n, 0, bits, ctx->stream));
    *result_in_alt = (kb.current() == keys_alt);
}

void sort_pairs_u64(vt_ctx* ctx, uint64_t* keys, uint32_t* vals, int64_t n, int bits, uint64_t* keys_alt,
                    uint32_t* vals_alt, bool* result_in_alt)
{
    *result_in_alt = false;
    if (n <= 1) return;
    rocprim::double_buffer<uint64_t> kb(keys, keys_alt);
    rocprim::double_buffer<uint32_t> vb(vals, vals_alt);
    size_t tb = 0;
    VT_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tb, kb, vb, (size_t)n, 0, bits, ctx->stream));
    void* tmp = scratch(ctx, "sort_tmp", tb);
    TimedScope ts(ctx, "sort_u64");
    VT_HIP(ctx, rocprim::radix_sort_pairs(tmp, tb, kb, vb, (size_t)n, 0, bits, ctx->stream));
    *result_in_alt = (kb.current() == keys_alt);
}

// ---- graphs ------------------------------------------------------------------------------

vt_graph* graph_new(vt_ctx* ctx, int64_t nv, int64_t ne)
{
    vt_graph* g = new vt_graph();
    g->nv = nv;
    g->ne = ne;
    try {
        g->offs = (int64_t*)dalloc(ctx, sizeof(int64_t) * (nv + 1));
        g->edges = (DEdge*)dalloc(ctx, sizeof(DEdge) * (ne ? ne : 1));
    } catch (...) {
        graph_delete(ctx, g);
        throw;
    }
    g->bytes = sizeof(int64_t) * (nv + 1) + sizeof(DEdge) * (ne ? ne : 1);
    return g;
}

void graph_delete(vt_ctx* ctx, vt_graph* g)
{
    if (!g) return;
    dfree(ctx, g->offs, sizeof(int64_t) * (g->nv + 1));
    dfree(ctx, g->edges, sizeof(DEdge) * (g->ne ? g->ne : 1));
    if (g->part) dfree(ctx, g->part, sizeof(int64_t) * (g->nv ? g->nv : 1));
    delete g;
}

__global__ void k_max_degree(const int64_t* __restrict__ offs, int64_t nv, unsigned long long* __restrict__ out)
{
    int64_t m = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t d = offs[i + 1] - offs[i];
        m = d > m ? d : m;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const int64_t x = __shfl_xor(m, o);
        m = x > m ? x : m;
    }
    // one atomic per block: every block's atomics go to the one address
    __shared__ int64_t s_max[256 / 64];
    if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); k++) m = s_max[k] > m ? s_max[k] : m;
        atomicMax(out, (unsigned long long)m);
    }
}

// Validation: offsets monotone from 0 to ne, tails in range.  out[0] counts bad rows,
// out[1] bad tails.
__global__ void k_validate(const int64_t* __restrict__ offs, const DEdge* __restrict__ edges, int64_t nv,
                           int64_t ne, unsigned long long* __restrict__ out)
{
    unsigned long long badr = 0, badt = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t a = offs[i], b = offs[i + 1];
        if (a > b || a < 0 || b > ne) { badr++; continue; }
    }
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = edges[e].tail;
        if (t < 0 || t >= nv) badt++;
    }
    if (badr) atomicAdd(out, badr);
    if (badt) atomicAdd(out + 1, badt);
}

void graph_finalize(vt_ctx* ctx, vt_graph* g)
{
    unsigned long long* d = (unsigned long long*)scratch(ctx, "small_u64", 64);
    dzero(ctx, d, 8);
    if (g->nv)
        VT_LAUNCH(ctx, "max_degree", k_max_degree, grid_for(g->nv, 256, 512), 256, 0, g->offs, g->nv, d);
    unsigned long long h = 0;
    d2h(ctx, &h, d, 8);
    g->max_degree = (int64_t)h;
}

}  // namespace vt

using namespace vt;

// ======================================================================================
// Allocation-free modularity (vt_modularity_noalloc): its partial sums live in a fixed
// device array instead of memory the context allocates.
// ======================================================================================

// The tree's block partials (one double2 per block, at most kSumGrid blocks: 16 KiB) and
// the two final sums.  Every device holds its own copy until the process ends.  Calls on
// one device share that copy, so two calls that overlap on one device race.
constexpr size_t kNoAllocPartBytes = sizeof(double2) * kSumGrid;
constexpr size_t kNoAllocBytes = kNoAllocPartBytes + 256;
static __device__ __align__(256) unsigned char noalloc_scratch[kNoAllocBytes];

namespace {

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoAllocMaxDevices] = {};  // device address of each device's copy

// Resolves the current device's copy of noalloc_scratch, once per device.  In host code the
// array's name is only a stand-in for the symbol; kernels get this address instead.
hipError_t noalloc_setup(int device)
{
    if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch));
    if (e == hipSuccess) g_noalloc_ptr[device] = p;
    return e;
}

void* noalloc_ptr(int device)
{
    if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    return g_noalloc_ptr[device];
}

// k_modularity with 1/2m passed by value: |(Σcw)c − (Σa²)c·c|, the same expression.
__global__ void k_modularity_value(const double* __restrict__ sums, double constant, double* __restrict__ q)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) q[0] = fabs((sums[0] * constant) - (sums[1] * constant * constant));
}

}  // namespace

// ======================================================================================
// C ABI: context and graphs
// ======================================================================================

extern "C" {

int vt_ctx_create(int device, uint64_t hbm_budget_bytes, vt_ctx** out)
{
    if (!out) return VT_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return VT_ERR_HIP;
    if (device < 0 || device >= n) return VT_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return VT_ERR_HIP;
    if (noalloc_setup(device) != hipSuccess) return VT_ERR_HIP;
    vt_ctx* c = new vt_ctx();
    c->device = device;
    c->budget = hbm_budget_bytes;
    c->tuning = vt::default_tuning();
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return VT_ERR_HIP;
    }
    // Keep freed blocks in the device's stream-ordered pool instead of unmapping them at
    // every synchronisation: a phase loop frees and reallocates the same large buffers
    // (phase state, rebuild keys) level after level.
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
        uint64_t keep = ~0ull;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    *out = c;
    return VT_OK;
}

int vt_ctx_destroy(vt_ctx* ctx)
{
    if (!ctx) return VT_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    release_scratch(ctx);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& p : ctx->pending) { ctx->event_pool.push_back(p.e0); ctx->event_pool.push_back(p.e1); }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    if (ctx->stage) (void)hipHostFree(ctx->stage);
    delete ctx;
    return VT_OK;
}

const char* vt_last_error(const vt_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : "null context";
}

uint64_t vt_ctx_bytes_in_use(const vt_ctx* ctx)
{
    return ctx ? ctx->in_use : 0;
}

int vt_ctx_set_profiling(vt_ctx* ctx, int on)
{
    if (!ctx) return VT_ERR_ARG;
    return guarded(ctx, [&] {
        flush_timings(ctx);
        ctx->profiling = on != 0;
    });
}

int vt_ctx_kernel_stats(const vt_ctx* ctx, const char* name, double* avg_ms, int64_t* count, double* total_ms)
{
    if (!ctx || !name) return VT_ERR_ARG;
    flush_timings(const_cast<vt_ctx*>(ctx));
    auto it = ctx->stats.find(name);
    const int64_t c = (it == ctx->stats.end()) ? 0 : it->second.count;
    const double t = (it == ctx->stats.end()) ? 0.0 : it->second.total_ms;
    if (count) *count = c;
    if (total_ms) *total_ms = t;
    if (avg_ms) *avg_ms = c ? t / (double)c : 0.0;
    return VT_OK;
}

int vt_ctx_reset_stats(vt_ctx* ctx)
{
    if (!ctx) return VT_ERR_ARG;
    flush_timings(ctx);
    ctx->stats.clear();
    return VT_OK;
}

int vt_graph_upload(vt_ctx* ctx, int64_t nv, int64_t ne, const int64_t* offsets, const vt_edge* edges,
                    vt_graph** out)
{
    if (!ctx || !out) return VT_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        VT_HIP(ctx, hipSetDevice(ctx->device));
        if (nv < 0 || ne < 0 || !offsets || (ne > 0 && !edges))
            fail(ctx, VT_ERR_ARG, "vt_graph_upload: bad arguments (nv=%lld ne=%lld)", (long long)nv, (long long)ne);
        if (nv >= (int64_t(1) << 31) - 1)
            fail(ctx, VT_ERR_LIMIT, "vt_graph_upload: %lld vertices; this build supports fewer than 2^31 - 1",
                 (long long)nv);
        if (offsets[0] != 0 || offsets[nv] != ne)
            fail(ctx, VT_ERR_ARG, "vt_graph_upload: offsets must run from 0 to ne (got %lld..%lld, ne=%lld)",
                 (long long)offsets[0], (long long)offsets[nv], (long long)ne);
        vt_graph* g = graph_new(ctx, nv, ne);
        try {
            h2d(ctx, g->offs, offsets, sizeof(int64_t) * (nv + 1));
            h2d(ctx, g->edges, edges, sizeof(DEdge) * ne);
            unsigned long long* d = (unsigned long long*)scratch(ctx, "small_u64", 64);
            dzero(ctx, d, 16);
            VT_LAUNCH(ctx, "validate", k_validate, grid_for(std::max(nv, ne), 256, 4096), 256, 0, g->offs,
                      g->edges, nv, ne, d);
            unsigned long long bad[2] = {0, 0};
            d2h(ctx, bad, d, 16);
            if (bad[0] || bad[1])
                fail(ctx, VT_ERR_ARG, "vt_graph_upload: malformed CSR (%llu bad rows, %llu tails out of range)",
                     bad[0], bad[1]);
            graph_finalize(ctx, g);
        } catch (...) {
            graph_delete(ctx, g);
            throw;
        }
        *out = g;
    });
}

int vt_graph_info(const vt_graph* g, int64_t* nv, int64_t* ne)
{
    if (!g) return VT_ERR_ARG;
    if (nv) *nv = g->nv;
    if (ne) *ne = g->ne;
    return VT_OK;
}

int vt_graph_download(vt_ctx* ctx, const vt_graph* g, int64_t* offsets, vt_edge* edges)
{
    if (!ctx || !g) return VT_ERR_ARG;
    return guarded(ctx, [&] {
        VT_HIP(ctx, hipSetDevice(ctx->device));
        if (offsets) d2h(ctx, offsets, g->offs, sizeof(int64_t) * (g->nv + 1));
        if (edges && g->ne) d2h(ctx, edges, g->edges, sizeof(DEdge) * g->ne);
    });
}

int vt_graph_free(vt_ctx* ctx, vt_graph* g)
{
    if (!ctx) return VT_ERR_ARG;
    return guarded(ctx, [&] {
        VT_HIP(ctx, hipSetDevice(ctx->device));
        sync(ctx);
        graph_delete(ctx, g);
    });
}

int vt_modularity_noalloc(vt_ctx* ctx, int64_t nv, const double* d_cw, const double* d_cdeg, double constant,
                          int32_t flags, double* d_q_out)
{
    if (!ctx) return VT_ERR_ARG;
    return guarded(ctx, [&] {
        if (nv < 0 || !d_q_out || (nv > 0 && (!d_cw || !d_cdeg)))
            fail(ctx, VT_ERR_ARG, "vt_modularity_noalloc: bad arguments (nv=%lld)", (long long)nv);
        VT_HIP(ctx, hipSetDevice(ctx->device));
        unsigned char* base = (unsigned char*)noalloc_ptr(ctx->device);
        if (!base) fail(ctx, VT_ERR_STATE, "vt_modularity_noalloc: no scratch address for device %d", ctx->device);
        int64_t grid = 0, chunk = 0;
        sum2_plan(nv, &grid, &chunk);
        if ((size_t)grid * sizeof(double2) > kNoAllocPartBytes)
            fail(ctx, VT_ERR_LIMIT, "vt_modularity_noalloc: %lld partial sums exceed the %zu-byte scratch",
                 (long long)grid, kNoAllocPartBytes);
        double2* part = (double2*)base;
        double* sums = (double*)(base + kNoAllocPartBytes);
        // the same kernels, grid and chunk as vt_modularity, so the same bits
        if (flags & VT_FLAG_SEQUENTIAL_SUMS) {
            hipLaunchKernelGGL(k_seq_sum2, dim3(1), dim3(64), 0, ctx->stream, d_cw, d_cdeg, nv, sums);
        } else {
            hipLaunchKernelGGL(k_sum2_partial, dim3((unsigned)grid), dim3(kSumBlock), 0, ctx->stream, d_cw, d_cdeg, nv,
                               chunk, part);
            VT_HIP(ctx, hipGetLastError());
            hipLaunchKernelGGL(k_sum2_final, dim3(1), dim3(kSumBlock), 0, ctx->stream, (const double2*)part, (int)grid,
                               sums);
        }
        VT_HIP(ctx, hipGetLastError());
        hipLaunchKernelGGL(k_modularity_value, dim3(1), dim3(64), 0, ctx->stream, (const double*)sums, constant,
                           d_q_out);
        VT_HIP(ctx, hipGetLastError());
        VT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    });
}

}  // extern "C"
