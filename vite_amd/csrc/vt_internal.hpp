// Internal declarations shared by the vite_amd translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/vite_amd.h"

namespace vt {

// Device edge record: identical layout to vt_edge / the reference Edge (edge.hpp:72-80),
// 16-byte aligned so a row element is one dwordx4 load.
struct __align__(16) DEdge {
    int64_t tail;
    double weight;
};
static_assert(sizeof(DEdge) == 16, "edge record must be 16 bytes");

struct Error {
    int code;
    std::string msg;
};

struct KernelStat {
    int64_t count = 0;
    double total_ms = 0.0;
};

struct PendingTiming {
    std::string name;
    hipEvent_t e0, e1;
};

}  // namespace vt

struct vt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    uint64_t in_use = 0;
    uint64_t peak = 0;    // high-water mark of in_use + cached since the last reset
    uint64_t cached = 0;  // bytes of freed blocks kept for reuse
    std::multimap<size_t, void*> cache;    // freed blocks by size
    std::map<void*, size_t> blocks;        // live blocks and their sizes
    std::string err;
    // reusable scratch buffers keyed by purpose
    std::map<std::string, std::pair<void*, size_t>> scratch;
    // kernel timing
    bool profiling = false;
    std::vector<vt::PendingTiming> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, vt::KernelStat> stats;
    // per-iteration trace of the last phase
    std::vector<double> trace_q;
    std::vector<int64_t> trace_moves;
    // pinned host staging for small host-to-device copies: a copy from pageable memory
    // waits for the stream's earlier work; a copy from here is queued like a kernel.  The
    // ring restarts whenever the stream has drained (sync, d2h).
    unsigned char* stage = nullptr;
    size_t stage_off = 0;
    // sweep launch shape (performance only: every setting gives the same results)
    vt_tuning tuning = {{7, 15, 31, 63, 127, 1023}, 16384};  // replaced by default_tuning() at creation
};

struct vt_graph {
    int64_t nv = 0, ne = 0;
    int64_t* offs = nullptr;     // device, nv+1
    vt::DEdge* edges = nullptr;  // device, ne
    int64_t max_degree = 0;
    int64_t* part = nullptr;     // device partition left by the last phase (nv), or null
    size_t bytes = 0;
};

namespace vt {

// ---- errors -------------------------------------------------------------------------
struct Fail {
    int code;
};

[[noreturn]] void fail(vt_ctx* ctx, int code, const char* fmt, ...);
void hip_check(vt_ctx* ctx, hipError_t e, const char* what, const char* file, int line);
#define VT_HIP(ctx, call) ::vt::hip_check((ctx), (call), #call, __FILE__, __LINE__)

// Runs `body` converting vt::Fail into the status code; the message stays on ctx.
template <class F>
int guarded(vt_ctx* ctx, F&& body)
{
    try {
        body();
        return VT_OK;
    } catch (const Fail& f) {
        return f.code;
    } catch (const std::exception& e) {
        if (ctx) ctx->err = std::string("internal error: ") + e.what();
        return VT_ERR_HIP;
    }
}

// ---- memory ---------------------------------------------------------------------------
void* dalloc(vt_ctx* ctx, size_t bytes);
void dfree(vt_ctx* ctx, void* p, size_t bytes);
// Grow-only named scratch buffer.
void* scratch(vt_ctx* ctx, const std::string& key, size_t bytes);
void release_scratch(vt_ctx* ctx);  // also returns cached blocks to the driver
void trim_cache(vt_ctx* ctx);

template <class T>
struct DBuf {
    vt_ctx* ctx = nullptr;
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(vt_ctx* c, size_t count) : ctx(c), n(count) { p = (T*)dalloc(c, sizeof(T) * (count ? count : 1)); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : ctx(o.ctx), p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept
    {
        reset();
        ctx = o.ctx; p = o.p; n = o.n;
        o.p = nullptr; o.n = 0;
        return *this;
    }
    void reset()
    {
        if (p) dfree(ctx, p, sizeof(T) * (n ? n : 1));
        p = nullptr;
        n = 0;
    }
    ~DBuf() { reset(); }
};

// ---- launches with optional event timing ---------------------------------------------
struct TimedScope {
    vt_ctx* ctx;
    const char* name;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TimedScope(vt_ctx* c, const char* n);
    ~TimedScope();
};

void flush_timings(vt_ctx* ctx);

#define VT_LAUNCH(ctx, name, kernel, grid, block, shmem, ...)                                  \
    do {                                                                                      \
        ::vt::TimedScope ts_(ctx, name);                                                      \
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), shmem, (ctx)->stream, __VA_ARGS__); \
        VT_HIP(ctx, hipGetLastError());                                                       \
    } while (0)

void sync(vt_ctx* ctx);
void h2d(vt_ctx* ctx, void* dst, const void* src, size_t bytes);
void d2h(vt_ctx* ctx, void* dst, const void* src, size_t bytes);
void d2d(vt_ctx* ctx, void* dst, const void* src, size_t bytes);
void dzero(vt_ctx* ctx, void* dst, size_t bytes);

inline unsigned grid_for(int64_t n, int block, int64_t cap = 1 << 20)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// Folds the staged weights of the lanes in `rest` (ascending lane = edge order) onto acc,
// U terms per round so the LDS latency is paid once per round.  Rounds past the last term
// read wst[-1] = -0.0, and x + (-0.0) == x for every x, so no select is needed.
template <class mask_t, int U>
__device__ __forceinline__ double fold_rest(double acc, mask_t rest, const double* wst)
{
    while (rest) {
        int k[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            k[u] = rest ? (int)(sizeof(mask_t) == 8 ? __builtin_ctzll((unsigned long long)rest)
                                                    : __builtin_ctz((unsigned)rest))
                        : -1;
            rest = (mask_t)(rest & (rest - 1));
        }
        double w[U];
#pragma unroll
        for (int u = 0; u < U; u++) w[u] = wst[k[u]];
#pragma unroll
        for (int u = 0; u < U; u++) acc = acc + w[u];
    }
    return acc;
}

// Lane j's double, read as a wave-uniform value (two 32-bit lane reads).
__device__ __forceinline__ double readlane_f64(double x, int j)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, j);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), j);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// ---- device primitives (vt_prims.hip) -------------------------------------------------
// Deterministic fixed-order sums: every block reduces a fixed contiguous chunk with a
// fixed tree; block partials are reduced by one block in index order.
void tree_sum2(vt_ctx* ctx, const double* a, const double* b_sq, int64_t n, double* d_out2);
void seq_sum2(vt_ctx* ctx, const double* a, const double* b_sq, int64_t n, double* d_out2);
// The same sums per graph of a batch of disjoint graphs (vertex ranges voff[g]..voff[g+1]),
// each bit-identical to tree_sum2 / seq_sum2 on that graph alone.  d_out gets 2 per graph.
struct SegSums {
    int32_t ng = 0;
    int64_t nblocks = 0;
    DBuf<int64_t> voff, chunk, pbase;
    DBuf<int32_t> bgraph, bsub;
    DBuf<double2> part;
};
void seg_sums_plan(vt_ctx* ctx, const std::vector<int64_t>& voff, SegSums& s);
// gact (optional): graphs whose flag is 0 are skipped and their sums left as they were.
void seg_tree_sum2(vt_ctx* ctx, const SegSums& s, const double* a, const double* b_sq, double* d_out,
                   const unsigned char* gact = nullptr);
void seg_seq_sum2(vt_ctx* ctx, const SegSums& s, const double* a, const double* b_sq, double* d_out,
                  const unsigned char* gact = nullptr);
void exclusive_scan_i64(vt_ctx* ctx, const int64_t* in, int64_t* out, int64_t n);
void exclusive_scan_i32(vt_ctx* ctx, const int32_t* in, int32_t* out, int64_t n);
// Stable radix sort of (key, value) pairs on key bits [0, bits).
void sort_pairs_u32(vt_ctx* ctx, uint32_t* keys, uint32_t* vals, int64_t n, int bits,
                    uint32_t* keys_alt, uint32_t* vals_alt, bool* result_in_alt);
void sort_pairs_u64(vt_ctx* ctx, uint64_t* keys, uint32_t* vals, int64_t n, int bits,
                    uint64_t* keys_alt, uint32_t* vals_alt, bool* result_in_alt);
// The same scan and sort with the element count in device memory (vt_dsort.hip): grids are
// sized for n_max and blocks past the count exit, so a sequence of them can be captured in a
// HIP graph and replayed for any count up to n_max.  Their scratch is allocated on the first
// call of each size (which must not be under capture).
//   dev_scan_i32: out = exclusive scan of in[0, *d_n); *d_total (if not null) = the sum.
//   dev_sort_pairs_u32: stable sort of the first (*d_count) * mult pairs by key bits
//   [0, bits); returns whether the result is in the alt buffers (fixed by bits).
void dev_scan_i32(vt_ctx* ctx, const int32_t* in, int32_t* out, const int64_t* d_n, int64_t n_max, int64_t* d_total,
                  const char* part_key);
bool dev_sort_pairs_u32(vt_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                        const int64_t* d_count, int64_t mult, int64_t n_max, int bits);

// ---- graph helpers ---------------------------------------------------------------------
vt_graph* graph_new(vt_ctx* ctx, int64_t nv, int64_t ne);
void graph_finalize(vt_ctx* ctx, vt_graph* g);  // computes max_degree
// The largest offs[i + 1] - offs[i] over i < nv, atomically maxed into *out (cleared first).
__global__ void k_max_degree(const int64_t* __restrict__ offs, int64_t nv, unsigned long long* __restrict__ out);
void graph_delete(vt_ctx* ctx, vt_graph* g);

int64_t bits_for(int64_t maxval);
vt_tuning default_tuning();  // bucket bounds at each kernel's capacity (vt_louvain.hip)

// ---- batch of disjoint graphs (vt_louvain.hip) ----------------------------------------------
void batch_phase(vt_ctx* ctx, vt_graph* u, const int32_t* d_gid, const SegSums& seg, const std::vector<double>& lower,
                 const std::vector<double>& thresh, int flags, bool use_graph, std::vector<double>& q_out,
                 std::vector<int32_t>& iters_out, int* rounds_out, int* relists_out = nullptr);
// check_ids: d_part comes from a caller and its ids are checked to lie in [0, nv) (one host
// wait); a partition a phase left on the device needs no check.
vt_graph* rebuild(vt_ctx* ctx, const vt_graph* g, const int64_t* d_part, bool check_ids);

}  // namespace vt
