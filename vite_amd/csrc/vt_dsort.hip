// Scan and stable radix sort whose element count lives in device memory.
//
// The batched phase loop replays each Louvain round from one captured HIP graph.  The round's
// event count (two events per vertex that moves) is only known on the device, so the sort of
// the events cannot take its size from the host.  These kernels read the count from device
// memory, size their grids for a host-known capacity and let the blocks past the count exit
// at once.  They record no memset, allocate nothing once their scratch exists, and make no
// host read, so a sequence of them can be captured and replayed.
//
// Sort: least-significant-digit radix sort, 8- or 9-bit digits, tiles of 2048 pairs.  Per pass a
// histogram kernel counts each tile's digits (digit-major table: digit d of tile b at
// d * ntiles + b), the table is scanned, and a scatter kernel places each pair at its digit's
// offset for its tile plus its rank among the tile's pairs of that digit in input order.
// Ranks come from wave ballots over the digit bits, so equal keys keep their input order
// (stable), which is what the per-community event folds need.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "vt_internal.hpp"

namespace vt {

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int64_t kScanChunk = (int64_t)kScanBlock * kScanItems;  // 2048 values per block
constexpr int kPartBlock = 1024;

constexpr int kSortBlock = 256;
constexpr int kSortRounds = 8;
constexpr int64_t kTile = (int64_t)kSortBlock * kSortRounds;  // 2048 pairs per tile
constexpr int kWaves = kSortBlock / 64;
constexpr int kMaxDigitBits = 9;                // digits of 8 or 9 bits
constexpr int kMaxBins = 1 << kMaxDigitBits;    // bins of a 9-bit digit
constexpr int kMaxBinsPerThread = kMaxBins / kSortBlock;

__device__ __forceinline__ int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Inclusive prefix sum over the 256 threads of a block; returns the thread's inclusive sum
// and leaves the block total in *total.
__device__ __forceinline__ int64_t block_inclusive_scan(int64_t x, int64_t* s_wave, int64_t* total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wave[w] = x;
    __syncthreads();
    int64_t before = 0, all = 0;
    for (int k = 0; k < kScanBlock / 64; k++) {
        if (k < w) before += s_wave[k];
        all += s_wave[k];
    }
    __syncthreads();
    *total = all;
    return x + before;
}

// Chunk partial sums of in[0, n), chunks of 2048 values dealt to the blocks in turn.
__global__ void __launch_bounds__(kScanBlock) k_dscan_reduce(const int32_t* __restrict__ in, const int64_t* __restrict__ d_n,
                                                           int64_t* __restrict__ part)
{
    __shared__ int64_t s_wave[kScanBlock / 64];
    const int64_t n = *d_n;
    const int64_t nb = ceil_div(n, kScanChunk);
    for (int64_t c = blockIdx.x; c < nb; c += gridDim.x) {
        const int64_t base = c * kScanChunk;
        int64_t x = 0;
        for (int k = 0; k < kScanItems; k++) {
            const int64_t i = base + (int64_t)k * kScanBlock + threadIdx.x;
            if (i < n) x += in[i];
        }
        int64_t total = 0;
        block_inclusive_scan(x, s_wave, &total);
        if (threadIdx.x == 0) part[c] = total;
    }
}

// Exclusive scan of the block partials in place (one block), and the grand total.
__global__ void __launch_bounds__(kPartBlock) k_dscan_parts(const int64_t* __restrict__ d_n, int64_t* __restrict__ part,
                                                          int64_t* __restrict__ d_total)
{
    __shared__ int64_t s_wave[kPartBlock / 64];
    const int64_t nb = ceil_div(*d_n, kScanChunk);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int64_t carry = 0;
    for (int64_t base = 0; base < nb; base += kPartBlock) {
        const int64_t i = base + threadIdx.x;
        const int64_t v = i < nb ? part[i] : 0;
        int64_t x = v;
        for (int o = 1; o < 64; o <<= 1) {
            const int64_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) s_wave[w] = x;
        __syncthreads();
        int64_t before = 0, all = 0;
        for (int k = 0; k < kPartBlock / 64; k++) {
            if (k < w) before += s_wave[k];
            all += s_wave[k];
        }
        if (i < nb) part[i] = carry + before + x - v;
        carry += all;
        __syncthreads();
    }
    if (threadIdx.x == 0 && d_total) *d_total = carry;
}

// Exclusive scan of each chunk's values, offset by the chunk's scanned partial.
__global__ void __launch_bounds__(kScanBlock) k_dscan_down(const int32_t* __restrict__ in, const int64_t* __restrict__ d_n,
                                                         const int64_t* __restrict__ part, int32_t* __restrict__ out)
{
    __shared__ int64_t s_wave[kScanBlock / 64];
    const int64_t n = *d_n;
    const int64_t nb = ceil_div(n, kScanChunk);
    for (int64_t c = blockIdx.x; c < nb; c += gridDim.x) {
        // thread t scans the contiguous run [t*8, t*8+8) of the chunk
        const int64_t lo = c * kScanChunk + (int64_t)threadIdx.x * kScanItems;
        int32_t v[kScanItems];
        int64_t x = 0;
        for (int k = 0; k < kScanItems; k++) {
            v[k] = (lo + k < n) ? in[lo + k] : 0;
            x += v[k];
        }
        int64_t total = 0;
        int64_t run = block_inclusive_scan(x, s_wave, &total) - x + part[c];
        for (int k = 0; k < kScanItems; k++) {
            if (lo + k < n) out[lo + k] = (int32_t)run;
            run += v[k];
        }
    }
}

// Sizes of one sort: pairs, tiles and histogram entries (nbins per tile).
__global__ void k_dsort_sizes(const int64_t* __restrict__ d_count, int64_t mult, int nbins, int64_t* __restrict__ sizes)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const int64_t n = *d_count * mult;
        const int64_t nt = ceil_div(n, kTile);
        sizes[0] = n;
        sizes[1] = nt;
        sizes[2] = nbins * nt;
    }
}

// The pass's digit: key bits [shift, shift + digit bits), cut at the sort's last bit (dmask).
__device__ __forceinline__ uint32_t digit_of(uint32_t key, int shift, uint32_t dmask) { return (key >> shift) & dmask; }

// Digit counts of each tile, digit-major (nbins = 256 or 512 bins).
__global__ void __launch_bounds__(kSortBlock) k_dsort_hist(const uint32_t* __restrict__ keys,
                                                         const int64_t* __restrict__ sizes, int shift, uint32_t dmask,
                                                         int nbins, int32_t* __restrict__ hist)
{
    __shared__ int32_t s_hist[kMaxBins];
    const int64_t n = sizes[0], nt = sizes[1];
    for (int64_t b = blockIdx.x; b < nt; b += gridDim.x) {
        for (int d = threadIdx.x; d < nbins; d += kSortBlock) s_hist[d] = 0;
        __syncthreads();
        for (int r = 0; r < kSortRounds; r++) {
            const int64_t i = b * kTile + (int64_t)r * kSortBlock + threadIdx.x;
            if (i < n) atomicAdd(&s_hist[digit_of(keys[i], shift, dmask)], 1);
        }
        __syncthreads();
        for (int d = threadIdx.x; d < nbins; d += kSortBlock) hist[(int64_t)d * nt + b] = s_hist[d];
        __syncthreads();
    }
}

// Places the pairs of one tile in two steps.  First every pair gets its place in the tile
// sorted by digit (the tile's pairs of smaller digits, then those of its digit in earlier
// rounds, earlier waves of this round and earlier lanes of this wave) and is staged there in
// LDS.  Then the staged tile is written out in that order: a pair at staged place p with
// digit d goes to (scanned table entry of (d, tile)) + (p - first staged place of d), so the
// pairs of one digit leave as one contiguous run and the writes coalesce.
__global__ void __launch_bounds__(kSortBlock) k_dsort_scatter(const uint32_t* __restrict__ keys,
                                                            const uint32_t* __restrict__ vals,
                                                            const int64_t* __restrict__ sizes, int shift,
                                                            uint32_t dmask, int dbits, const int32_t* __restrict__ hist,
                                                            const int32_t* __restrict__ offs,
                                                            uint32_t* __restrict__ keys_out,
                                                            uint32_t* __restrict__ vals_out)
{
    __shared__ int32_t s_run[kMaxBins];            // digit's pairs placed by earlier rounds
    __shared__ int32_t s_wcnt[kWaves][kMaxBins];   // this round's pairs per wave and digit
    __shared__ int32_t s_first[kMaxBins];          // first staged place of each digit
    __shared__ int32_t s_out[kMaxBins];            // global place of each digit's first pair, minus s_first
    __shared__ int64_t s_wave[kWaves];
    __shared__ uint32_t s_key[kTile], s_val[kTile];
    const int64_t n = sizes[0], nt = sizes[1];
    const int nbins = 1 << dbits;
    const int bpt = nbins / kSortBlock;  // bins per thread, consecutive: 1 or 2
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1;
    for (int64_t b = blockIdx.x; b < nt; b += gridDim.x) {
        // the tile's digit counts, scanned: first staged place of each digit
        int32_t cnt[kMaxBinsPerThread];
        int32_t mine = 0;
        for (int j = 0; j < bpt; j++) {
            cnt[j] = hist[(int64_t)(threadIdx.x * bpt + j) * nt + b];
            mine += cnt[j];
        }
        int64_t total = 0;
        int32_t first = (int32_t)(block_inclusive_scan(mine, s_wave, &total) - mine);
        for (int j = 0; j < bpt; j++) {
            const int d = threadIdx.x * bpt + j;
            s_first[d] = first;
            s_out[d] = offs[(int64_t)d * nt + b] - first;
            s_run[d] = 0;
            first += cnt[j];
        }
        const int64_t base = b * kTile;
        const int tile_n = (int)(n - base < kTile ? n - base : kTile);
        for (int r = 0; r < kSortRounds; r++) {
            for (int k = 0; k < kWaves; k++)
                for (int d = threadIdx.x; d < nbins; d += kSortBlock) s_wcnt[k][d] = 0;
            __syncthreads();
            const int li = r * kSortBlock + threadIdx.x;
            const bool valid = li < tile_n;
            uint32_t key = 0, val = 0;
            if (valid) {
                key = keys[base + li];
                val = vals[base + li];
            }
            const uint32_t d = digit_of(key, shift, dmask);
            // lanes of this wave holding the same digit
            unsigned long long peers = __ballot(valid);
            for (int bit = 0; bit < dbits; bit++) {
                const bool set = (d >> bit) & 1u;
                const unsigned long long m = __ballot(set);
                peers &= set ? m : ~m;
            }
            const int rank = __popcll(peers & below);
            if (valid && rank == 0) s_wcnt[w][d] = __popcll(peers);
            __syncthreads();
            if (valid) {
                int place = s_first[d] + s_run[d] + rank;
                for (int k = 0; k < w; k++) place += s_wcnt[k][d];
                s_key[place] = key;
                s_val[place] = val;
            }
            __syncthreads();
            for (int dd = threadIdx.x; dd < nbins; dd += kSortBlock) {
                int add = 0;
                for (int k = 0; k < kWaves; k++) add += s_wcnt[k][dd];
                s_run[dd] += add;
            }
        }
        __syncthreads();
        for (int p = threadIdx.x; p < tile_n; p += kSortBlock) {
            const uint32_t key = s_key[p];
            const int64_t pos = (int64_t)s_out[digit_of(key, shift, dmask)] + p;
            keys_out[pos] = key;
            vals_out[pos] = s_val[p];
        }
        __syncthreads();  // the staging arrays are refilled for the next tile
    }
}

}  // namespace

// Blocks per launch: enough to fill the device, few enough that a small count leaves few
// blocks with nothing to do (the count is not known on the host).
constexpr int64_t kMaxGrid = 2048;

void dev_scan_i32(vt_ctx* ctx, const int32_t* in, int32_t* out, const int64_t* d_n, int64_t n_max, int64_t* d_total,
                  const char* part_key)
{
    const int64_t nb_max = std::max<int64_t>(1, (n_max + kScanChunk - 1) / kScanChunk);
    const unsigned grid = (unsigned)std::min(nb_max, kMaxGrid);
    int64_t* part = (int64_t*)scratch(ctx, part_key, 8 * (size_t)nb_max);
    VT_LAUNCH(ctx, "dscan_reduce", k_dscan_reduce, grid, kScanBlock, 0, in, d_n, part);
    VT_LAUNCH(ctx, "dscan_parts", k_dscan_parts, 1, kPartBlock, 0, d_n, part, d_total);
    VT_LAUNCH(ctx, "dscan_down", k_dscan_down, grid, kScanBlock, 0, in, d_n, (const int64_t*)part, out);
}

bool dev_sort_pairs_u32(vt_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                        const int64_t* d_count, int64_t mult, int64_t n_max, int bits)
{
    if (n_max >= (int64_t(1) << 31))
        fail(ctx, VT_ERR_LIMIT, "device sort of up to %lld pairs: positions must fit 31 bits", (long long)n_max);
    const int64_t nt_max = std::max<int64_t>(1, (n_max + kTile - 1) / kTile);
    const unsigned grid = (unsigned)std::min(nt_max, kMaxGrid);
    // 8-bit digits up to 24 key bits (three passes at most), else 9-bit digits: 25-27 bits
    // (the move events of up to 2^27 vertices) in three passes instead of four
    const int dbits = bits > 24 ? kMaxDigitBits : 8;
    const int nbins = 1 << dbits;
    int64_t* sizes = (int64_t*)scratch(ctx, "dsort_sizes", 8 * 4);
    int32_t* hist = (int32_t*)scratch(ctx, "dsort_hist", 4 * (size_t)nbins * (size_t)nt_max);
    int32_t* offs = (int32_t*)scratch(ctx, "dsort_offs", 4 * (size_t)nbins * (size_t)nt_max);
    VT_LAUNCH(ctx, "dsort_sizes", k_dsort_sizes, 1, 64, 0, d_count, mult, nbins, sizes);
    bool alt = false;
    for (int shift = 0; shift < bits; shift += dbits) {
        const uint32_t dmask = bits - shift >= dbits ? (uint32_t)nbins - 1u : (1u << (bits - shift)) - 1u;
        uint32_t* ki = alt ? keys_alt : keys;
        uint32_t* vi = alt ? vals_alt : vals;
        uint32_t* ko = alt ? keys : keys_alt;
        uint32_t* vo = alt ? vals : vals_alt;
        VT_LAUNCH(ctx, "dsort_hist", k_dsort_hist, grid, kSortBlock, 0, (const uint32_t*)ki,
                  (const int64_t*)sizes, shift, dmask, nbins, hist);
        dev_scan_i32(ctx, hist, offs, sizes + 2, (int64_t)nbins * nt_max, nullptr, "dsort_part");
        VT_LAUNCH(ctx, "dsort_scatter", k_dsort_scatter, grid, kSortBlock, 0, (const uint32_t*)ki,
                  (const uint32_t*)vi, (const int64_t*)sizes, shift, dmask, dbits, (const int32_t*)hist,
                  (const int32_t*)offs, ko, vo);
        alt = !alt;
    }
    return alt;
}

}  // namespace vt

using namespace vt;

extern "C" {

int vt_dev_sort_pairs(vt_ctx* ctx, int64_t n, int64_t capacity, int32_t bits, const uint32_t* keys,
                      const uint32_t* vals, uint32_t* keys_out, uint32_t* vals_out)
{
    if (!ctx) return VT_ERR_ARG;
    return guarded(ctx, [&] {
        if (n < 0 || capacity < n || bits < 1 || bits > 32 || (n > 0 && (!keys || !vals || !keys_out || !vals_out)))
            fail(ctx, VT_ERR_ARG, "vt_dev_sort_pairs: bad arguments (n=%lld capacity=%lld bits=%d)", (long long)n,
                 (long long)capacity, bits);
        VT_HIP(ctx, hipSetDevice(ctx->device));
        const size_t c1 = (size_t)std::max<int64_t>(capacity, 1);
        DBuf<uint32_t> k0(ctx, c1), v0(ctx, c1), k1(ctx, c1), v1(ctx, c1);
        DBuf<int64_t> dn(ctx, 1);
        h2d(ctx, k0.p, keys, 4 * (size_t)n);
        h2d(ctx, v0.p, vals, 4 * (size_t)n);
        h2d(ctx, dn.p, &n, 8);
        const bool alt = dev_sort_pairs_u32(ctx, k0.p, v0.p, k1.p, v1.p, dn.p, 1, capacity, bits);
        d2h(ctx, keys_out, alt ? k1.p : k0.p, 4 * (size_t)n);
        d2h(ctx, vals_out, alt ? v1.p : v0.p, 4 * (size_t)n);
        sync(ctx);
    });
}

int vt_dev_scan(vt_ctx* ctx, int64_t n, int64_t capacity, const int32_t* in, int32_t* out, int64_t* total_out)
{
    if (!ctx) return VT_ERR_ARG;
    return guarded(ctx, [&] {
        if (n < 0 || capacity < n || (n > 0 && (!in || !out)))
            fail(ctx, VT_ERR_ARG, "vt_dev_scan: bad arguments (n=%lld capacity=%lld)", (long long)n,
                 (long long)capacity);
        VT_HIP(ctx, hipSetDevice(ctx->device));
        const size_t c1 = (size_t)std::max<int64_t>(capacity, 1);
        DBuf<int32_t> a(ctx, c1), b(ctx, c1);
        DBuf<int64_t> dn(ctx, 2);
        h2d(ctx, a.p, in, 4 * (size_t)n);
        h2d(ctx, dn.p, &n, 8);
        dev_scan_i32(ctx, a.p, b.p, dn.p, capacity, dn.p + 1, "dscan_test_part");
        d2h(ctx, out, b.p, 4 * (size_t)n);
        if (total_out) d2h(ctx, total_out, dn.p + 1, 8);
        sync(ctx);
    });
}

}  // extern "C"
