// Scan and stable radix sort whose element count lives in device memory.
//
// The batched phase loop replays each Louvain round from one captured HIP graph.  The round's
// event count (two events per vertex that moves) is only known on the device, so the sort of
// the events cannot take its size from the host.  These kernels read the count from device
// memory, size their grids for a host-known capacity and let the blocks past the count exit
// at once.  They record no memset, allocate nothing once their scratch exists, and make no
// host read, so a sequence of them can be captured and replayed.  A single graph's
// iterations use them too, so that an iteration reads nothing back before its modularity.
//
// Sort: least-significant-digit radix sort, 8- or 9-bit digits, tiles of 4096 pairs, each
// block sorting a contiguous run of tiles.  Per pass a histogram kernel counts each block's
// digits (digit-major table: digit d of block b at d * blocks + b), one block per digit scans
// its row and totals it, and a scatter kernel places each pair at its digit's global offset
// for its block plus its rank among the block's pairs of that digit in input order.  Ranks
// come from wave ballots over the digit bits, so equal keys keep their input order (stable),
// which is what the per-community event folds need.
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
// 4096-pair tiles: on 18 M pairs with 24-bit keys a sort takes 428 us against 472 us with
// 2048-pair tiles (longer runs per digit in each tile's writes) and 500 us with 6144
#ifndef VT_SORT_ROUNDS
#define VT_SORT_ROUNDS 16
#endif
constexpr int kSortRounds = VT_SORT_ROUNDS;
constexpr int64_t kTile = (int64_t)kSortBlock * kSortRounds;  // pairs per tile
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

// Sizes of one sort: [0] pairs, [1] tiles, [2] table entries (nbins per block), [3] blocks
// that take part, [4] tiles per block (each block sorts a contiguous run of tiles).
__global__ void k_dsort_sizes(const int64_t* __restrict__ d_count, int64_t mult, int nbins, int64_t grid,
                              int64_t* __restrict__ sizes)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const int64_t n = *d_count * mult;
        const int64_t nt = ceil_div(n, kTile);
        const int64_t per = ceil_div(nt, grid);
        const int64_t nb = per ? ceil_div(nt, per) : 0;
        sizes[0] = n;
        sizes[1] = nt;
        sizes[2] = nbins * nb;
        sizes[3] = nb;
        sizes[4] = per;
    }
}

// The pass's digit: key bits [shift, shift + digit bits), cut at the sort's last bit (dmask).
__device__ __forceinline__ uint32_t digit_of(uint32_t key, int shift, uint32_t dmask) { return (key >> shift) & dmask; }

// Digit counts of each block's run of tiles, digit-major (digit d of block b at d * nb + b).
__global__ void __launch_bounds__(kSortBlock) k_dsort_hist(const uint32_t* __restrict__ keys,
                                                         const int64_t* __restrict__ sizes, int shift, uint32_t dmask,
                                                         int nbins, int32_t* __restrict__ hist)
{
    __shared__ int32_t s_hist[kMaxBins];
    const int64_t n = sizes[0], nb = sizes[3], per = sizes[4];
    const int64_t b = blockIdx.x;
    if (b >= nb) return;
    for (int d = threadIdx.x; d < nbins; d += kSortBlock) s_hist[d] = 0;
    __syncthreads();
    const int64_t lo = b * per * kTile;
    const int64_t hi = std::min<int64_t>(n, lo + per * kTile);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kSortBlock) atomicAdd(&s_hist[digit_of(keys[i], shift, dmask)], 1);
    __syncthreads();
    for (int d = threadIdx.x; d < nbins; d += kSortBlock) hist[(int64_t)d * nb + b] = s_hist[d];
}

// One block per digit: the exclusive scan of the digit's block counts (offs, same layout as
// hist) and the digit's total (dtot).
__global__ void __launch_bounds__(kSortBlock) k_dsort_offsets(const int64_t* __restrict__ sizes,
                                                            const int32_t* __restrict__ hist,
                                                            int32_t* __restrict__ offs, int32_t* __restrict__ dtot)
{
    __shared__ int64_t s_wave[kWaves];
    const int64_t nb = sizes[3];
    const int d = blockIdx.x;
    const int64_t per = ceil_div(nb, kSortBlock);  // at most kMaxGrid / kSortBlock = 8
    const int64_t lo = (int64_t)threadIdx.x * per;
    const int64_t hi = std::min<int64_t>(nb, lo + per);
    const int32_t* h = hist + (int64_t)d * nb;
    int64_t x = 0;
    for (int64_t i = lo; i < hi; i++) x += h[i];
    int64_t total = 0;
    int64_t run = block_inclusive_scan(x, s_wave, &total) - x;
    for (int64_t i = lo; i < hi; i++) {
        const int32_t c = h[i];
        offs[(int64_t)d * nb + i] = (int32_t)run;
        run += c;
    }
    if (threadIdx.x == 0) dtot[d] = (int32_t)total;
}

// Each block places its run of tiles in order.  Per tile: every wave ranks its 1024 pairs (16
// steps of 64 lanes, in input order) among the wave's pairs of the same digit, from ballots
// over the digit bits and a per-wave running count in LDS; the waves' counts give each
// digit's place in the tile (a scan over digits) and each wave's place inside the digit; the
// pairs are staged in LDS in (digit, input order) and written out with consecutive lanes on
// consecutive positions.  The block's place for each digit starts at the digit's global
// offset (all smaller digits, then the earlier blocks' pairs of the digit) and advances by
// each tile's count.
__global__ void __launch_bounds__(kSortBlock) k_dsort_scatter(const uint32_t* __restrict__ keys,
                                                            const uint32_t* __restrict__ vals,
                                                            const int64_t* __restrict__ sizes, int shift,
                                                            uint32_t dmask, int dbits,
                                                            const int32_t* __restrict__ offs,
                                                            const int32_t* __restrict__ dtot,
                                                            uint32_t* __restrict__ keys_out,
                                                            uint32_t* __restrict__ vals_out)
{
    constexpr int kSteps = kSortRounds;                 // steps of 64 pairs per wave and tile
    __shared__ int32_t s_wrun[kWaves][kMaxBins];        // per wave: running count, then place
    __shared__ int32_t s_outt[kMaxBins];                // output position of the tile's staged index 0
    __shared__ int64_t s_wave[kWaves];
    __shared__ uint32_t s_key[kTile], s_val[kTile];
    const int64_t n = sizes[0], nb = sizes[3], per = sizes[4];
    const int64_t b = blockIdx.x;
    if (b >= nb) return;
    const int nbins = 1 << dbits;
    const int bpt = nbins / kSortBlock;  // bins per thread, consecutive: 1 or 2
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1;

    // the block's global place per digit: smaller digits' totals, then the earlier blocks
    int64_t gbase[kMaxBinsPerThread];
    {
        int64_t mine = 0;
        for (int j = 0; j < bpt; j++) mine += dtot[threadIdx.x * bpt + j];
        int64_t total = 0;
        int64_t run = block_inclusive_scan(mine, s_wave, &total) - mine;
        for (int j = 0; j < bpt; j++) {
            const int d = threadIdx.x * bpt + j;
            gbase[j] = run + offs[(int64_t)d * nb + b];
            run += dtot[d];
            for (int k = 0; k < kWaves; k++) s_wrun[k][d] = 0;
        }
    }
    __syncthreads();
    const int64_t t_end = std::min<int64_t>(sizes[1], (b + 1) * per);
    for (int64_t t = b * per; t < t_end; t++) {
        const int64_t base = t * kTile;
        const int tile_n = (int)(n - base < kTile ? n - base : kTile);
        // (1) ranks inside the wave, in input order
        uint32_t key[kSteps], val[kSteps];
        int32_t rank[kSteps];
        for (int r = 0; r < kSteps; r++) {
            const int li = w * (kSteps * 64) + r * 64 + lane;
            const bool valid = li < tile_n;
            key[r] = valid ? keys[base + li] : 0u;
            val[r] = valid ? vals[base + li] : 0u;
        }
        for (int r = 0; r < kSteps; r++) {
            const int li = w * (kSteps * 64) + r * 64 + lane;
            const bool valid = li < tile_n;
            const uint32_t d = digit_of(key[r], shift, dmask);
            unsigned long long peers = __ballot(valid);
            for (int bit = 0; bit < dbits; bit++) {
                const bool set = (d >> bit) & 1u;
                const unsigned long long m = __ballot(set);
                peers &= set ? m : ~m;
            }
            const int within = __popcll(peers & below);
            const int32_t before = valid ? s_wrun[w][d] : 0;
            rank[r] = before + within;
            // every peer has read the count before its lowest lane moves it on (one wave, in order)
            __builtin_amdgcn_wave_barrier();
            if (valid && within == 0) s_wrun[w][d] = before + __popcll(peers);
            __builtin_amdgcn_wave_barrier();
        }
        __syncthreads();
        // (2) per digit: the tile's count, its first staged place, each wave's place
        {
            int32_t cnt[kMaxBinsPerThread];
            int64_t mine = 0;
            for (int j = 0; j < bpt; j++) {
                const int d = threadIdx.x * bpt + j;
                int32_t c = 0;
                for (int k = 0; k < kWaves; k++) c += s_wrun[k][d];
                cnt[j] = c;
                mine += c;
            }
            int64_t total = 0;
            int32_t first = (int32_t)(block_inclusive_scan(mine, s_wave, &total) - mine);
            for (int j = 0; j < bpt; j++) {
                const int d = threadIdx.x * bpt + j;
                int32_t run = first;
                for (int k = 0; k < kWaves; k++) {
                    const int32_t c = s_wrun[k][d];
                    s_wrun[k][d] = run;
                    run += c;
                }
                s_outt[d] = (int32_t)(gbase[j] - first);
                gbase[j] += cnt[j];
                first += cnt[j];
            }
        }
        __syncthreads();
        // (3) stage in (digit, input order)
        for (int r = 0; r < kSteps; r++) {
            const int li = w * (kSteps * 64) + r * 64 + lane;
            if (li < tile_n) {
                const int place = s_wrun[w][digit_of(key[r], shift, dmask)] + rank[r];
                s_key[place] = key[r];
                s_val[place] = val[r];
            }
        }
        __syncthreads();
        // (4) write out; (5) clear the running counts for the next tile (not read in (4))
        for (int p = threadIdx.x; p < tile_n; p += kSortBlock) {
            const uint32_t k = s_key[p];
            const int64_t pos = (int64_t)s_outt[digit_of(k, shift, dmask)] + p;
            keys_out[pos] = k;
            vals_out[pos] = s_val[p];
        }
        for (int j = 0; j < bpt; j++)
            for (int k = 0; k < kWaves; k++) s_wrun[k][threadIdx.x * bpt + j] = 0;
        __syncthreads();  // staging, s_outt and the cleared counts are reused by the next tile
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
    int64_t* sizes = (int64_t*)scratch(ctx, "dsort_sizes", 8 * 8);
    int32_t* hist = (int32_t*)scratch(ctx, "dsort_hist", 4 * (size_t)nbins * grid);
    int32_t* offs = (int32_t*)scratch(ctx, "dsort_offs", 4 * (size_t)nbins * grid);
    int32_t* dtot = (int32_t*)scratch(ctx, "dsort_dtot", 4 * (size_t)nbins);
    VT_LAUNCH(ctx, "dsort_sizes", k_dsort_sizes, 1, 64, 0, d_count, mult, nbins, (int64_t)grid, sizes);
    bool alt = false;
    for (int shift = 0; shift < bits; shift += dbits) {
        const uint32_t dmask = bits - shift >= dbits ? (uint32_t)nbins - 1u : (1u << (bits - shift)) - 1u;
        uint32_t* ki = alt ? keys_alt : keys;
        uint32_t* vi = alt ? vals_alt : vals;
        uint32_t* ko = alt ? keys : keys_alt;
        uint32_t* vo = alt ? vals : vals_alt;
        VT_LAUNCH(ctx, "dsort_hist", k_dsort_hist, grid, kSortBlock, 0, (const uint32_t*)ki,
                  (const int64_t*)sizes, shift, dmask, nbins, hist);
        VT_LAUNCH(ctx, "dsort_offsets", k_dsort_offsets, (unsigned)nbins, kSortBlock, 0, (const int64_t*)sizes,
                  (const int32_t*)hist, offs, dtot);
        VT_LAUNCH(ctx, "dsort_scatter", k_dsort_scatter, grid, kSortBlock, 0, (const uint32_t*)ki,
                  (const uint32_t*)vi, (const int64_t*)sizes, shift, dmask, dbits, (const int32_t*)offs,
                  (const int32_t*)dtot, ko, vo);
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
