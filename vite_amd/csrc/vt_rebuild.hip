// Coarsening: distbuildNextLevelGraph (rebuild.cpp:480-504) at world size 1.
//
//  1. renumber: new id of community c = rank of its first appearance in cvect as the
//     vertex index ascends (distReNumber, rebuild.cpp:121-135) -- atomicMin of the first
//     vertex per community, a flag scan over vertices, one gather;
//  2. coarse edge (C1, C2) weight = left fold over the vertices of C1 ascending, each row in
//     CSR order (fill_newEdgesMap, rebuild.cpp:294-329), the first contribution as is: one
//     wave per community walks its members' rows in that order and folds the self-loop
//     weight and, in an LDS table, every external (C1, C2) entry; communities with too many
//     distinct neighbours for the table take a stable radix sort of their external entries
//     by (C1, C2) whose runs are folded left to right;
//  3. rows sorted by C2 (std::map order, rebuild.cpp:469-476), the self-loop entry placed at
//     its column.  Self-loops are kept.
#include <hip/hip_runtime.h>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include <algorithm>
#include <climits>
#include <vector>

#include "vt_internal.hpp"

namespace vt {

__global__ void k_first_vertex(const int64_t* __restrict__ cvect, int64_t nv, unsigned long long* __restrict__ first,
                               unsigned long long* __restrict__ bad)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = cvect[i];
        if (c < 0 || c >= nv) { atomicAdd(bad, 1ull); continue; }
        // vertex i - 1 in the same community: i is not the community's first vertex (on
        // graphs numbered by locality most vertices follow one of their own community)
        if (i > 0 && cvect[i - 1] == c) continue;
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

// ---- per-community fold --------------------------------------------------------------------
// The members of each community (a stable sort of vertices by new id) in vertex order, each
// member's row in CSR order, form one concatenated row per community: exactly the order in
// which rebuild.cpp:303-327 adds the community's fine edges into its coarse row.  One wave
// walks a community's concatenated row once, 64 entries at a time with coalesced loads, and
// folds every coarse entry (C2 = new id of the tail, the community itself included) into a
// per-wave LDS table as the local move folds its counters: lanes insert their key by CAS,
// mark their lane in the slot's mask, and the lowest lane of each key adds the chunk's
// weights for it in lane (= row) order onto the slot's value, which starts at -0.0
// (-0.0 + w == w, so each entry is the left fold with the first weight as is).  The coarse
// row, the table's keys in ascending order, goes to a staging area at the community's
// offset in the concatenated rows, which is at least as long.  A community with more than
// kCfMaxKeys distinct keys takes a stable radix sort of its entries by (C1, C2) instead,
// whose runs are folded left to right.

// Each community's first position among the members sorted by community: every community
// id below nc has a member, so cstart[c] is where the run of c begins, and cstart[nc] = nv.
__global__ void k_member_starts(const uint32_t* __restrict__ keys, int64_t nv, int64_t nc, int64_t* __restrict__ cstart)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) cstart[nc] = nv;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nv; p += (int64_t)gridDim.x * blockDim.x)
        if (p == 0 || keys[p] != keys[p - 1]) cstart[keys[p]] = p;
}

__global__ void k_iota_u32(uint32_t* __restrict__ p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)i;
}

// Each member's row length (for the scan into concatenated-row offsets) and row start.
__global__ void k_member_degrees(const uint32_t* __restrict__ mem, int64_t nv, const int64_t* __restrict__ offs,
                                 int64_t* __restrict__ mdeg, int64_t* __restrict__ mrow)
{
    for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m <= nv; m += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r0 = m < nv ? offs[mem[m]] : 0;
        mdeg[m] = m < nv ? offs[mem[m] + 1] - r0 : 0;
        if (m < nv) mrow[m] = r0;
    }
}

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Positions of community c's concatenated row -> CSR indices (moff: each member's start in
// its community's concatenated row; mrow: its row start in the CSR).  Up to 64 members
// (whose rows total under 2^31 entries), lane i holds member i's start relative to the
// community's (32 bits) and its row start minus its start, and lanes find their member by a
// binary search over shuffles that starts at the member count's highest power of two;
// larger communities search moff.  c is the same in every lane, so the community's bounds
// come in through scalar loads.
struct MemberRows {
    int64_t m0, m1, s0, s1, my_base = 0;
    int32_t my_rel = INT32_MAX, top = 0;
    bool small;

    __device__ MemberRows(const int64_t* __restrict__ cstart, const int64_t* __restrict__ moff,
                          const int64_t* __restrict__ mrow, int64_t c, int lane)
    {
        m0 = cstart[c];
        m1 = cstart[c + 1];
        s0 = moff[m0];
        s1 = moff[m1];
        small = m1 - m0 <= 64 && s1 - s0 < INT32_MAX;
        if (small) {
            const int nm = (int)(m1 - m0);
            top = nm > 1 ? 1 << (31 - __builtin_clz((unsigned)(nm - 1))) : 0;
            if (lane < nm) {
                const int64_t ms = moff[m0 + lane];
                my_rel = (int32_t)(ms - s0);
                my_base = mrow[m0 + lane] - ms;
            }
        }
    }

    // CSR index of position p (p < s1); every lane of the wave calls it.
    __device__ int64_t csr_index(int64_t p, const int64_t* __restrict__ moff, const int64_t* __restrict__ mrow) const
    {
        if (small) {
            const int nm = (int)(m1 - m0);
            const int64_t q64 = p - s0;
            const int32_t q = q64 < INT32_MAX ? (int32_t)q64 : INT32_MAX;
            int lo = 0;
            for (int step = top; step > 0; step >>= 1) {
                const int cand = lo + step;
                const int32_t cs = __shfl(my_rel, cand < 64 ? cand : 63);
                if (cand < nm && cs <= q) lo = cand;
            }
            return p + __shfl(my_base, lo);
        }
        if (p >= s1) return 0;
        int64_t lo = m0, hi = m1 - 1;  // the member whose row holds position p
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (moff[mid] <= p) lo = mid;
            else hi = mid - 1;
        }
        return mrow[lo] + (p - moff[lo]);
    }
};

constexpr int kCfWaves = 4;
// The fold's (and the overflow emission's) blocks are one wave each: a block holds its
// resources until its last wave ends, and the waves of a block take communities of very
// different sizes, so 4-wave blocks left about half of the resident-wave slots idle.
constexpr int kFoldWaves = 1;
constexpr int kCfLog = 8;
constexpr int kCfCap = 1 << kCfLog;       // table slots per wave
constexpr int kCfMaxKeys = 128;           // distinct keys a community may hold in the table
constexpr int kCfList = kCfMaxKeys + 64;  // a chunk adds at most 64 keys before the check
constexpr uint32_t kNoKey = 0xffffffffu;

// Per community c: its coarse row staged at rkey/rval[s0, s0 + rowlen[c]), with s0 its
// offset in the concatenated rows; or, with more distinct keys than the table holds,
// rowlen[c] = 0 and xcnt[c] = its entry count, for the sort path.
__global__ void __launch_bounds__(64 * kFoldWaves) k_comm_fold(
    int64_t nc, const int64_t* __restrict__ cstart, const int64_t* __restrict__ moff, const int64_t* __restrict__ mrow,
    const DEdge* __restrict__ edges, const uint32_t* __restrict__ newid,
    uint32_t* __restrict__ rkey, double* __restrict__ rval, int64_t* __restrict__ rowlen, int64_t* __restrict__ xcnt)
{
    __shared__ uint32_t tk[kFoldWaves][kCfCap];
    __shared__ double tv[kFoldWaves][kCfCap];
    __shared__ unsigned long long tm[kFoldWaves][kCfCap];
    __shared__ uint32_t tl[kFoldWaves][kCfList];  // keys in insertion order
    __shared__ uint16_t ts[kFoldWaves][kCfList];  // and their slots
    __shared__ double wst[kFoldWaves][65];        // [0] = -0.0, then the chunk's weights
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const unsigned long long below = (1ull << lane) - 1;
    for (int i = lane; i < kCfCap; i += 64) {
        tk[w][i] = kNoKey;
        tv[w][i] = -0.0;
        tm[w][i] = 0;
    }
    if (lane == 0) wst[w][0] = -0.0;
    wave_lds_sync();
    for (int64_t c = (int64_t)blockIdx.x * kFoldWaves + w; c < nc; c += (int64_t)gridDim.x * kFoldWaves) {
        const MemberRows mr(cstart, moff, mrow, c, lane);
        int nd = 0;  // distinct external keys so far (the table never holds more than kCfList)
        double self = -0.0;  // the internal entries' fold, the same in every lane
        bool any_self = false;
        // This lane's entry of the current chunk and its coarse column, and the next chunk's
        // entry: a chunk's column reads and the entries of the chunk after it are in flight
        // while the chunk before is folded, so a chunk waits on no load of its own.
        DEdge e{}, en{};
        uint32_t c2 = 0;
        if (mr.s0 < mr.s1) {
            const int64_t k = mr.csr_index(mr.s0 + lane, moff, mrow);
            if (mr.s0 + lane < mr.s1) e = edges[k];
            if (mr.s0 + 64 < mr.s1) {
                const int64_t kn = mr.csr_index(mr.s0 + 64 + lane, moff, mrow);
                if (mr.s0 + 64 + lane < mr.s1) en = edges[kn];
            }
            if (mr.s0 + lane < mr.s1) c2 = newid[e.tail];
        }
        for (int64_t base = mr.s0; base < mr.s1 && nd <= kCfMaxKeys; base += 64) {
            const int64_t p = base + lane;
            const bool valid = p < mr.s1;
            const double wt = valid ? e.weight : -0.0;
            uint32_t c2n = 0;
            if (p + 64 < mr.s1) c2n = newid[en.tail];
            DEdge enn{};
            if (base + 128 < mr.s1) {
                const int64_t kn = mr.csr_index(p + 128, moff, mrow);
                if (p + 128 < mr.s1) enn = edges[kn];
            }
            // internal entries (about half after a phase) form one long run per chunk: add
            // them in lane order with the lane read as a scalar, without the table
            const bool in = valid && c2 == (uint32_t)c;
            const bool ext = valid && !in;
            const unsigned long long im = __ballot(in);
            any_self = any_self || im != 0;
            for (unsigned long long mm = im; mm; mm &= mm - 1) self = self + readlane_f64(wt, __builtin_ctzll(mm));
            int slot = 0;
            bool fresh = false;
            if (ext) {
                uint32_t h = (c2 * 0x9E3779B1u) >> (32 - kCfLog);
                for (;;) {  // ends: at most kCfList < kCfCap keys are ever present
                    const uint32_t old = atomicCAS(&tk[w][h], kNoKey, c2);
                    if (old == kNoKey || old == c2) {
                        fresh = old == kNoKey;
                        break;
                    }
                    h = (h + 1) & (kCfCap - 1);
                }
                slot = (int)h;
            }
            const unsigned long long fm = __ballot(fresh);
            if (fresh) {
                const int idx = nd + __popcll(fm & below);
                tl[w][idx] = c2;
                ts[w][idx] = (uint16_t)slot;
            }
            nd += __popcll(fm);
            wst[w][1 + lane] = wt;
            if (ext) atomicOr(&tm[w][slot], 1ull << lane);
            wave_lds_sync();
            const unsigned long long m = ext ? tm[w][slot] : 0ull;
            wave_lds_sync();
            if (ext && (m & below) == 0) {
                tv[w][slot] = fold_rest<unsigned long long, 2>(tv[w][slot], m, &wst[w][1]);
                tm[w][slot] = 0;
            }
            wave_lds_sync();
            e = en;
            en = enn;
            c2 = c2n;
        }
        if (nd <= kCfMaxKeys) {
            // each external key's column is its rank among the (distinct) keys, one further
            // when the self-loop entry sorts before it
            int lt = 0;  // external keys below c
            for (int i0 = 0; i0 < nd; i0 += 64) {
                const int i = i0 + lane;
                const uint32_t ki = i < nd ? tl[w][i] : kNoKey;
                lt += __popcll(__ballot(ki < (uint32_t)c));
                if (i < nd) {
                    int rank = 0;
                    for (int j = 0; j < nd; j++) rank += tl[w][j] < ki ? 1 : 0;
                    const int64_t q = mr.s0 + rank + ((any_self && ki > (uint32_t)c) ? 1 : 0);
                    rkey[q] = ki;
                    rval[q] = tv[w][ts[w][i]];
                }
            }
            if (lane == 0) {
                if (any_self) {
                    rkey[mr.s0 + lt] = (uint32_t)c;
                    rval[mr.s0 + lt] = self;
                }
                rowlen[c] = nd + (any_self ? 1 : 0);
                xcnt[c] = 0;
            }
        } else if (lane == 0) {
            rowlen[c] = 0;
            xcnt[c] = mr.s1 - mr.s0;
        }
        wave_lds_sync();
        for (int i = lane; i < nd; i += 64) {
            const int sl = ts[w][i];
            tk[w][sl] = kNoKey;
            tv[w][sl] = -0.0;
        }
        wave_lds_sync();
    }
}

// Every entry of the communities that overflowed the table, as ((C1 << bits) | C2, weight),
// at xoff[C1] + its position in C1's concatenated row.
// Overflowing communities of more than this many entries (a giant community on a skewed
// graph) are emitted by every wave of k_ovf_emit_big instead of one wave each: emitting
// needs no order, and one wave walking such a community held the launch for 4 ms on BA
// n = 2^22.
constexpr int64_t kOvfBig = 1 << 16;

// The ids of those communities, in any order (their entries' places do not depend on it).
__global__ void k_ovf_big_list(int64_t nc, const int64_t* __restrict__ xcnt, unsigned long long* __restrict__ count,
                               int64_t* __restrict__ list)
{
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
        if (xcnt[c] > kOvfBig) list[atomicAdd(count, 1ull)] = c;
}

__global__ void __launch_bounds__(64 * kFoldWaves) k_ovf_emit(
    int64_t nc, const int64_t* __restrict__ cstart, const int64_t* __restrict__ moff, const int64_t* __restrict__ mrow,
    const DEdge* __restrict__ edges, const uint32_t* __restrict__ newid,
    const int64_t* __restrict__ xcnt, const int64_t* __restrict__ xoff, int bits, unsigned long long* __restrict__ keys,
    double* __restrict__ vals)
{
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int64_t c = (int64_t)blockIdx.x * kFoldWaves + w; c < nc; c += (int64_t)gridDim.x * kFoldWaves) {
        if (xcnt[c] == 0 || xcnt[c] > kOvfBig) continue;  // k_ovf_emit_big takes the giant ones
        const MemberRows mr(cstart, moff, mrow, c, lane);
        const int64_t o = xoff[c] - mr.s0;
        for (int64_t base = mr.s0; base < mr.s1; base += 64) {
            const int64_t p = base + lane;
            const int64_t k = mr.csr_index(p, moff, mrow);
            if (p < mr.s1) {
                const DEdge e = edges[k];
                keys[o + p] = ((unsigned long long)c << bits) | newid[e.tail];
                vals[o + p] = e.weight;
            }
        }
    }
}

// The giant overflowing communities: for each listed community, the whole grid deals its
// 64-entry chunks to the waves; each lane finds its entry's CSR index by a search of the
// member offsets, as k_ovf_emit does for communities of more than 64 members.
__global__ void __launch_bounds__(64 * kCfWaves) k_ovf_emit_big(
    const unsigned long long* __restrict__ count, const int64_t* __restrict__ list, const int64_t* __restrict__ cstart,
    const int64_t* __restrict__ moff, const int64_t* __restrict__ mrow, const DEdge* __restrict__ edges,
    const uint32_t* __restrict__ newid, const int64_t* __restrict__ xoff, int bits, unsigned long long* __restrict__ keys,
    double* __restrict__ vals)
{
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t wave = (int64_t)blockIdx.x * kCfWaves + w, waves = (int64_t)gridDim.x * kCfWaves;
    const int64_t nbig = (int64_t)*count;
    for (int64_t i = 0; i < nbig; i++) {
        const int64_t c = list[i];
        const MemberRows mr(cstart, moff, mrow, c, lane);
        const int64_t o = xoff[c] - mr.s0;
        const int64_t nchunks = (mr.s1 - mr.s0 + 63) / 64;
        for (int64_t ch = wave; ch < nchunks; ch += waves) {
            const int64_t p = mr.s0 + ch * 64 + lane;
            const int64_t k = mr.csr_index(p, moff, mrow);
            if (p < mr.s1) {
                const DEdge e = edges[k];
                keys[o + p] = ((unsigned long long)c << bits) | newid[e.tail];
                vals[o + p] = e.weight;
            }
        }
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

// An overflowing community's first run and run count (= its coarse row length); runs are
// sorted by C1, then C2 -- binary searches, no atomics.
__global__ void k_ovf_rows(int64_t nc, const int64_t* __restrict__ xcnt, const uint32_t* __restrict__ rc1, int64_t nu,
                           int64_t* __restrict__ run_first, int64_t* __restrict__ rowlen)
{
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
        if (xcnt[c] == 0) continue;
        const int64_t f = lower_bound_u32(rc1, 0, nu, (uint32_t)c);
        run_first[c] = f;
        rowlen[c] = lower_bound_u32(rc1, f, nu, (uint32_t)c + 1) - f;
    }
}

// Run r of community C1 goes to its rank among C1's runs in C1's staged row.
__global__ void k_ovf_place(const uint32_t* __restrict__ rc1, const uint32_t* __restrict__ rc2,
                            const double* __restrict__ rw, int64_t nu, const int64_t* __restrict__ cstart,
                            const int64_t* __restrict__ moff, const int64_t* __restrict__ run_first,
                            uint32_t* __restrict__ rkey, double* __restrict__ rval)
{
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nu; r += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t c1 = rc1[r];
        const int64_t q = moff[cstart[c1]] + (r - run_first[c1]);
        rkey[q] = rc2[r];
        rval[q] = rw[r];
    }
}

// Staged rows -> the coarse CSR, one wave per community.
__global__ void __launch_bounds__(64 * kCfWaves) k_row_copy(int64_t nc, const int64_t* __restrict__ cstart,
                                                           const int64_t* __restrict__ moff,
                                                           const int64_t* __restrict__ rowlen,
                                                           const int64_t* __restrict__ out_offs,
                                                           const uint32_t* __restrict__ rkey,
                                                           const double* __restrict__ rval, DEdge* __restrict__ out)
{
    // 16 lanes per community (coarse rows average a few dozen entries)
    constexpr int kLanes = 16;
    const int sub = threadIdx.x % kLanes;
    const int64_t groups = (int64_t)gridDim.x * blockDim.x / kLanes;
    for (int64_t c = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kLanes; c < nc; c += groups) {
        const int64_t s0 = moff[cstart[c]], len = rowlen[c], o = out_offs[c];
        for (int64_t i = sub; i < len; i += kLanes) {
            DEdge e;
            e.tail = (int64_t)rkey[s0 + i];
            e.weight = rval[s0 + i];
            out[o + i] = e;
        }
    }
}

__global__ void k_pair_totals(const int64_t* __restrict__ a, const int64_t* __restrict__ b, int64_t* __restrict__ out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        out[0] = *a;
        out[1] = *b;
    }
}

// Overflowing communities: their entries sorted stably by (C1, C2), runs folded left to
// right, and each community's runs placed in its staged row.
static void overflow_rows(vt_ctx* ctx, const vt_graph* g, const uint32_t* newid, int64_t nc, int bits,
                          const int64_t* cstart, const int64_t* mrow, const int64_t* moff, const int64_t* xcnt,
                          const int64_t* xoff, int64_t nx, uint32_t* rkey, double* rval, int64_t* rowlen)
{
    const size_t x1 = (size_t)std::max<int64_t>(nx, 1), c1n = (size_t)nc + 1;
    DBuf<unsigned long long> k0(ctx, x1), k1(ctx, x1);
    DBuf<double> v0(ctx, x1), v1(ctx, x1);
    VT_LAUNCH(ctx, "rb_ovf_emit", k_ovf_emit, grid_for(nc, kFoldWaves, 262144), 64 * kFoldWaves, 0, nc, cstart, moff, mrow,
              (const DEdge*)g->edges, newid, xcnt, xoff, bits, k0.p, v0.p);
    if (nx > kOvfBig) {  // a community that big may exist
        DBuf<unsigned long long> nbig(ctx, 1);
        DBuf<int64_t> big(ctx, (size_t)std::max<int64_t>(nx / kOvfBig, 1));
        dzero(ctx, nbig.p, 8);
        VT_LAUNCH(ctx, "rb_ovf_big_list", k_ovf_big_list, grid_for(nc, 256, 4096), 256, 0, nc, xcnt, nbig.p, big.p);
        VT_LAUNCH(ctx, "rb_ovf_emit_big", k_ovf_emit_big, 1024, 64 * kCfWaves, 0, (const unsigned long long*)nbig.p,
                  (const int64_t*)big.p, cstart, moff, mrow, (const DEdge*)g->edges, newid, xoff, bits, k0.p, v0.p);
    }
    const unsigned long long* ks = k0.p;
    const double* vs = v0.p;
    if (nx > 1) {
        rocprim::double_buffer<unsigned long long> kb(k0.p, k1.p);
        rocprim::double_buffer<double> vb(v0.p, v1.p);
        // Many entries per community (the coarse levels, where nearly every community
        // overflows): each community's entries, already together at xoff[C1], are sorted by C2
        // alone, a workgroup each; otherwise one sort of every entry by (C1, C2).  Both are
        // stable, so the order is the same.
        const bool seg = nx >= 16 * nc;
        auto sort = [&](void* tmp, size_t& tb) {
            if (seg)
                return rocprim::segmented_radix_sort_pairs(tmp, tb, kb, vb, (unsigned)nx, (unsigned)nc, xoff, xoff + 1, 0u,
                                                           (unsigned)bits, ctx->stream);
            return rocprim::radix_sort_pairs(tmp, tb, kb, vb, (size_t)nx, 0, 2 * bits, ctx->stream);
        };
        size_t tb = 0;
        VT_HIP(ctx, sort(nullptr, tb));
        void* tmp = scratch(ctx, "sort_tmp", tb);
        {
            TimedScope ts(ctx, "rb_sort");
            VT_HIP(ctx, sort(tmp, tb));
        }
        ks = kb.current();
        vs = vb.current();
    }
    int64_t nu = 0;
    DBuf<int64_t> starts, tot(ctx, 2);
    {
        DBuf<int64_t> head(ctx, x1), hpos(ctx, x1);
        VT_LAUNCH(ctx, "rb_heads", k_run_heads, grid_for(nx, 256, 16384), 256, 0, ks, nx, head.p);
        exclusive_scan_i64(ctx, head.p, hpos.p, nx);
        VT_LAUNCH(ctx, "rb_total", k_total, 1, 64, 0, (const int64_t*)head.p, (const int64_t*)hpos.p, nx, tot.p);
        d2h(ctx, &nu, tot.p, 8);
        starts = DBuf<int64_t>(ctx, (size_t)nu + 1);
        VT_LAUNCH(ctx, "rb_starts", k_run_starts, grid_for(nx, 256, 16384), 256, 0, (const int64_t*)head.p,
                  (const int64_t*)hpos.p, nx, nu, starts.p);
    }
    DBuf<uint32_t> rc1(ctx, (size_t)nu), rc2(ctx, (size_t)nu);
    DBuf<double> rw(ctx, (size_t)nu);
    VT_LAUNCH(ctx, "rb_fold", k_ext_runs, grid_for(nu, 256, 16384), 256, 0, ks, vs, (const int64_t*)starts.p, nu, bits,
              rc1.p, rc2.p, rw.p);
    DBuf<int64_t> run_first(ctx, c1n);
    VT_LAUNCH(ctx, "rb_ovf_rows", k_ovf_rows, grid_for(nc, 256, 16384), 256, 0, nc, xcnt, (const uint32_t*)rc1.p, nu,
              run_first.p, rowlen);
    VT_LAUNCH(ctx, "rb_ovf_place", k_ovf_place, grid_for(nu, 256, 16384), 256, 0, (const uint32_t*)rc1.p,
              (const uint32_t*)rc2.p, (const double*)rw.p, nu, cstart, moff, (const int64_t*)run_first.p, rkey, rval);
}

static vt_graph* rebuild_fold(vt_ctx* ctx, const vt_graph* g, const uint32_t* newid, int64_t nc)
{
    const int64_t nv = g->nv, ne = g->ne;
    const int bits = (int)bits_for(nc > 1 ? nc - 1 : 1);
    const size_t c1n = (size_t)nc + 1, n1 = (size_t)std::max<int64_t>(nv, 1), e1 = (size_t)std::max<int64_t>(ne, 1);

    // members of each community in ascending vertex order (stable sort by new id), and each
    // member's start in its community's concatenated row
    DBuf<int64_t> cstart(ctx, c1n), moff(ctx, n1 + 1), mrow(ctx, n1);
    {
        DBuf<uint32_t> mk(ctx, n1), mv(ctx, n1), mk2(ctx, n1), mv2(ctx, n1);
        d2d(ctx, mk.p, newid, 4 * (size_t)nv);
        VT_LAUNCH(ctx, "rb_iota", k_iota_u32, grid_for(nv, 256, 8192), 256, 0, mv.p, nv);
        bool malt = false;
        sort_pairs_u32(ctx, mk.p, mv.p, nv, bits, mk2.p, mv2.p, &malt);
        VT_LAUNCH(ctx, "rb_member_starts", k_member_starts, grid_for(nv, 256, 8192), 256, 0,
                  (const uint32_t*)(malt ? mk2.p : mk.p), nv, nc, cstart.p);
        DBuf<int64_t> mdeg(ctx, n1 + 1);
        VT_LAUNCH(ctx, "rb_mdeg", k_member_degrees, grid_for(nv + 1, 256, 8192), 256, 0,
                  (const uint32_t*)(malt ? mv2.p : mv.p), nv, (const int64_t*)g->offs, mdeg.p, mrow.p);
        exclusive_scan_i64(ctx, mdeg.p, moff.p, nv + 1);
    }

    // every community's coarse row, staged at its offset in the concatenated rows
    DBuf<uint32_t> rkey(ctx, e1);
    DBuf<double> rval(ctx, e1);
    DBuf<int64_t> rowlen(ctx, c1n), xcnt(ctx, c1n), out_offs(ctx, c1n), xoff(ctx, c1n), tot(ctx, 3);
    dzero(ctx, rowlen.p + nc, 8);
    dzero(ctx, xcnt.p + nc, 8);
    if (nc)
        VT_LAUNCH(ctx, "rb_comm_fold", k_comm_fold, grid_for(nc, kFoldWaves, 262144), 64 * kFoldWaves, 0, nc,
                  (const int64_t*)cstart.p, (const int64_t*)moff.p, (const int64_t*)mrow.p, (const DEdge*)g->edges, newid,
                  rkey.p, rval.p, rowlen.p, xcnt.p);
    exclusive_scan_i64(ctx, rowlen.p, out_offs.p, nc + 1);
    exclusive_scan_i64(ctx, xcnt.p, xoff.p, nc + 1);
    VT_LAUNCH(ctx, "rb_totals", k_pair_totals, 1, 64, 0, (const int64_t*)out_offs.p + nc, (const int64_t*)xoff.p + nc,
              tot.p);
    // the coarse graph's largest row is read with its entry count (the overflowing
    // communities' rows are added below, and the count and the largest row read again)
    auto max_row = [&]() {
        dzero(ctx, tot.p + 2, 8);
        if (nc)
            VT_LAUNCH(ctx, "max_degree", k_max_degree, grid_for(nc, 256, 512), 256, 0, (const int64_t*)out_offs.p, nc,
                      (unsigned long long*)(tot.p + 2));
    };
    max_row();
    int64_t totals[3] = {0, 0, 0};
    d2h(ctx, totals, tot.p, 24);
    if (totals[1]) {
        overflow_rows(ctx, g, newid, nc, bits, cstart.p, mrow.p, moff.p, xcnt.p, xoff.p, totals[1], rkey.p, rval.p,
                      rowlen.p);
        exclusive_scan_i64(ctx, rowlen.p, out_offs.p, nc + 1);
        d2d(ctx, tot.p, out_offs.p + nc, 8);
        max_row();
        d2h(ctx, totals, tot.p, 24);
    }

    vt_graph* out = graph_new(ctx, nc, totals[0]);
    try {
        d2d(ctx, out->offs, out_offs.p, 8 * c1n);
        if (nc)
            VT_LAUNCH(ctx, "rb_row_copy", k_row_copy, grid_for(nc, kCfWaves, 65536), 64 * kCfWaves, 0, nc,
                      (const int64_t*)cstart.p, (const int64_t*)moff.p, (const int64_t*)rowlen.p,
                      (const int64_t*)out_offs.p, (const uint32_t*)rkey.p, (const double*)rval.p, out->edges);
        out->max_degree = totals[2];
    } catch (...) {
        graph_delete(ctx, out);
        throw;
    }
    return out;
}

vt_graph* rebuild(vt_ctx* ctx, const vt_graph* g, const int64_t* d_cvect, bool check_ids)
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
    if (check_ids) {  // a caller's partition; a phase's own has every id in [0, nv)
        unsigned long long nbad = 0;
        d2h(ctx, &nbad, bad.p, 8);
        if (nbad) fail(ctx, VT_ERR_ARG, "vt_rebuild: %llu community ids outside [0, nv)", nbad);
    }
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
    return rebuild_fold(ctx, g, newid.p, nc);
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
        *out = rebuild(ctx, g, src, cvect != nullptr);
        sync(ctx);
    });
}
