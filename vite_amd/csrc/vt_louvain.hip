// Louvain phase on one MI355X: distLouvainMethod and its eight sibling overloads
// (louvain.cpp:7-1938) as one device-resident loop.
//
// Hot kernel: the local move (distExecuteLouvainIteration + distBuildLocalMapCounter +
// distGetMaxIndex, louvain.cpp:2220-2466).  A group of G lanes owns one vertex at a time;
// vertices are bucketed by degree so that G and the hash-table capacity fit the row.  The
// per-vertex community table is open-addressed in LDS (HBM for hub rows).  Edges of a
// row arrive G at a time (coalesced 16-byte records); every distinct community in a chunk
// elects the lowest lane holding it, and that lane folds the chunk's weights for the
// community in edge order onto the running value.  This reproduces the reference's left
// fold in CSR order exactly, so the gains -- and therefore the moves -- are bit-identical.
//
// All FP64 here must be compiled with -ffp-contract=off: the reference build forms no
// FMA, and a contracted gain can flip a near-tie.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include <algorithm>
#include <array>
#include <initializer_list>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "vt_internal.hpp"

namespace vt {

// ------------------------------------------------------------------------------------------
// Constants and small device helpers
// ------------------------------------------------------------------------------------------

// "No value yet" marker for a table slot: a quiet NaN with a payload no arithmetic makes.
constexpr unsigned long long kSentBits = 0x7ff8dead5eed0001ull;

__device__ __forceinline__ double sent_value() { return __longlong_as_double((long long)kSentBits); }
__device__ __forceinline__ bool is_sent(double x) { return (unsigned long long)__double_as_longlong(x) == kSentBits; }

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t slot_hash(int32_t c) { return (uint32_t)c * 0x9E3779B1u; }

// Degree buckets: (lanes per vertex G, table capacity CAP, max degree).  CAP >= 2*(deg+1)
// keeps the table at most half full (a row has at most deg+1 distinct communities).
constexpr int kNumBuckets = 10;
constexpr int kLpvBucket = 7;      // lane per vertex (degree <= lpv_max)
// The wide bucket (one wave per row, up to 1023 entries) is split by degree so that each row
// gets the smallest table that keeps it at most half full: 512 slots up to degree 255, 1024
// up to 511, else 2048.  Smaller tables leave room for more resident waves per CU.
constexpr int kWide512Bucket = 8;
constexpr int kWide1024Bucket = 9;
constexpr int kLpvMaxDeg = 64;     // the largest row k_local_move_lpv is instantiated for
constexpr int kLpvArena = 2560;
constexpr int kHubBucket = 6;

// Bucket of a row of degree d under per-bucket maximum degrees mx[0..5] (the tuning; the
// default is 7, 15, 31, 63, 127, 1023): rows above mx[5] take the HBM-table kernel.
struct BucketBounds {
    int32_t mx[6];
    int32_t wide[2];  // rows of bucket 5 up to wide[0] take the 512-slot table, up to wide[1] the 1024-slot one
};

__host__ __device__ inline int degree_bucket(int64_t d, const BucketBounds& bb)
{
    for (int b = 0; b < 6; b++)
        if (d <= bb.mx[b]) return b;
    return 6;
}

template <int G> struct MaskT { using type = uint32_t; };
template <> struct MaskT<64> { using type = unsigned long long; };

// Table storage.  LDS tables use plain accesses (one wave owns a group; LDS operations of a
// wave execute in order).  HBM tables use agent-scope atomics so no lane reads a stale L1
// line written by another lane's atomic.
template <bool HBM>
struct TMem {
    template <class T> __device__ static T ld(T* p)
    {
        if constexpr (HBM) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else return *p;
    }
    template <class T> __device__ static void st(T* p, T v)
    {
        if constexpr (HBM) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else *p = v;
    }
    __device__ static double ldd(double* p)
    {
        if constexpr (HBM) {
            unsigned long long b = __hip_atomic_load((unsigned long long*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return __longlong_as_double((long long)b);
        } else return *p;
    }
    __device__ static void std_(double* p, double v)
    {
        if constexpr (HBM)
            __hip_atomic_store((unsigned long long*)p, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        else *p = v;
    }
};

// Per-group table view.
template <int G, bool HBM>
struct Table {
    using mask_t = typename MaskT<G>::type;
    int32_t* keys;   // CAP, -1 = empty
    double* vals;    // CAP, sentinel = empty
    mask_t* masks;   // CAP (LDS) or unused (HBM)
    int32_t* occ;    // CAP: occupied slot list
    double* wst;     // G staged weights (LDS)
    int32_t* sst;    // G staged slots (LDS, HBM path)
    int32_t* nocc;   // 1 (LDS)
    int32_t* cslot;  // 1 (LDS)
    uint32_t capmask;
};

// Community state indexed by vertex id, one 16-byte record per id: the vertex's community
// (currComm) and, for the community whose id is this vertex id, its size and degree
// (localCinfo).  A neighbour's record holds its community id; when that community is the
// neighbour's own id (the neighbour founded it, or still is alone), the community's degree
// comes in the same line, so the candidate's degree read that follows finds it in cache.
struct __align__(16) VRec {
    int32_t comm;
    int32_t size;
    double deg;
};
static_assert(sizeof(VRec) == 16, "record must be 16 bytes");

struct MoveArgs {
    const int64_t* offs;
    const DEdge* edges;
    const int64_t* curr;
    const double* vdeg;
    const double* cdeg;
    const int64_t* csize;
    const unsigned char* active;  // null when ET is off
    double* frozen;               // null when ET is off
    int64_t* target;
    double* cw;
    const double* constant;       // 1/2m: a device scalar, or one per graph of a batch
    unsigned long long* inactive_count;
    // Batch of disjoint graphs: graph of each vertex and whether that graph is still
    // iterating (null for a single graph).
    const int32_t* gid;
    const unsigned char* gact;
    // Structure-of-arrays copy of the edges (32-bit tails, weights) and a 32-bit copy of
    // currComm, read by the group buckets; null when the build reads the records directly.
    const int32_t* tail32;
    const double* wgt;
    const int32_t* curr32;
    // Records of the same community state (null unless the launch tuning asks for them); the
    // group kernels then read communities, sizes and degrees from them.
    const VRec* rec;
};

// Inserts or finds community c; returns its slot.
template <int G, bool HBM>
__device__ __forceinline__ int32_t table_slot(const Table<G, HBM>& t, int32_t c)
{
    uint32_t h = (slot_hash(c) >> 7) & t.capmask;
    for (;;) {
        const int32_t k = TMem<HBM>::ld(&t.keys[h]);
        if (k == c) return (int32_t)h;
        if (k == -1) {
            const int32_t old = atomicCAS(&t.keys[h], -1, c);
            if (old == -1) {
                const int32_t pos = atomicAdd(t.nocc, 1);
                TMem<HBM>::st(&t.occ[pos], (int32_t)h);
                return (int32_t)h;
            }
            if (old == c) return (int32_t)h;
        }
        h = (h + 1) & t.capmask;
    }
}

// One vertex, G lanes.  `lane` is the lane index inside the group, `gbase` the group's
// first lane inside the wave.  Follows louvain.cpp:2281-2466 and 2220-2279.
template <int G, bool HBM>
__device__ __forceinline__ void move_vertex(const MoveArgs& a, const Table<G, HBM>& t, int32_t v, int lane,
                                            int gbase, unsigned* inactive_local)
{
    using mask_t = typename MaskT<G>::type;
    const mask_t gmask = (G == 64) ? (mask_t)~0ull : (mask_t)((1ull << G) - 1);

    if (a.gid && !a.gact[a.gid[v]]) return;  // its graph has left the phase
    if (a.active && !a.active[v]) {
        // Frozen vertex: clusterWeight[i] = frozenClusterWeight[i] (louvain.cpp:101-104).
        if (lane == 0) {
            a.cw[v] = a.frozen[v];
            atomicAdd(inactive_local, 1u);
        }
        return;
    }
    const int64_t e0 = a.offs[v], e1 = a.offs[v + 1];
    const int64_t cc = a.curr[v];
    if (e0 == e1) {
        // Isolated vertex: target = cc, clusterWeight stays 0 (louvain.cpp:2323-2335).
        if (lane == 0) {
            a.target[v] = cc;
            a.cw[v] = 0.0;
            if (a.frozen) a.frozen[v] = 0.0;
        }
        return;
    }

    // The own community is the first map entry, counter 0.0 (louvain.cpp:2326-2327).
    if (lane == 0) {
        const int32_t s = table_slot<G, HBM>(t, (int32_t)cc);
        TMem<HBM>::std_(&t.vals[s], 0.0);
        *t.cslot = s;
    }
    wave_sync();

    double selfLoop = 0.0;
    for (int64_t base = e0; base < e1; base += G) {
        const int64_t j = base + lane;
        const bool valid = j < e1;
        int32_t slot = -1;
        double w = 0.0;
        bool self = false;
        if (valid) {
            const DEdge ed = a.edges[j];
            w = ed.weight;
            self = (ed.tail == (int64_t)v);
            slot = table_slot<G, HBM>(t, (int32_t)a.curr[ed.tail]);
        }
        t.wst[lane] = w;
        mask_t m = 0;
        if constexpr (!HBM) {
            if (valid) atomicOr(&t.masks[slot], (mask_t)1 << lane);
            wave_sync();
            if (valid) m = t.masks[slot];
            wave_sync();
        } else {
            t.sst[lane] = slot;
            wave_sync();
            if (valid)
                for (int k = 0; k < G; k++) m |= (t.sst[k] == slot) ? ((mask_t)1 << k) : (mask_t)0;
            wave_sync();
        }
        // The lowest lane holding a community folds its weights in edge order.
        if (valid && (m & (((mask_t)1 << lane) - 1)) == 0) {
            double acc = TMem<HBM>::ldd(&t.vals[slot]);
            acc = is_sent(acc) ? w : acc + w;
            mask_t rest = m & (m - 1);
            while (rest) {
                const int k = (G == 64) ? __builtin_ctzll((unsigned long long)rest) : __builtin_ctz((unsigned)rest);
                rest &= rest - 1;
                acc = acc + t.wst[k];
            }
            TMem<HBM>::std_(&t.vals[slot], acc);
            if constexpr (!HBM) t.masks[slot] = 0;
        }
        // selfLoop += weight for tail == vertex, in edge order (louvain.cpp:2440-2441).
        mask_t sm = (mask_t)((__ballot(self) >> gbase) & gmask);
        while (sm) {
            const int k = (G == 64) ? __builtin_ctzll((unsigned long long)sm) : __builtin_ctz((unsigned)sm);
            sm &= sm - 1;
            selfLoop = selfLoop + t.wst[k];
        }
        wave_sync();
    }

    // distGetMaxIndex: strict-improvement argmax over candidates != cc with the
    // smallest-id tie-break; only gains > 0 can win (louvain.cpp:2246-2273).
    const int32_t nocc = *t.nocc;
    const int32_t cs = *t.cslot;
    const double kv = a.vdeg[v];
    const double ccDeg = a.cdeg[cc];
    const int64_t ccSize = a.csize[cc];
    const double constant = a.gid ? a.constant[a.gid[v]] : *a.constant;
    const double ecc = TMem<HBM>::ldd(&t.vals[cs]);
    const double eix = ecc - selfLoop;
    const double ax = ccDeg - kv;
    double bg = 0.0;
    long long bc = LLONG_MAX;
    for (int p = lane; p < nocc; p += G) {
        const int32_t s = TMem<HBM>::ld(&t.occ[p]);
        const int32_t c = TMem<HBM>::ld(&t.keys[s]);
        if ((int64_t)c == cc) continue;
        const double ay = a.cdeg[c];
        const double eiy = TMem<HBM>::ldd(&t.vals[s]);
        const double gain = 2.0 * (eiy - eix) - 2.0 * kv * (ay - ax) * constant;
        if (gain > 0.0 && (gain > bg || (gain == bg && (long long)c < bc))) {
            bg = gain;
            bc = c;
        }
    }
    for (int o = G / 2; o > 0; o >>= 1) {
        const double og = __shfl_xor(bg, o, G);
        const long long oc = __shfl_xor(bc, o, G);
        if (og > bg || (og == bg && oc < bc)) {
            bg = og;
            bc = oc;
        }
    }
    if (lane == 0) {
        int64_t tgt = cc;
        int64_t maxSize = ccSize;
        if (bg > 0.0) {
            tgt = bc;
            maxSize = a.csize[bc];
        }
        // Singleton rule (louvain.cpp:2275-2276).
        if (maxSize == 1 && ccSize == 1 && tgt > cc) tgt = cc;
        a.target[v] = tgt;
        a.cw[v] = ecc;
        if (a.frozen) a.frozen[v] = ecc;
    }
    // Restore the touched slots for the next vertex.
    for (int p = lane; p < nocc; p += G) {
        const int32_t s = TMem<HBM>::ld(&t.occ[p]);
        TMem<HBM>::st(&t.keys[s], (int32_t)-1);
        TMem<HBM>::std_(&t.vals[s], sent_value());
    }
    wave_sync();
    if (lane == 0) *t.nocc = 0;
    wave_sync();
}

// Insert-or-find without an occupancy list: CAS first (one LDS round trip whether the key
// is new or already present).
// Edge chunks per row for the 8- and 16-lane groups (the others take two): a chunk pass
// costs a wave about the same instructions whatever G is, so narrower groups with more
// chunks serve more vertices per wave.
#ifndef VT_NCH_G8
#define VT_NCH_G8 2
#endif
#ifndef VT_NCH_G16
#define VT_NCH_G16 3
#endif
#ifndef VT_CAP_G8
#define VT_CAP_G8 32
#endif
#ifndef VT_NCH_G32
#define VT_NCH_G32 2
#endif
#ifndef VT_NCH_G64
#define VT_NCH_G64 3
#endif
constexpr int kNch[5] = {2, VT_NCH_G8, VT_NCH_G16, VT_NCH_G32, VT_NCH_G64};
// Largest row each bucket's kernel takes: NCH * G - 1 (G = 4, 8, 16, 32, 64), then the
// one-wave 2048-slot kernel.
constexpr int32_t kBucketCap[6] = {2 * 4 - 1, VT_NCH_G8 * 8 - 1, VT_NCH_G16 * 16 - 1, VT_NCH_G32 * 32 - 1,
                                   VT_NCH_G64 * 64 - 1, 1023};

#ifndef VT_SPLIT_EDGES
#define VT_SPLIT_EDGES 1
#endif
constexpr bool kSplitEdges = VT_SPLIT_EDGES;

#ifndef VT_DPP_ARGMAX
#define VT_DPP_ARGMAX 1
#endif
constexpr bool kDppArgmax = VT_DPP_ARGMAX;


// (gain, community) argmax step: take the other candidate if its gain is larger, or equal
// with a smaller id (the reference's order-independent rule, louvain.cpp:2246-2273).
__device__ __forceinline__ void argmax_take(double& bg, int32_t& bc, double og, int32_t oc)
{
    // bitwise, not short-circuit: no divergent branches
    const bool take = (og > bg) | ((og == bg) & (oc < bc));
    bg = take ? og : bg;
    bc = take ? oc : bc;
}

// One DPP step: every lane combines with the lane the control selects; lanes whose source
// lies outside the row (or the masked rows) see the identity (0.0, INT_MAX).
template <int CTRL, int ROWS>
__device__ __forceinline__ void argmax_dpp_step(double& bg, int32_t& bc)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(bg), CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(bg), CTRL, ROWS, 0xF, false);
    const int oc = __builtin_amdgcn_update_dpp(INT_MAX, bc, CTRL, ROWS, 0xF, false);
    argmax_take(bg, bc, __hiloint2double(hi, lo), oc);
}

// Argmax over the G lanes of each group, reduced into the group's highest lane with DPP
// (row_shr within 16-lane rows, then row_bcast15 / row_bcast31 across rows): no LDS traffic.
template <int G>
__device__ __forceinline__ void group_argmax_dpp(double& bg, int32_t& bc)
{
    if constexpr (G >= 2) argmax_dpp_step<0x111, 0xF>(bg, bc);   // row_shr:1
    if constexpr (G >= 4) argmax_dpp_step<0x112, 0xF>(bg, bc);   // row_shr:2
    if constexpr (G >= 8) argmax_dpp_step<0x114, 0xF>(bg, bc);   // row_shr:4
    if constexpr (G >= 16) argmax_dpp_step<0x118, 0xF>(bg, bc);  // row_shr:8
    if constexpr (G >= 32) argmax_dpp_step<0x142, 0xA>(bg, bc);  // row_bcast15 into rows 1, 3
    if constexpr (G >= 64) argmax_dpp_step<0x143, 0xC>(bg, bc);  // row_bcast31 into rows 2, 3
}

#ifndef VT_ARGMAX_MAXMIN
#define VT_ARGMAX_MAXMIN 1
#endif
constexpr bool kArgmaxMaxMin = VT_ARGMAX_MAXMIN;

// All-reduce step inside 16-lane rows: every lane reads the lane CTRL selects in its own row
// (quad_perm, half-row mirror, row mirror), so no lane sees an identity.
template <int CTRL>
__device__ __forceinline__ double dpp_row_f64(double x)
{
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(x), __double2loint(x), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(x), __double2hiint(x), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ int32_t dpp_row_i32(int32_t x)
{
    return __builtin_amdgcn_update_dpp(x, x, CTRL, 0xF, 0xF, false);
}

// The same argmax for groups of 4, 8 or 16 lanes in two all-reductions: the largest gain
// over the group, then the smallest community id among the lanes holding exactly that gain
// (the lexicographic rule argmax_take applies, done as max then min).  Every lane of the
// group ends with the result.  Lane steps: quad_perm [1,0,3,2] (0xB1), quad_perm [2,3,0,1]
// (0x4E), half-row mirror (0x141), row mirror (0x140).
template <int G>
__device__ __forceinline__ void group_argmax_maxmin(double& bg, int32_t& bc)
{
    static_assert(G == 4 || G == 8 || G == 16, "groups inside one 16-lane row");
    double m = bg;
    m = fmax(m, dpp_row_f64<0xB1>(m));
    m = fmax(m, dpp_row_f64<0x4E>(m));
    if constexpr (G >= 8) m = fmax(m, dpp_row_f64<0x141>(m));
    if constexpr (G >= 16) m = fmax(m, dpp_row_f64<0x140>(m));
    int32_t c = (bg == m) ? bc : INT_MAX;
    c = min(c, dpp_row_i32<0xB1>(c));
    c = min(c, dpp_row_i32<0x4E>(c));
    if constexpr (G >= 8) c = min(c, dpp_row_i32<0x141>(c));
    if constexpr (G >= 16) c = min(c, dpp_row_i32<0x140>(c));
    bg = m;
    bc = c;
}

// Table of one group (G lanes, one vertex at a time) in LDS.
// Lanes with nothing to store or read point at the block's trash cells instead, so LDS
// accesses need no divergent branches (the trash is written and read racily and never used).
template <int G>
struct GroupTable {
    using mask_t = typename MaskT<G>::type;
    int32_t* keys;    // CAP slots, -1 = empty
    double* vals;     // CAP, -0.0 = no value yet (the additive identity)
    mask_t* masks;    // CAP: lanes of the current chunk holding the slot's community
    uint8_t* occ;     // 2G: occupied slots in insertion order
    double* wst;      // G staged chunk weights; wst[-1] holds -0.0
    uint32_t capmask;
    int32_t* tkey;    // trash cells shared by the block
    double* tval;
    mask_t* tmask;
    uint8_t* tocc;
};

#ifndef VT_FOLD_UNROLL
#define VT_FOLD_UNROLL 2
#endif

// One chunk of at most G edges of a row already in registers: insert / find each
// community (recording new slots in occ), elect the lowest lane per community, and let it
// fold the chunk's weights for that community in edge order onto the running counter.  A
// counter created here starts as the first weight, or 0.0 + w for the vertex's own
// community, whose counter the reference seeds with 0.0 (louvain.cpp:2326-2327, 2456-2462).
// `nocc` (uniform over the group) counts the slots recorded so far.
template <int G>
__device__ __forceinline__ void fold_chunk(const GroupTable<G>& t, double* ecc_cell, int32_t c, int32_t cc, double w,
                                           bool valid, bool self, int lane, int gbase, double& selfLoop, int& nocc)
{
    using mask_t = typename MaskT<G>::type;
    const mask_t gmask = (G == 64) ? (mask_t)~0ull : (mask_t)((1ull << G) - 1);
    const mask_t below = (mask_t)(((mask_t)1 << lane) - 1);
    int32_t slot = -1;
    bool fresh = false;
    if (valid) {
        uint32_t h = (slot_hash(c) >> 7) & t.capmask;
        for (;;) {
            const int32_t old = atomicCAS(&t.keys[h], -1, c);
            if (old == -1 || old == c) {
                fresh = old == -1;
                break;
            }
            h = (h + 1) & t.capmask;
        }
        slot = (int32_t)h;
    }
    const mask_t fm = (mask_t)((__ballot(fresh) >> gbase) & gmask);
    *(fresh ? &t.occ[nocc + __popcll((unsigned long long)(fm & below))] : t.tocc) = (uint8_t)slot;
    nocc += __popcll((unsigned long long)fm);
    t.wst[lane] = w;
    mask_t* mp = valid ? &t.masks[slot] : t.tmask;
    atomicOr(mp, (mask_t)1 << lane);
    wave_sync();
    const mask_t m = *mp;
    wave_sync();
    if (valid && (m & below) == 0) {
        // An empty slot holds -0.0, and -0.0 + w == w for every w, so a new counter starts
        // as its first weight.  The vertex's own community starts from 0.0 + w: adding +0.0
        // turns an empty slot's -0.0 into +0.0 and leaves a running sum unchanged (a sum
        // that started from +0.0 is never -0.0 under round-to-nearest).
        double acc = t.vals[slot];
        if (c == cc) acc = acc + 0.0;
        acc = acc + w;
        acc = fold_rest<mask_t, VT_FOLD_UNROLL>(acc, (mask_t)(m & (m - 1)), t.wst);
        t.vals[slot] = acc;
        t.masks[slot] = 0;
        if (c == cc) *ecc_cell = acc;
    }
    const unsigned long long sb = __ballot(self);
    if (sb) {  // wave-uniform: most chunks hold no self-loop
        mask_t sm = (mask_t)((sb >> gbase) & gmask);
        while (sm) {
            const int k = (G == 64) ? __builtin_ctzll((unsigned long long)sm) : __builtin_ctz((unsigned)sm);
            sm &= sm - 1;
            selfLoop = selfLoop + t.wst[k];
        }
    }
    wave_sync();
}

// Per-vertex header, prefetched one vertex ahead.
// A vertex as the fused buckets see it: row start, degree (< 0 for a vertex early
// termination froze), own community and weighted degree.  Rows here have degree <= 127 and
// community ids fit 32 bits (vt_graph_upload rejects nv >= 2^31).
// In a batch, degree -2 marks a vertex whose graph has left the phase (nothing is written)
// and `cst` carries its graph's 1/2m.
struct VHeaderBase {
    int32_t v;
    int32_t deg;
    int64_t e0;
    int32_t cc;
    double kv;
};
template <bool BATCH>
struct VHeader : VHeaderBase {};
template <>
struct VHeader<true> : VHeaderBase {
    double cst;
};

template <bool BATCH, bool REC>
__device__ __forceinline__ VHeader<BATCH> load_header(const MoveArgs& a, int32_t v, bool have)
{
    VHeader<BATCH> h;
    h.v = v;
    h.deg = -1;
    h.e0 = 0;
    h.cc = 0;
    h.kv = 0.0;
    if constexpr (BATCH) h.cst = 0.0;
    if (have) {
        const bool live = !(a.active && !a.active[v]);
        h.e0 = a.offs[v];
        h.deg = live ? (int32_t)(a.offs[v + 1] - h.e0) : -1;
        if constexpr (REC) h.cc = a.rec[v].comm;
        else if constexpr (kSplitEdges) h.cc = a.curr32[v];
        else h.cc = (int32_t)a.curr[v];
        h.kv = a.vdeg[v];
        if constexpr (BATCH) {
            const int32_t gg = a.gid[v];
            h.cst = a.constant[gg];
            if (!a.gact[gg]) h.deg = -2;
        }
    }
    return h;
}

#ifndef VT_PIPE_ROWS
#define VT_PIPE_ROWS 0
#endif
constexpr bool kPipeRows = VT_PIPE_ROWS;

// Rows of degree <= NCH*G - 1 read in NCH chunks of G edges, table CAP = 4G slots,
// software-pipelined across the vertices a group processes: while vertex i is folded in
// LDS and its candidates' community degrees are gathered, vertex i+1's edge chunks and
// neighbour-community gathers are already in flight.  Same results as move_vertex.

template <int NCH>
struct RowRegs {
    int32_t t[NCH];   // neighbour ids (low words; ids < 2^31)
    double w[NCH];
    int32_t c[NCH];   // their communities
    bool ok[NCH], self[NCH];
};

// Only the tail (32 bits) and the weight of an edge are read, which keeps a chunk in
// flight at three registers a lane.
template <int G, int NCH, bool BATCH>
__device__ __forceinline__ void row_issue_edges(const MoveArgs& a, const VHeader<BATCH>& h, int lane,
                                                RowRegs<NCH>& r)
{
#pragma unroll
    for (int u = 0; u < NCH; u++) {
        const int64_t j = h.e0 + u * G + lane;
        r.ok[u] = u * G + lane < h.deg;
        r.t[u] = -1;
        r.w[u] = 0.0;
        if (r.ok[u]) {
            if constexpr (kSplitEdges) {
                r.t[u] = a.tail32[j];
                r.w[u] = a.wgt[j];
            } else {
                r.t[u] = reinterpret_cast<const int32_t*>(a.edges)[4 * j];
                r.w[u] = reinterpret_cast<const double*>(a.edges)[2 * j + 1];
            }
        }
    }
}

template <bool REC, int NCH>
__device__ __forceinline__ void row_issue_gathers(const MoveArgs& a, int32_t v, RowRegs<NCH>& r)
{
#pragma unroll
    for (int u = 0; u < NCH; u++) {
        r.c[u] = 0;
        if (r.ok[u]) {
            if constexpr (REC) r.c[u] = a.rec[r.t[u]].comm;
            else if constexpr (kSplitEdges) r.c[u] = a.curr32[r.t[u]];
            else r.c[u] = reinterpret_cast<const int32_t*>(a.curr)[2 * (int64_t)r.t[u]];
        }
        r.self[u] = r.ok[u] && r.t[u] == v;
    }
}

template <int G, int CAP, int BLOCK, bool PIPE, bool BATCH, int NCH, bool REC>
__device__ __forceinline__ void run_two_pipelined(const MoveArgs& a, const GroupTable<G>& t, double* ecc_cell,
                                                  const int32_t* __restrict__ list, int64_t n, int64_t idx,
                                                  int64_t stride, int lane, int gbase, unsigned* inactive_local)
{
    static_assert(CAP >= 4 * G && CAP <= 256, "occ holds 8-bit slot indexes");
    static_assert(NCH * G <= 3 * CAP / 4, "at most 3/4 of the table filled by distinct communities");
    if (idx >= n) return;
    VHeader<BATCH> h = load_header<BATCH, REC>(a, list[idx], true);
    int32_t vn = (idx + stride < n) ? list[idx + stride] : 0;
    RowRegs<NCH> r;
    if constexpr (PIPE) row_issue_edges<G>(a, h, lane, r);
    VHeader<BATCH> hn = load_header<BATCH, REC>(a, vn, idx + stride < n);
    if constexpr (PIPE) row_issue_gathers<REC>(a, h.v, r);
    const double constant0 = BATCH ? 0.0 : *a.constant;
    while (idx < n) {
        const bool has_next = PIPE && idx + stride < n;
        // list entry two ahead, header one ahead (and with PIPE the next vertex's edges)
        const int64_t i2 = idx + 2 * stride;
        const int32_t vnn = (i2 < n) ? list[i2] : 0;
        RowRegs<NCH> rn;
        if constexpr (PIPE) {
            row_issue_edges<G>(a, hn, lane, rn);
        } else {
            row_issue_edges<G>(a, h, lane, r);
            row_issue_gathers<REC>(a, h.v, r);
        }
        const VHeader<BATCH> hnn = load_header<BATCH, REC>(a, vnn, i2 < n);
        double constant = constant0;
        if constexpr (BATCH) constant = h.cst;

        const int32_t v = h.v;
        const int64_t cc = h.cc;
        if (BATCH && h.deg < -1) {
            // a batch graph that has left the phase: nothing to do
        } else if (h.deg < 0) {
            if (lane == 0) {
                a.cw[v] = a.frozen[v];
                atomicAdd(inactive_local, 1u);
            }
        } else if (h.deg == 0) {
            if (lane == 0) {
                a.target[v] = cc;
                a.cw[v] = 0.0;
                if (a.frozen) a.frozen[v] = 0.0;
            }
        } else {
            // the loads of this vertex have been issued: its LDS work and argmax take the
            // SIMD ahead of waves still waiting on memory, so it reaches its next loads sooner
            __builtin_amdgcn_s_setprio(1);
            double ccDeg;
            int64_t ccSize;
            if constexpr (REC) {
                ccDeg = a.rec[cc].deg;
                ccSize = a.rec[cc].size;
            } else {
                ccDeg = a.cdeg[cc];
                ccSize = a.csize[cc];
            }
            if (lane == 0) *ecc_cell = 0.0;
            double selfLoop = 0.0;
            int nocc = 0;
            fold_chunk<G>(t, ecc_cell, r.c[0], (int32_t)cc, r.w[0], r.ok[0], r.self[0], lane, gbase, selfLoop, nocc);
#pragma unroll
            for (int u = 1; u < NCH; u++)
                if (h.deg > u * G)
                    fold_chunk<G>(t, ecc_cell, r.c[u], (int32_t)cc, r.w[u], r.ok[u], r.self[u], lane, gbase, selfLoop,
                                  nocc);
            const double ecc = *ecc_cell;
            // the next vertex's neighbour communities, before waiting on this one's gathers
            if (has_next) row_issue_gathers<REC>(a, hn.v, rn);
            // distGetMaxIndex (louvain.cpp:2246-2276) over the occupied slots, clearing them
            const double kv = h.kv;
            const double eix = ecc - selfLoop;
            const double ax = ccDeg - kv;
            double bg = 0.0;
            int32_t bc = INT_MAX;
            for (int base = 0; base < nocc; base += G) {
                const bool ok = base + lane < nocc;
                const int sl = t.occ[ok ? base + lane : 0];
                int32_t* kp = ok ? &t.keys[sl] : t.tkey;
                double* vp = ok ? &t.vals[sl] : t.tval;
                const int32_t ck = *kp;
                const double ey = *vp;
                *kp = -1;
                *vp = -0.0;
                const bool cand = ok & ((int64_t)ck != cc);
                const int32_t ci = cand ? ck : (int32_t)cc;
                double ay;
                if constexpr (REC) ay = a.rec[ci].deg;
                else ay = a.cdeg[ci];
                const double gain = 2.0 * (ey - eix) - 2.0 * kv * (ay - ax) * constant;
                const bool take = cand & (gain > 0.0) & ((gain > bg) | ((gain == bg) & (ck < bc)));
                bg = take ? gain : bg;
                bc = take ? ck : bc;
            }
            constexpr int kWriter = kDppArgmax ? G - 1 : 0;
            if constexpr (kArgmaxMaxMin && G <= 16) {
                group_argmax_maxmin<G>(bg, bc);
            } else if constexpr (kDppArgmax) {
                group_argmax_dpp<G>(bg, bc);
            } else {
                for (int o = G / 2; o > 0; o >>= 1) {
                    const double og = __shfl_xor(bg, o, G);
                    const int32_t oc = __shfl_xor(bc, o, G);
                    if (og > bg || (og == bg && oc < bc)) {
                        bg = og;
                        bc = oc;
                    }
                }
            }
            if (lane == kWriter) {
                int64_t tgt = cc;
                if (bg > 0.0) {
                    tgt = bc;
                    if constexpr (REC) {
                        if (ccSize == 1 && tgt > cc && a.rec[tgt].size == 1) tgt = cc;  // louvain.cpp:2275-2276
                    } else {
                        if (ccSize == 1 && tgt > cc && a.csize[tgt] == 1) tgt = cc;
                    }
                }
                a.target[v] = tgt;
                a.cw[v] = ecc;
                if (a.frozen) a.frozen[v] = ecc;
            }
            wave_sync();
            __builtin_amdgcn_s_setprio(0);
        }
        if (has_next && h.deg <= 0) row_issue_gathers<REC>(a, hn.v, rn);
        h = hn;
        hn = hnn;
        if constexpr (PIPE) r = rn;
        idx += stride;
    }
}

// LDS carve-out for BLOCK/G groups of CAP-slot tables.
constexpr size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

template <int G, int CAP, int BLOCK>
struct LdsLayout {
    using mask_t = typename MaskT<G>::type;
    static constexpr int NG = BLOCK / G;
    static constexpr size_t keys = 0;
    static constexpr size_t vals = align16(keys + sizeof(int32_t) * NG * CAP);
    static constexpr size_t masks = align16(vals + sizeof(double) * NG * CAP);
    static constexpr size_t occ = align16(masks + sizeof(mask_t) * NG * CAP);
    static constexpr size_t wst = align16(occ + sizeof(int32_t) * NG * CAP);
    static constexpr size_t nocc = align16(wst + sizeof(double) * NG * G);
    static constexpr size_t cslot = align16(nocc + sizeof(int32_t) * NG);
    static constexpr size_t ecc = align16(cslot + sizeof(int32_t) * NG);
    static constexpr size_t bytes = align16(ecc + sizeof(double) * NG);
};

#ifndef VT_XCD_RANGES
#define VT_XCD_RANGES 0
#endif
constexpr bool kXcdRanges = VT_XCD_RANGES;
constexpr int kXcds = 8;

// LDS carve-out of the two-chunk group buckets.
template <int G, int CAP, int BLOCK, int NCH = 2>
struct GroupLds {
    using mask_t = typename MaskT<G>::type;
    static constexpr int NG = BLOCK / G;
    static constexpr size_t keys = 0;
    static constexpr size_t vals = align16(keys + sizeof(int32_t) * NG * CAP);
    static constexpr size_t masks = align16(vals + sizeof(double) * NG * CAP);
    static constexpr size_t wst = align16(masks + sizeof(mask_t) * NG * CAP);
    static constexpr size_t ecc = align16(wst + sizeof(double) * NG * (G + 1));
    static constexpr size_t trash = align16(ecc + sizeof(double) * NG);  // val, mask, key, occ
    static constexpr size_t occ = trash + 32;
    static constexpr size_t bytes = align16(occ + (size_t)NG * NCH * G);
};

// Runs one degree bucket inside a block: block `blk` of `nblk` owns list entries
// blk*NG + g, stepping by nblk*NG.  Every thread of the block must call it.
template <int G, int CAP, int BLOCK, bool TWO, bool BATCH = false, int NCH = 2, bool REC = false>
__device__ __forceinline__ void run_bucket(const MoveArgs& a, char* lds, const int32_t* __restrict__ list, int64_t n,
                                           int64_t blk, int64_t nblk, unsigned* inactive)
{
    const int g = threadIdx.x / G;
    const int lane = threadIdx.x % G;
    const int gbase = (threadIdx.x % 64) / G * G;
    const int64_t stride = nblk * (BLOCK / G);
    int64_t idx = blk * (BLOCK / G) + g;
    if constexpr (TWO) {
        using L = GroupLds<G, CAP, BLOCK, NCH>;
        using mask_t = typename L::mask_t;
        constexpr int NG = L::NG;
        int32_t* keys = (int32_t*)(lds + L::keys);
        double* vals = (double*)(lds + L::vals);
        mask_t* masks = (mask_t*)(lds + L::masks);
        double* wst = (double*)(lds + L::wst);
        for (int i = threadIdx.x; i < NG * CAP; i += BLOCK) {
            keys[i] = -1;
            vals[i] = -0.0;
            masks[i] = 0;
        }
        if (threadIdx.x < NG) wst[threadIdx.x * (G + 1)] = -0.0;
        __syncthreads();
        char* tr = lds + L::trash;
        GroupTable<G> t{keys + g * CAP, vals + g * CAP, masks + g * CAP, (uint8_t*)(lds + L::occ) + g * NCH * G,
                        wst + g * (G + 1) + 1, (uint32_t)(CAP - 1), (int32_t*)(tr + 16), (double*)tr,
                        (mask_t*)(tr + 8), (uint8_t*)(tr + 20)};
        double* ecc_cell = (double*)(lds + L::ecc) + g;
        if (kXcdRanges && nblk % kXcds == 0) {
            // Workgroups are dealt to the XCDs round-robin, so block blk runs on XCD
            // blk % 8 (block counts and bucket starts are multiples of 8).  Give each XCD a
            // contiguous eighth of the list: neighbouring vertices then share one L2.
            const int64_t x = blk % kXcds, per = (n + kXcds - 1) / kXcds;
            const int64_t lo = x * per < n ? x * per : n, hi = lo + per < n ? lo + per : n;
            run_two_pipelined<G, CAP, BLOCK, kPipeRows, BATCH, NCH, REC>(a, t, ecc_cell, list + lo, hi - lo,
                                                               (blk / kXcds) * (BLOCK / G) + g,
                                                               (nblk / kXcds) * (BLOCK / G), lane, gbase, inactive);
        } else {
            run_two_pipelined<G, CAP, BLOCK, kPipeRows, BATCH, NCH, REC>(a, t, ecc_cell, list, n, idx, stride, lane, gbase,
                                                                    inactive);
        }
    } else {
        using L = LdsLayout<G, CAP, BLOCK>;
        using mask_t = typename L::mask_t;
        constexpr int NG = L::NG;
        int32_t* keys = (int32_t*)(lds + L::keys);
        double* vals = (double*)(lds + L::vals);
        mask_t* masks = (mask_t*)(lds + L::masks);
        int32_t* occ = (int32_t*)(lds + L::occ);
        double* wst = (double*)(lds + L::wst);
        int32_t* nocc = (int32_t*)(lds + L::nocc);
        int32_t* cslot = (int32_t*)(lds + L::cslot);
        for (int i = threadIdx.x; i < NG * CAP; i += BLOCK) {
            keys[i] = -1;
            vals[i] = sent_value();
            masks[i] = 0;
        }
        if (threadIdx.x < NG) nocc[threadIdx.x] = 0;
        __syncthreads();
        Table<G, false> t{keys + g * CAP, vals + g * CAP, masks + g * CAP, occ + g * CAP, wst + g * G, nullptr,
                          nocc + g, cslot + g, (uint32_t)(CAP - 1)};
        for (; idx < n; idx += stride) move_vertex<G, false>(a, t, list[idx], lane, gbase, inactive);
    }
}

// Buckets 0-3 (degree <= 63) of one sweep in two launches over one block plan: blocks
// [bbeg[b], bbeg[b+1]) run bucket b over its list range.  Bucket 2 (16-lane groups) runs
// alone and buckets 0, 1 and 3 share the other launch: each kernel then fits 64 VGPRs,
// so a SIMD holds eight waves (one kernel for all four needs 70 and holds seven).
// Bucket 4 (64 <= degree <= 127) has its own launch so that its larger tables do not set
// the LDS footprint, and so the occupancy, of these.
constexpr int kFusedBuckets = 4;
constexpr int kFusedBlock = 256;
constexpr size_t kFusedLds = std::max({GroupLds<4, 16, 256, kNch[0]>::bytes, GroupLds<8, VT_CAP_G8, 256, kNch[1]>::bytes,
                                       GroupLds<16, 64, 256, kNch[2]>::bytes, GroupLds<32, 128, 256, kNch[3]>::bytes});
constexpr size_t kMidLds = GroupLds<64, 256, 256, kNch[4]>::bytes;
static_assert(kFusedLds <= 20480 || VT_CAP_G8 > 32, "eight 256-lane blocks per CU");

struct SweepPlan {
    int64_t bbeg[kFusedBuckets + 1];  // block ranges
    int64_t off[kFusedBuckets];       // list offsets
    int64_t n[kFusedBuckets];         // list lengths
};

enum FusedPart { kPartRest = 1, kPartG16 = 2 };

template <bool BATCH, int PART, bool REC>
__global__ void __launch_bounds__(kFusedBlock) __attribute__((amdgpu_waves_per_eu(8)))
k_local_move(MoveArgs a, const int32_t* __restrict__ lists, SweepPlan plan)
{
    __shared__ __align__(16) char lds[kFusedLds];
    __shared__ unsigned s_inactive;
    if (threadIdx.x == 0) s_inactive = 0;
    // kPartRest skips bucket 2's block range; kPartG16 covers only that range
    int64_t bid = blockIdx.x;
    if (PART == kPartRest && bid >= plan.bbeg[2]) bid += plan.bbeg[3] - plan.bbeg[2];
    if (PART == kPartG16) bid += plan.bbeg[2];
    int b = 0;
    while (b < kFusedBuckets - 1 && bid >= plan.bbeg[b + 1]) b++;
    const int64_t blk = bid - plan.bbeg[b], nblk = plan.bbeg[b + 1] - plan.bbeg[b];
    const int32_t* list = lists + plan.off[b];
    if (PART == kPartG16) {
        run_bucket<16, 64, 256, true, BATCH, kNch[2], REC>(a, lds, list, plan.n[2], blk, nblk, &s_inactive);
    } else {
        switch (b) {
            case 0: run_bucket<4, 16, 256, true, BATCH, kNch[0], REC>(a, lds, list, plan.n[0], blk, nblk, &s_inactive); break;
            case 1: run_bucket<8, VT_CAP_G8, 256, true, BATCH, kNch[1], REC>(a, lds, list, plan.n[1], blk, nblk, &s_inactive); break;
            default: run_bucket<32, 128, 256, true, BATCH, kNch[3], REC>(a, lds, list, plan.n[3], blk, nblk, &s_inactive); break;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_inactive && a.inactive_count) atomicAdd(a.inactive_count, (unsigned long long)s_inactive);
}

// Bucket 4 (64 <= degree <= 127): four 64-lane groups per block, 256-slot tables.
template <bool BATCH, bool REC>
__global__ void __launch_bounds__(256) k_local_move_mid(MoveArgs a, const int32_t* __restrict__ list, int64_t n)
{
    __shared__ __align__(16) char lds[kMidLds];
    __shared__ unsigned s_inactive;
    if (threadIdx.x == 0) s_inactive = 0;
    run_bucket<64, 256, 256, true, BATCH, kNch[4], REC>(a, lds, list, n, blockIdx.x, gridDim.x, &s_inactive);
    __syncthreads();
    if (threadIdx.x == 0 && s_inactive && a.inactive_count) atomicAdd(a.inactive_count, (unsigned long long)s_inactive);
}

// Bucket 5 (128 <= degree <= 1023): one wave per vertex, 2048-slot LDS table.
template <int CAP>
__global__ void __launch_bounds__(64) k_local_move_wide(MoveArgs a, const int32_t* __restrict__ list, int64_t n)
{
    __shared__ __align__(16) char lds[LdsLayout<64, CAP, 64>::bytes];
    __shared__ unsigned s_inactive;
    if (threadIdx.x == 0) s_inactive = 0;
    run_bucket<64, CAP, 64, false>(a, lds, list, n, blockIdx.x, gridDim.x, &s_inactive);
    __syncthreads();
    if (threadIdx.x == 0 && s_inactive && a.inactive_count) atomicAdd(a.inactive_count, (unsigned long long)s_inactive);
}

// ---- lane per vertex ---------------------------------------------------------------------
//
// For rows of degree <= MAXD one lane owns one vertex and walks its row in CSR order, so the
// per-community counters are the reference's left folds with no cross-lane traffic.  The
// lane's map (louvain.cpp:2297-2298) is a compact key/value array in LDS holding at most
// deg+1 communities, searched linearly; the wave carves lane regions out of an ARENA-entry
// LDS pool by a prefix sum of deg+1 and runs extra rounds for lanes that did not fit.  Row
// loads run three stages deep: edges two batches ahead, the community gather one batch
// ahead of the map updates.  Blocks map to list ranges so that each XCD (blocks are dealt
// to the eight XCDs round robin) sweeps one contiguous range and keeps its gathers in its L2.

__device__ __forceinline__ int64_t xcd_logical_block(int64_t b, int64_t nb)
{
    const int64_t x = b % 8, i = b / 8, per = nb / 8, rem = nb % 8;
    return x * per + (x < rem ? x : rem) + i;
}

template <int MAXD, int ARENA, int U>
__global__ void __launch_bounds__(64) k_local_move_lpv(MoveArgs a, const int32_t* __restrict__ list, int64_t n,
                                                       int64_t chunk)
{
    __shared__ int32_t s_keys[ARENA];
    __shared__ double s_vals[ARENA];
    const int lane = threadIdx.x;
    const int64_t lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int64_t lo = lb * chunk;
    const int64_t hi = lo + chunk < n ? lo + chunk : n;
    const double constant = *a.constant;
    unsigned inactive = 0;

    for (int64_t base = lo; base < hi; base += 64) {
        const int64_t idx = base + lane;
        bool pending = false;
        int32_t v = 0;
        int64_t e0 = 0, e1 = 0, cc = 0;
        if (idx < hi) {
            v = list[idx];
            if (a.active && !a.active[v]) {
                a.cw[v] = a.frozen[v];  // louvain.cpp:101-104
                inactive++;
            } else {
                e0 = a.offs[v];
                e1 = a.offs[v + 1];
                cc = a.curr[v];
                if (e0 == e1) {  // isolated: stays, clusterWeight 0 (louvain.cpp:2323-2335)
                    a.target[v] = cc;
                    a.cw[v] = 0.0;
                    if (a.frozen) a.frozen[v] = 0.0;
                } else {
                    pending = true;
                }
            }
        }
        while (__ballot(pending)) {
            const int need = pending ? (int)(e1 - e0 + 1) : 0;
            int incl = need;
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(incl, o);
                if (lane >= o) incl += t;
            }
            if (pending && incl <= ARENA) {
                pending = false;
                int32_t* keys = s_keys + (incl - need);
                double* vals = s_vals + (incl - need);
                const double kv = a.vdeg[v];
                const double ccDeg = a.cdeg[cc];
                const int64_t ccSize = a.csize[cc];
                keys[0] = (int32_t)cc;  // own community first, counter 0.0 (louvain.cpp:2326-2327)
                vals[0] = 0.0;
                int nk = 1;
                double selfLoop = 0.0;

                DEdge eA[U], eB[U], eC[U];
                int32_t cA[U], cB[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    if (e0 + u < e1) eA[u] = a.edges[e0 + u];
                    if (e0 + U + u < e1) eB[u] = a.edges[e0 + U + u];
                }
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (e0 + u < e1) cA[u] = (int32_t)a.curr[eA[u].tail];
                for (int64_t j = e0; j < e1; j += U) {
#pragma unroll
                    for (int u = 0; u < U; u++)
                        if (j + 2 * U + u < e1) eC[u] = a.edges[j + 2 * U + u];
#pragma unroll
                    for (int u = 0; u < U; u++)
                        if (j + U + u < e1) cB[u] = (int32_t)a.curr[eB[u].tail];
                    // distBuildLocalMapCounter (louvain.cpp:2431-2463) for this batch, in order
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        if (j + u < e1) {
                            const double w = eA[u].weight;
                            if (eA[u].tail == (int64_t)v) selfLoop = selfLoop + w;
                            const int32_t c = cA[u];
                            int k = 0;
                            while (k < nk && keys[k] != c) k++;
                            if (k < nk) {
                                vals[k] = vals[k] + w;
                            } else {
                                keys[nk] = c;
                                vals[nk] = w;
                                nk++;
                            }
                        }
                    }
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        eA[u] = eB[u];
                        cA[u] = cB[u];
                        eB[u] = eC[u];
                    }
                }

                // distGetMaxIndex (louvain.cpp:2246-2276): gathers in batches of 2U.
                const double ecc = vals[0];
                const double eix = ecc - selfLoop;
                const double ax = ccDeg - kv;
                double bg = 0.0;
                int32_t bc = -1;
                for (int k0 = 1; k0 < nk; k0 += 2 * U) {
                    int32_t ck[2 * U];
                    double ay[2 * U];
#pragma unroll
                    for (int u = 0; u < 2 * U; u++)
                        if (k0 + u < nk) ck[u] = keys[k0 + u];
#pragma unroll
                    for (int u = 0; u < 2 * U; u++)
                        if (k0 + u < nk) ay[u] = a.cdeg[ck[u]];
#pragma unroll
                    for (int u = 0; u < 2 * U; u++) {
                        if (k0 + u < nk) {
                            const double gain = 2.0 * (vals[k0 + u] - eix) - 2.0 * kv * (ay[u] - ax) * constant;
                            if (gain > 0.0 && (gain > bg || (gain == bg && ck[u] < bc))) {
                                bg = gain;
                                bc = ck[u];
                            }
                        }
                    }
                }
                int64_t tgt = cc;
                if (bg > 0.0) {
                    tgt = bc;
                    // singleton rule (louvain.cpp:2275-2276)
                    if (ccSize == 1 && tgt > cc && a.csize[tgt] == 1) tgt = cc;
                }
                a.target[v] = tgt;
                a.cw[v] = ecc;
                if (a.frozen) a.frozen[v] = ecc;
            }
            wave_sync();  // arena reuse in the next round
        }
    }
    for (int o = 32; o > 0; o >>= 1) inactive += __shfl_xor(inactive, o);
    if (lane == 0 && inactive && a.inactive_count) atomicAdd(a.inactive_count, (unsigned long long)inactive);
}

// Hub kernel: one wave per vertex, table in HBM (capacity `cap` per wave slot).
__global__ void __launch_bounds__(64) k_move_hbm(MoveArgs a, const int32_t* __restrict__ list, int64_t n,
                                                 int32_t* __restrict__ hkeys, double* __restrict__ hvals,
                                                 int32_t* __restrict__ hocc, uint32_t cap)
{
    __shared__ double s_wst[64];
    __shared__ int32_t s_sst[64];
    __shared__ int32_t s_nocc, s_cslot;
    __shared__ unsigned s_inactive;
    const int lane = threadIdx.x;
    int32_t* keys = hkeys + (size_t)blockIdx.x * cap;
    double* vals = hvals + (size_t)blockIdx.x * cap;
    int32_t* occ = hocc + (size_t)blockIdx.x * cap;
    if (lane == 0) { s_nocc = 0; s_inactive = 0; }
    __syncthreads();
    Table<64, true> t{keys, vals, nullptr, occ, s_wst, s_sst, &s_nocc, &s_cslot, cap - 1};
    for (int64_t idx = blockIdx.x; idx < n; idx += gridDim.x) move_vertex<64, true>(a, t, list[idx], lane, 0, &s_inactive);
    __syncthreads();
    if (lane == 0 && s_inactive && a.inactive_count) atomicAdd(a.inactive_count, (unsigned long long)s_inactive);
}

// ---- hub rows by sorting ----------------------------------------------------------------
// Rows above the wide bound: every (hub, neighbour community) pair becomes a key, one
// stable radix sort groups each hub's edges by community with edge order kept inside a
// group, and a workgroup per hub folds every group (a run) left to right, scores it and
// reduces the argmax.  The same left folds in CSR order as distBuildLocalMapCounter
// (louvain.cpp:2419-2466), so the same gains as the table kernels.

// Keys (hub j << cbits | community of the tail; communities < 2^cbits) and values (edge
// index) for hub j's row.
// Hub keys are (hub index << cbits | neighbour community): 32-bit when they fit, else 64-bit.
// One thread per hub entry q (hub entries are the hubs' rows back to back, hoffs their
// offsets), the value the entry's CSR index.  Each wave finds the hub of its first entry (the
// last j with hoffs[j] <= q) by a wave-uniform search; its lanes step from there.
template <class K>
__global__ void k_hub_keys(const int32_t* __restrict__ hubs, int32_t nh, const int64_t* __restrict__ hoffs,
                           const int64_t* __restrict__ offs, const int32_t* __restrict__ tail32,
                           const int32_t* __restrict__ curr32, int cbits, K* __restrict__ keys,
                           uint32_t* __restrict__ vals)
{
    const int64_t total = hoffs[nh];
    const int lane = threadIdx.x & 63;
    const int64_t wave0 = blockIdx.x * (int64_t)blockDim.x + 64 * __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int64_t base = wave0; base < total; base += (int64_t)gridDim.x * blockDim.x) {
        int32_t lo = 0, hi = nh - 1;
        while (lo < hi) {
            const int32_t mid = (lo + hi + 1) >> 1;
            if (hoffs[mid] <= base) lo = mid;
            else hi = mid - 1;
        }
        const int64_t q = base + lane;
        if (q < total) {
            int32_t j = lo;
            while (hoffs[j + 1] <= q) j++;  // hoffs[nh] = total > q
            const int64_t k = offs[hubs[j]] + (q - hoffs[j]);
            keys[q] = ((K)j << cbits) | (K)(uint32_t)curr32[tail32[k]];
            vals[q] = (uint32_t)k;
        }
    }
}

// The same keys with a workgroup per hub row (sweeps with many hubs: no search, and each
// row's entries are read in order).
template <class K>
__global__ void k_hub_keys_rows(const int32_t* __restrict__ hubs, int32_t nh, const int64_t* __restrict__ hoffs,
                                const int64_t* __restrict__ offs, const int32_t* __restrict__ tail32,
                                const int32_t* __restrict__ curr32, int cbits, K* __restrict__ keys,
                                uint32_t* __restrict__ vals)
{
    for (int32_t j = blockIdx.x; j < nh; j += gridDim.x) {
        const int64_t q0 = hoffs[j], n = hoffs[j + 1] - q0, e0 = offs[hubs[j]];
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
            const int64_t k = e0 + i;
            keys[q0 + i] = ((K)j << cbits) | (K)(uint32_t)curr32[tail32[k]];
            vals[q0 + i] = (uint32_t)k;
        }
    }
}

template <class K>
__global__ void k_hub_heads(const K* __restrict__ keys, int64_t n, int32_t* __restrict__ head)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        head[q] = (q == 0 || keys[q] != keys[q - 1]) ? 1 : 0;
}

// starts[r] = first position of run r, and starts[nruns] = n.
__global__ void k_hub_starts(const int32_t* __restrict__ head, const int32_t* __restrict__ pos, int64_t n,
                             int32_t* __restrict__ starts)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        if (head[q]) starts[pos[q]] = (int32_t)q;
        if (q == n - 1) starts[pos[q] + head[q]] = (int32_t)n;
    }
}

__global__ void k_hub_degrees(const int32_t* __restrict__ hubs, int32_t nh, const int64_t* __restrict__ offs,
                              int64_t* __restrict__ deg)
{
    for (int32_t j = blockIdx.x * blockDim.x + threadIdx.x; j <= nh; j += gridDim.x * blockDim.x)
        deg[j] = j < nh ? offs[hubs[j] + 1] - offs[hubs[j]] : 0;
}

// Left fold of a run's weights: the first as is, or from 0.0 for the vertex's own community
// (whose counter the reference seeds with 0.0); the self-loop weights folded from 0.0 too.
__device__ __forceinline__ double hub_run_fold(const DEdge* __restrict__ edges, const uint32_t* __restrict__ vals,
                                               int64_t b, int64_t e, bool own, int32_t v, double& self)
{
    double acc = own ? 0.0 : edges[vals[b]].weight;
    int64_t k = own ? b : b + 1;
    for (; k + 4 <= e; k += 4) {
        const DEdge x0 = edges[vals[k]], x1 = edges[vals[k + 1]], x2 = edges[vals[k + 2]], x3 = edges[vals[k + 3]];
        acc = acc + x0.weight;
        if (own && x0.tail == v) self = self + x0.weight;
        acc = acc + x1.weight;
        if (own && x1.tail == v) self = self + x1.weight;
        acc = acc + x2.weight;
        if (own && x2.tail == v) self = self + x2.weight;
        acc = acc + x3.weight;
        if (own && x3.tail == v) self = self + x3.weight;
    }
    for (; k < e; k++) {
        const DEdge x = edges[vals[k]];
        acc = acc + x.weight;
        if (own && x.tail == v) self = self + x.weight;
    }
    return acc;
}

// The hub local move as four flat launches, for sweeps with few hubs or one hub holding much
// of the work (a workgroup per hub would leave most of the chip idle: a few hundred hubs, the
// largest with tens of thousands of runs).  Runs are the sorted hub entries' runs of equal
// key, starts[] their bounds, pos[] the run index of each entry; own[2j], own[2j+1] = hub j's
// own-community counter and self-loop weight.
__device__ __forceinline__ int32_t hub_run_count(const int64_t* __restrict__ hoffs, int32_t nh,
                                                 const int32_t* __restrict__ pos, const int32_t* __restrict__ head)
{
    const int64_t total = hoffs[nh];
    return total ? pos[total - 1] + head[total - 1] : 0;
}

// Per hub: its own community's run folded from 0.0 with the self-loops (louvain.cpp:2326-2327,
// 2440-2441), and the argmax cells cleared.
template <class K>
__global__ void k_hub_own(MoveArgs a, const int32_t* __restrict__ hubs, int32_t nh, const int64_t* __restrict__ hoffs,
                          const K* __restrict__ keys, const uint32_t* __restrict__ vals, const int32_t* __restrict__ pos,
                          const int32_t* __restrict__ head, const int32_t* __restrict__ starts, int cbits,
                          double* __restrict__ own, unsigned long long* __restrict__ best, int32_t* __restrict__ bestc)
{
    const int64_t total = hoffs[nh];
    const int32_t nruns = hub_run_count(hoffs, nh, pos, head);
    for (int32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nh; j += gridDim.x * blockDim.x) {
        best[j] = 0ull;  // the bits of +0.0: only gains > 0 can win
        bestc[j] = INT_MAX;
        const int32_t v = hubs[j];
        if (a.active && !a.active[v]) continue;
        const int32_t r0 = hoffs[j] < total ? pos[hoffs[j]] : nruns;
        const int32_t r1 = hoffs[j + 1] < total ? pos[hoffs[j + 1]] : nruns;
        const K ckey = ((K)j << cbits) | (K)(uint32_t)a.curr[v];
        int32_t lo = r0, hi = r1;
        while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (keys[starts[mid]] < ckey) lo = mid + 1;
            else hi = mid;
        }
        double self = 0.0, ecc = 0.0;
        if (lo < r1 && keys[starts[lo]] == ckey) ecc = hub_run_fold(a.edges, vals, starts[lo], starts[lo + 1], true, v, self);
        own[2 * j] = ecc;
        own[2 * j + 1] = self;
    }
}

// Per run of another community: its counter (the run's weights folded left to right, the
// first as is) and its gain (louvain.cpp:2246-2273); best[j] = the largest positive gain's
// bits (non-negative doubles order as their bit patterns).
template <class K>
__global__ void k_hub_runs(MoveArgs a, const int32_t* __restrict__ hubs, int32_t nh, const int64_t* __restrict__ hoffs,
                           const K* __restrict__ keys, const uint32_t* __restrict__ vals, const int32_t* __restrict__ pos,
                           const int32_t* __restrict__ head, const int32_t* __restrict__ starts, int cbits,
                           const double* __restrict__ own, double* __restrict__ gains,
                           unsigned long long* __restrict__ best)
{
    const int32_t nruns = hub_run_count(hoffs, nh, pos, head);
    const double constant = *a.constant;
    const K cmask = (K)(((unsigned long long)1 << cbits) - 1);
    const int lane = threadIdx.x & 63;
    // whole waves step together (the bound is wave-uniform), so the shuffles see every lane
    for (int64_t base = blockIdx.x * (int64_t)blockDim.x + (threadIdx.x & ~63); base < nruns;
         base += (int64_t)gridDim.x * blockDim.x) {
        const int32_t r = (int32_t)(base + lane);
        const bool valid = r < nruns;
        int32_t j = -1;
        double gain = 0.0;
        if (valid) {
            const int32_t b = starts[r], e = starts[r + 1];
            const K key = keys[b];
            j = (int32_t)(key >> cbits);
            const int32_t c = (int32_t)(uint32_t)(key & cmask);
            const int32_t v = hubs[j];
            if (!(a.active && !a.active[v])) {
                const int64_t cc = a.curr[v];
                if ((int64_t)c != cc) {
                    double unused = 0.0;
                    const double ey = hub_run_fold(a.edges, vals, b, e, false, v, unused);
                    const double kv = a.vdeg[v];
                    const double eix = own[2 * j] - own[2 * j + 1];
                    const double ax = a.cdeg[cc] - kv;
                    const double ay = a.cdeg[c];
                    gain = 2.0 * (ey - eix) - 2.0 * kv * (ay - ax) * constant;
                }
            }
            gains[r] = gain;
        }
        // A hub has thousands of runs: when the wave's runs all belong to one hub (lane 0's,
        // always valid), one atomic carries the wave's largest gain instead of one per run.
        const unsigned long long gb = gain > 0.0 ? (unsigned long long)__double_as_longlong(gain) : 0ull;
        const int32_t j0 = __shfl(j, 0);
        if (__ballot(valid && j != j0) == 0) {
            unsigned long long m = gb;
            for (int o = 32; o > 0; o >>= 1) {
                const unsigned long long x = __shfl_xor(m, o);
                m = x > m ? x : m;
            }
            if (lane == 0 && m) atomicMax(&best[j0], m);
        } else if (gb) {
            atomicMax(&best[j], gb);
        }
    }
}

// The smallest community among the runs holding their hub's largest gain (the reference's
// tie-break).
template <class K>
__global__ void k_hub_pick(const int32_t* __restrict__ pos, const int32_t* __restrict__ head,
                           const int64_t* __restrict__ hoffs, int32_t nh, const K* __restrict__ keys,
                           const int32_t* __restrict__ starts, int cbits, const double* __restrict__ gains,
                           const unsigned long long* __restrict__ best, int32_t* __restrict__ bestc)
{
    const int32_t nruns = hub_run_count(hoffs, nh, pos, head);
    const K cmask = (K)(((unsigned long long)1 << cbits) - 1);
    for (int32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nruns; r += gridDim.x * blockDim.x) {
        const double gain = gains[r];
        if (!(gain > 0.0)) continue;
        const K key = keys[starts[r]];
        const int32_t j = (int32_t)(key >> cbits);
        if ((unsigned long long)__double_as_longlong(gain) == best[j]) atomicMin(&bestc[j], (int32_t)(uint32_t)(key & cmask));
    }
}

// The same move with a workgroup per hub, for sweeps with enough hubs to fill the chip (the
// coarse levels: thousands of hubs; vt_tuning.hub_block_min sets how many): thread 0 folds the own community's run, the threads take
// the other runs in turn, and the workgroup reduces their (gain, community) argmax.
constexpr int kHubBlock = 256;

template <class K>
__global__ void __launch_bounds__(kHubBlock) k_hub_move(MoveArgs a, const int32_t* __restrict__ hubs, int32_t nh,
                                                        const int64_t* __restrict__ hoffs, const K* __restrict__ keys,
                                                        const uint32_t* __restrict__ vals,
                                                        const int32_t* __restrict__ pos,
                                                        const int32_t* __restrict__ head,
                                                        const int32_t* __restrict__ starts, int cbits)
{
    __shared__ double s_ecc, s_self;
    __shared__ double s_bg[kHubBlock / 64];
    __shared__ int32_t s_bc[kHubBlock / 64];
    unsigned inactive = 0;
    const double constant = *a.constant;
    const int64_t total = hoffs[nh];
    const int32_t nruns = hub_run_count(hoffs, nh, pos, head);
    const K cmask = (K)(((unsigned long long)1 << cbits) - 1);
    for (int32_t j = blockIdx.x; j < nh; j += gridDim.x) {
        const int32_t v = hubs[j];
        if (a.active && !a.active[v]) {
            if (threadIdx.x == 0) {
                a.cw[v] = a.frozen[v];  // louvain.cpp:101-104
                inactive++;
            }
            continue;
        }
        const int32_t r0 = hoffs[j] < total ? pos[hoffs[j]] : nruns;
        const int32_t r1 = hoffs[j + 1] < total ? pos[hoffs[j + 1]] : nruns;
        const int64_t cc = a.curr[v];
        if (threadIdx.x == 0) {
            const K ckey = ((K)j << cbits) | (K)(uint32_t)cc;
            int32_t lo = r0, hi = r1;
            while (lo < hi) {
                const int32_t mid = (lo + hi) >> 1;
                if (keys[starts[mid]] < ckey) lo = mid + 1;
                else hi = mid;
            }
            double self = 0.0, ecc = 0.0;
            if (lo < r1 && keys[starts[lo]] == ckey)
                ecc = hub_run_fold(a.edges, vals, starts[lo], starts[lo + 1], true, v, self);
            s_ecc = ecc;
            s_self = self;
        }
        __syncthreads();
        const double ecc = s_ecc;
        const double kv = a.vdeg[v];
        const double eix = ecc - s_self;
        const double ax = a.cdeg[cc] - kv;
        double bg = 0.0;
        int32_t bc = INT_MAX;
        for (int32_t r = r0 + threadIdx.x; r < r1; r += kHubBlock) {
            const int32_t b = starts[r], e = starts[r + 1];
            const int32_t c = (int32_t)(uint32_t)(keys[b] & cmask);
            if ((int64_t)c == cc) continue;
            double unused = 0.0;
            const double ey = hub_run_fold(a.edges, vals, b, e, false, v, unused);
            const double ay = a.cdeg[c];
            const double gain = 2.0 * (ey - eix) - 2.0 * kv * (ay - ax) * constant;
            argmax_take(bg, bc, gain > 0.0 ? gain : 0.0, gain > 0.0 ? c : INT_MAX);
        }
        for (int o = 32; o > 0; o >>= 1) argmax_take(bg, bc, __shfl_xor(bg, o), __shfl_xor(bc, o));
        if ((threadIdx.x & 63) == 0) {
            s_bg[threadIdx.x / 64] = bg;
            s_bc[threadIdx.x / 64] = bc;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < kHubBlock / 64; w++) argmax_take(bg, bc, s_bg[w], s_bc[w]);
            int64_t tgt = cc;
            if (bg > 0.0) {
                tgt = bc;
                if (a.csize[cc] == 1 && tgt > cc && a.csize[tgt] == 1) tgt = cc;  // louvain.cpp:2275-2276
            }
            a.target[v] = tgt;
            a.cw[v] = ecc;
            if (a.frozen) a.frozen[v] = ecc;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && inactive && a.inactive_count) atomicAdd(a.inactive_count, (unsigned long long)inactive);
}

// Per hub: target (with the singleton rule, louvain.cpp:2275-2276) and clusterWeight, or the
// frozen weight of a hub early termination froze (louvain.cpp:101-104).
__global__ void k_hub_final(MoveArgs a, const int32_t* __restrict__ hubs, int32_t nh, const double* __restrict__ own,
                            const unsigned long long* __restrict__ best, const int32_t* __restrict__ bestc)
{
    for (int32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nh; j += gridDim.x * blockDim.x) {
        const int32_t v = hubs[j];
        if (a.active && !a.active[v]) {
            a.cw[v] = a.frozen[v];
            if (a.inactive_count) atomicAdd(a.inactive_count, 1ull);
            continue;
        }
        const int64_t cc = a.curr[v];
        const double ecc = own[2 * j];
        int64_t tgt = cc;
        if (__longlong_as_double((long long)best[j]) > 0.0) {
            tgt = bestc[j];
            if (a.csize[cc] == 1 && tgt > cc && a.csize[tgt] == 1) tgt = cc;
        }
        a.target[v] = tgt;
        a.cw[v] = ecc;
        if (a.frozen) a.frozen[v] = ecc;
    }
}

__global__ void k_hbm_table_init(int32_t* keys, double* vals, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        keys[i] = -1;
        vals[i] = sent_value();
    }
}

// ------------------------------------------------------------------------------------------
// Init, bucketing, deltas, modularity, rotation
// ------------------------------------------------------------------------------------------

// distSumVertexDegree (louvain.cpp:2161-2186) + distInitComm (2564-2581) + ET state.
// vDegree: left fold of each row's weights from 0.0 in CSR order (distSumVertexDegree,
// louvain.cpp:2161-2186), one thread per row.  Eight loads are issued ahead of the dependent
// adds, so a long row (coarse levels, hubs) streams at memory rate instead of one load
// latency per edge.  Rows of degree >= long_min are left to k_row_fold_wave.
constexpr int kFoldRows = 256;
// Rows at least this long get a wave each when a graph has them (a hub of degree 36 K
// folded by one thread held the whole launch for 3 ms).
constexpr int64_t kFoldLongRow = 2048;

__global__ void __launch_bounds__(kFoldRows) k_row_fold(const int64_t* __restrict__ offs, const DEdge* __restrict__ edges,
                                                        int64_t nv, double* __restrict__ vdeg, int64_t long_min)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = offs[i + 1];
        int64_t k = offs[i];
        if (e - k >= long_min) continue;
        double acc = 0.0;
        for (; k + 8 <= e; k += 8) {
            double w[8];
#pragma unroll
            for (int u = 0; u < 8; u++) w[u] = edges[k + u].weight;
#pragma unroll
            for (int u = 0; u < 8; u++) acc = acc + w[u];
        }
        for (; k < e; k++) acc = acc + edges[k].weight;
        vdeg[i] = acc;
    }
}

// Long rows (the coarse levels, hubs): one wave per row loads 64 weights at a time with
// coalesced reads (the next 64 in flight while these are added), and every lane adds them in
// lane order, each read as a wave-uniform value -- the same left fold as k_row_fold.  Lanes
// past the row's end hold -0.0, which leaves any sum unchanged.  Rows shorter than long_min
// are left to k_row_fold.
constexpr int kFoldWaveRows = 4;
__global__ void __launch_bounds__(64 * kFoldWaveRows) k_row_fold_wave(const int64_t* __restrict__ offs,
                                                                     const DEdge* __restrict__ edges, int64_t nv,
                                                                     double* __restrict__ vdeg, int64_t long_min)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int64_t i = (int64_t)blockIdx.x * kFoldWaveRows + w; i < nv; i += (int64_t)gridDim.x * kFoldWaveRows) {
        const int64_t e0 = offs[i], e1 = offs[i + 1];
        if (e1 - e0 < long_min) continue;  // wave-uniform: i is the same in every lane
        double acc = 0.0;
        double x = (e0 + lane < e1) ? edges[e0 + lane].weight : -0.0;
        for (int64_t base = e0; base < e1; base += 64) {
            const double xn = (base + 64 + lane < e1) ? edges[base + 64 + lane].weight : -0.0;
#pragma unroll
            for (int u = 0; u < 64; u++) acc = acc + readlane_f64(x, u);
            x = xn;
        }
        if (lane == 0) vdeg[i] = acc;
    }
}

__global__ void k_init(int64_t nv, const double* __restrict__ vdeg, double* __restrict__ cdeg,
                       int64_t* __restrict__ csize, int64_t* __restrict__ past, int64_t* __restrict__ curr,
                       int64_t* __restrict__ target, unsigned char* __restrict__ active, double* __restrict__ frozen,
                       double* __restrict__ p_curr, double* __restrict__ p_prev)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x) {
        cdeg[i] = vdeg[i];
        csize[i] = 1;
        past[i] = i;
        curr[i] = i;
        target[i] = 0;
        if (active) {
            active[i] = 1;
            frozen[i] = 0.0;
            p_curr[i] = 0.0;
            p_prev[i] = 1.0;
        }
    }
}

__global__ void k_constant(const double* __restrict__ sums, double* __restrict__ constant)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) constant[0] = 1.0 / sums[0];
}

// Sort keys for the sweep lists: (color bucket, degree bucket).
// In a batch (gid), a vertex whose graph has left the phase gets key `excluded`, past every
// list, so the sweeps no longer visit it.
__global__ void k_list_keys(const int64_t* __restrict__ offs, int64_t nv, const int64_t* __restrict__ colors,
                            int64_t num_color, int lpv_max, BucketBounds bb, const int32_t* __restrict__ gid,
                            const unsigned char* __restrict__ gact, uint32_t excluded, uint32_t* __restrict__ keys,
                            uint32_t* __restrict__ vals)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t d = offs[i + 1] - offs[i];
        uint32_t b = (uint32_t)(d <= lpv_max ? kLpvBucket : degree_bucket(d, bb));
        if (b == kHubBucket - 1) b = d <= bb.wide[0] ? kWide512Bucket : d <= bb.wide[1] ? kWide1024Bucket : b;
        uint32_t c = 0;
        if (colors) c = (uint32_t)(colors[i] < 0 ? num_color - 1 : colors[i]);
        keys[i] = (gid && !gact[gid[i]]) ? excluded : c * kNumBuckets + b;
        vals[i] = (uint32_t)i;
    }
}

__global__ void k_color_keys(int64_t nv, const int64_t* __restrict__ colors, int64_t num_color,
                             uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x) {
        keys[i] = (uint32_t)(colors[i] < 0 ? num_color - 1 : colors[i]);
        vals[i] = (uint32_t)i;
    }
}

// Histogram of small keys: per-block counts in LDS, one global atomic per touched bin.
__global__ void k_key_histogram(const uint32_t* __restrict__ keys, int64_t n, int64_t* __restrict__ hist, int nbins)
{
    extern __shared__ unsigned s_bins[];
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) s_bins[i] = 0;
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (keys[i] < (uint32_t)nbins) atomicAdd(&s_bins[keys[i]], 1u);  // larger keys: excluded
    __syncthreads();
    for (int i = threadIdx.x; i < nbins; i += blockDim.x)
        if (s_bins[i]) atomicAdd((unsigned long long*)&hist[i], (unsigned long long)s_bins[i]);
}

// Movers in processing order: a vertex that was processed (active, or ET off) and whose
// target differs from its community contributes two delta events (louvain.cpp:2343-2359).
__global__ void k_mover_flags(const int32_t* __restrict__ order, int64_t n, const int64_t* __restrict__ target,
                              const int64_t* __restrict__ curr, const unsigned char* __restrict__ active,
                              int32_t* __restrict__ flags)
{
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = order ? order[p] : p;
        const bool live = !active || active[v];
        flags[p] = (live && target[v] != curr[v]) ? 1 : 0;
    }
}

__global__ void k_store_i64(int64_t* __restrict__ p, int64_t v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ void k_add_i64(int64_t* __restrict__ acc, const int64_t* __restrict__ x)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *acc += *x;
}

__global__ void k_scan_total(const int32_t* __restrict__ flags, const int32_t* __restrict__ pos, int64_t n,
                             int64_t* __restrict__ total)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) total[0] = n ? (int64_t)pos[n - 1] + flags[n - 1] : 0;
}

__global__ void k_emit_events(const int32_t* __restrict__ order, int64_t n, const int32_t* __restrict__ flags,
                              const int32_t* __restrict__ pos, const int64_t* __restrict__ target,
                              const int64_t* __restrict__ curr, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                              unsigned long long* __restrict__ long_runs)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *long_runs = 0;  // k_fold_events marks long runs
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        if (!flags[p]) continue;
        const int64_t v = order ? order[p] : p;
        const int64_t q = 2 * (int64_t)pos[p];
        keys[q] = (uint32_t)target[v];
        vals[q] = (uint32_t)v;
        keys[q + 1] = (uint32_t)curr[v];
        vals[q + 1] = (uint32_t)v;
    }
}

// One sorted event's contribution: +k_v when v arrives in the event's community c (its target
// is c), -k_v when it leaves (a - x == a + (-x) exactly); s = +-1, the member-count change.
__device__ __forceinline__ double event_delta(uint32_t c, uint32_t v, const int64_t* __restrict__ target,
                                              const double* __restrict__ vdeg, int& s)
{
    const bool in = (uint64_t)target[v] == (uint64_t)c;
    const double k = vdeg[v];
    s = in ? 1 : -1;
    return in ? k : -k;
}

// Events sorted stably by community keep processing order inside a community, so the
// head of each run folds localCupdate exactly as one thread's sequence of `omp atomic`
// updates does (+= k_i on arrival, -= k_i on departure), then localCinfo += localCupdate
// (louvain.cpp:2534-2547).
// Runs of at least this many events (a community that many movers join or leave: hubs on
// skewed graphs) are folded by a wave each in k_fold_events_long; a thread folding one such
// run held a whole launch for 7 ms on BA n = 2^22.
constexpr int64_t kFoldLongRun = 1024;

__global__ void k_fold_events(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                              const int64_t* __restrict__ target, const double* __restrict__ vdeg, int64_t n,
                              const int64_t* __restrict__ d_movers, double* __restrict__ cdeg,
                              int64_t* __restrict__ csize, uint32_t none, unsigned long long* __restrict__ long_runs)
{
    if (d_movers) n = 2 * d_movers[0];
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t c = keys[j];
        if (c == none || (j > 0 && keys[j - 1] == c)) continue;
        if (j + kFoldLongRun - 1 < n && keys[j + kFoldLongRun - 1] == c) {  // a long run
            *long_runs = 1;
            continue;
        }
        // the community's counters, read before the run so their (random) loads overlap
        // the run's; no other run touches community c
        const double deg0 = cdeg[c];
        const int64_t size0 = csize[c];
        double acc = 0.0;
        int64_t cnt = 0;
        int64_t k = j;
        // keys ascend, so keys[k + 3] == c means keys[k..k+3] all are c
        for (; k + 4 <= n && keys[k + 3] == c; k += 4) {
            int s0, s1, s2, s3;
            const double d0 = event_delta(c, vals[k], target, vdeg, s0);
            const double d1 = event_delta(c, vals[k + 1], target, vdeg, s1);
            const double d2 = event_delta(c, vals[k + 2], target, vdeg, s2);
            const double d3 = event_delta(c, vals[k + 3], target, vdeg, s3);
            acc = acc + d0;
            acc = acc + d1;
            acc = acc + d2;
            acc = acc + d3;
            cnt += s0 + s1 + s2 + s3;
        }
        for (; k < n && keys[k] == c; k++) {
            int s0;
            acc = acc + event_delta(c, vals[k], target, vdeg, s0);
            cnt += s0;
        }
        cdeg[c] = deg0 + acc;
        csize[c] = size0 + cnt;
    }
}

// The runs k_fold_events leaves (kFoldLongRun events or more): each wave scans 64-event
// windows for their starts; for each, the wave loads the run 64 events at a time and adds
// the deltas in event order as wave-uniform lane reads (lanes past the run's end hold -0.0,
// which leaves the sum unchanged) -- the same left fold from 0.0 as k_fold_events.  The
// member-count changes are integers, summed in any order.
__global__ void __launch_bounds__(256) k_fold_events_long(const uint32_t* __restrict__ keys,
                                                          const uint32_t* __restrict__ vals,
                                                          const int64_t* __restrict__ target,
                                                          const double* __restrict__ vdeg, int64_t n,
                                                          const int64_t* __restrict__ d_movers,
                                                          double* __restrict__ cdeg, int64_t* __restrict__ csize,
                                                          uint32_t none, const unsigned long long* __restrict__ long_runs)
{
    if (*long_runs == 0) return;  // k_fold_events found none (the common case)
    if (d_movers) n = 2 * d_movers[0];
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t base = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 64; base < n; base += waves * 64) {
        const int64_t j = base + lane;
        bool start = false;
        if (j < n) {
            const uint32_t c = keys[j];
            start = c != none && (j == 0 || keys[j - 1] != c) && j + kFoldLongRun - 1 < n &&
                    keys[j + kFoldLongRun - 1] == c;
        }
        for (unsigned long long sm = __ballot(start); sm; sm &= sm - 1) {
            const int64_t j0 = base + __builtin_ctzll(sm);
            const uint32_t c = keys[j0];
            double acc = 0.0;
            int64_t cnt = 0;
            for (int64_t k = j0;; k += 64) {
                const int64_t idx = k + lane;
                const bool in = idx < n && keys[idx] == c;
                double d = -0.0;
                int v = 0;
                if (in) d = event_delta(c, vals[idx], target, vdeg, v);
#pragma unroll
                for (int u = 0; u < 64; u++) acc = acc + readlane_f64(d, u);
                for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
                cnt += v;
                if (__ballot(in) != ~0ull) break;  // the run ends inside this window
            }
            if (lane == 0) {
                cdeg[c] = cdeg[c] + acc;
                csize[c] += cnt;
            }
        }
    }
}

// currMod = |e_xx*c - a2_x*c*c| (louvain.cpp:2518-2524).
// movers (or null): the iteration's mover count, copied to movers_out so that one read
// returns it with the value.
__global__ void k_modularity(const double* __restrict__ sums, const double* __restrict__ constant,
                             double* __restrict__ q, const int64_t* __restrict__ movers,
                             int64_t* __restrict__ movers_out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const double c = constant[0];
        q[0] = fabs((sums[0] * c) - (sums[1] * c * c));
        if (movers) movers_out[0] = movers[0];
    }
}

// ET types 1/3 rotation (louvain.cpp:173-183): the freeze test parses as
// ((target == curr) == past), a bool against a community id; vertices that fail it
// rotate even when already frozen.
__global__ void k_rotate_et1(int64_t nv, int num_iters, int64_t* __restrict__ past, int64_t* __restrict__ curr,
                             int64_t* __restrict__ target, unsigned char* __restrict__ active)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = target[i], c = curr[i], p = past[i];
        if (num_iters > 2 && (int64_t)(t == c) == p) active[i] = 0;
        else { past[i] = c; curr[i] = t; target[i] = p; }
    }
}

// ET types 2/4 rotation (louvain.cpp:380-400); the caller swaps p_curr/p_prev afterwards.
__global__ void k_rotate_et2(int64_t nv, int num_iters, double one_minus_delta, int64_t* __restrict__ past,
                             int64_t* __restrict__ curr, int64_t* __restrict__ target,
                             unsigned char* __restrict__ active, double* __restrict__ p_curr,
                             const double* __restrict__ p_prev)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x) {
        if (!active[i]) continue;
        const int64_t c = curr[i], p = past[i];
        bool act = true;
        if (num_iters > 2 && c == p) {
            const double pc = p_prev[i] * one_minus_delta;
            p_curr[i] = pc;
            if (pc <= 0.02) { active[i] = 0; act = false; }
        }
        if (act) { const int64_t t = target[i]; past[i] = c; curr[i] = t; target[i] = p; }
    }
}

__global__ void k_split_edges(const DEdge* __restrict__ edges, int64_t ne, int32_t* __restrict__ tail32,
                              double* __restrict__ wgt)
{
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < ne; j += (int64_t)gridDim.x * blockDim.x) {
        const DEdge e = edges[j];
        tail32[j] = (int32_t)e.tail;
        wgt[j] = e.weight;
    }
}

// The edge split and the degree fold in one pass over the edges, for graphs of short rows: a
// block takes kSfRows consecutive rows and walks their (contiguous) edges in tiles of
// kSfTile with coalesced loads, writing the split copy and staging the weights in LDS;
// thread t then adds row t's staged weights in order, carrying its sum across tiles -- the
// same left fold as k_row_fold.
constexpr int kSfRows = 64;
constexpr int kSfTile = 2048;
__global__ void __launch_bounds__(256) k_split_fold(const int64_t* __restrict__ offs, const DEdge* __restrict__ edges,
                                                    int64_t nv, int32_t* __restrict__ tail32, double* __restrict__ wgt,
                                                    double* __restrict__ vdeg)
{
    __shared__ double st[kSfTile];
    for (int64_t r0 = (int64_t)blockIdx.x * kSfRows; r0 < nv; r0 += (int64_t)gridDim.x * kSfRows) {
        const int nr = (int)(nv - r0 < kSfRows ? nv - r0 : kSfRows);
        const int64_t b0 = offs[r0], b1 = offs[r0 + nr];
        const int t = threadIdx.x;
        int64_t e0 = 0, e1 = 0;
        if (t < nr) {
            e0 = offs[r0 + t];
            e1 = offs[r0 + t + 1];
        }
        double acc = 0.0;
        for (int64_t base = b0; base < b1; base += kSfTile) {
            const int64_t top = base + kSfTile < b1 ? base + kSfTile : b1;
            for (int64_t k = base + t; k < top; k += 256) {
                const DEdge e = edges[k];
                tail32[k] = (int32_t)e.tail;
                wgt[k] = e.weight;
                st[k - base] = e.weight;
            }
            __syncthreads();
            if (t < nr) {
                const int64_t lo = e0 > base ? e0 : base, hi = e1 < top ? e1 : top;
                for (int64_t k = lo; k < hi; k++) acc = acc + st[k - base];
            }
            __syncthreads();
        }
        if (t < nr) vdeg[r0 + t] = acc;
    }
}

__global__ void k_narrow_ids(const int64_t* __restrict__ src, int64_t n, int32_t* __restrict__ dst)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (int32_t)src[i];
}

// Records of vertices 0..n-1, or of the listed vertices live[0..n-1] (a batch's graphs still
// iterating: the sweep reads no other records).
__global__ void k_pack_records(const int64_t* __restrict__ curr, const int64_t* __restrict__ csize,
                               const double* __restrict__ cdeg, int64_t n, const int32_t* __restrict__ live,
                               VRec* __restrict__ rec)
{
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = live ? live[p] : p;
        VRec r;
        r.comm = (int32_t)curr[i];
        r.size = (int32_t)csize[i];
        r.deg = cdeg[i];
        rec[i] = r;
    }
}

// ---- batch of disjoint graphs ------------------------------------------------------------

__global__ void k_seg_constant(const double* __restrict__ sums, int32_t ng, double* __restrict__ constant)
{
    for (int32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += gridDim.x * blockDim.x)
        constant[g] = 1.0 / sums[2 * g];
}

// Two events per vertex at fixed positions (arrival at target, departure from curr); a
// vertex that does not move, or whose graph has left the phase, gets key `none`, which sorts
// after every community.  A stable sort then gives each community its events in vertex
// order, as the one-graph path does with a compacted mover list.
// A batch vertex moves when its graph is still iterating and its target differs.  Flags are
// per position of the live list (the vertices of the graphs iterating at the last relist, in
// vertex order), so the scan and the events follow vertex order.
__global__ void k_batch_flags(int64_t n, const int32_t* __restrict__ live, const int64_t* __restrict__ target,
                              const int64_t* __restrict__ curr, const int32_t* __restrict__ gid,
                              const unsigned char* __restrict__ gact, int32_t* __restrict__ flags)
{
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = live[p];
        flags[p] = (gact[gid[v]] && target[v] != curr[v]) ? 1 : 0;
    }
}

// The live list: the vertices of the graphs still iterating, in vertex order (a flag pass,
// a scan of the flags, then each flagged vertex written at its scanned place).
__global__ void k_live_flags(int64_t nv, const int32_t* __restrict__ gid, const unsigned char* __restrict__ gact,
                             int32_t* __restrict__ flags)
{
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x)
        flags[v] = gact[gid[v]] ? 1 : 0;
}

__global__ void k_live_scatter(int64_t nv, const int32_t* __restrict__ flags, const int32_t* __restrict__ pos,
                               int32_t* __restrict__ live)
{
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x)
        if (flags[v]) live[pos[v]] = (int32_t)v;
}

// The end of the reference loop body per graph (louvain.cpp:540-577): currMod from the
// sums, leave when it gained less than thresh, else prevMod = max(currMod, lower) and rotate.
// status[0] counts the graphs that go on and status[1] their vertices (voff: vertex ranges).
__global__ void k_batch_decide(int32_t ng, const double* __restrict__ sums, const double* __restrict__ constant,
                               const double* __restrict__ thresh, const double* __restrict__ lower,
                               unsigned char* __restrict__ gact, unsigned char* __restrict__ grot,
                               double* __restrict__ prev, int32_t* __restrict__ iters, int32_t max_iters,
                               const int64_t* __restrict__ voff, unsigned long long* __restrict__ status)
{
    for (int32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += gridDim.x * blockDim.x) {
        grot[g] = 0;
        if (!gact[g]) continue;
        iters[g] += 1;
        const double c = constant[g];
        const double q = fabs((sums[2 * g] * c) - (sums[2 * g + 1] * c * c));
        if ((q - prev[g]) < thresh[g]) {
            gact[g] = 0;
            continue;
        }
        prev[g] = q < lower[g] ? lower[g] : q;
        if (max_iters > 0 && iters[g] >= max_iters) {
            gact[g] = 0;
            continue;
        }
        grot[g] = 1;
        atomicAdd(status, 1ull);
        atomicAdd(status + 1, (unsigned long long)(voff[g + 1] - voff[g]));
    }
}

// past <- curr <- target for the graphs that continue (louvain.cpp:565-570), over the live
// list; the 32-bit copy of currComm the sweep reads (curr32) is kept equal to curr here.
__global__ void k_batch_rotate(int64_t n, const int32_t* __restrict__ live, const int32_t* __restrict__ gid,
                               const unsigned char* __restrict__ grot, int64_t* __restrict__ past,
                               int64_t* __restrict__ curr, const int64_t* __restrict__ target,
                               int32_t* __restrict__ curr32)
{
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = live[p];
        if (!grot[gid[v]]) continue;
        const int64_t t = target[v];
        past[v] = curr[v];
        curr[v] = t;
        if (curr32) curr32[v] = (int32_t)t;
    }
}

// ------------------------------------------------------------------------------------------
// Host side
// ------------------------------------------------------------------------------------------

// One color's sweep: buckets 0-3 fused into one launch, then the mid, wide and hub buckets.
struct ColorSweep {
    SweepPlan plan{};
    int64_t fused_blocks = 0;
    int64_t mid_off = 0, mid_n = 0, hub_off = 0, hub_n = 0, lpv_off = 0, lpv_n = 0;
    int64_t wide_off[3] = {0, 0, 0}, wide_n[3] = {0, 0, 0};  // 512-, 1024- and 2048-slot tables
    int32_t index = 0;  // position in Phase::sweeps
};

// Rows up to this degree take the lane-per-vertex kernel when VITE_AMD_LPV_MAXDEG is set
// (default 0: every row goes to the group kernels, which measured faster on the RGG
// workloads because the lane's linear map search is a chain of dependent LDS reads).
static int lpv_max_degree()
{
    const char* e = std::getenv("VITE_AMD_LPV_MAXDEG");
    int d = e ? std::atoi(e) : 0;
    return std::max(0, std::min(d, kLpvMaxDeg));
}

struct Phase {
    vt_ctx* ctx;
    const vt_graph* g;
    int64_t nv;
    bool et = false;
    DBuf<int64_t> past, curr, target, csize;
    DBuf<double> vdeg, cdeg, cw, frozen, pa, pb;
    DBuf<unsigned char> active;
    DBuf<int32_t> lists;     // vertices grouped by (color, degree bucket), vertex order inside
    DBuf<int32_t> corder;    // vertices by color, vertex order inside (event order for color modes)
    DBuf<double> scal;       // [0..1] sums, [2] constant, [3] q, [4] the iteration's movers (int64 bits)
    DBuf<int64_t> iscal;     // [0] movers of the last apply, [1] of the iteration, [2] scan length
    DBuf<unsigned long long> ucount;
    std::vector<ColorSweep> sweeps;      // per color (one entry in plain/vorder)
    std::vector<int64_t> color_ptr;      // colorPtr (louvain.cpp:647-661)
    int64_t hub_cap = 0;
    int hub_blocks = 0;
    // Hub rows by sorting (k_hub_*): per color, the row offsets of its hubs in the key array.
    bool hub_sort = true;
    std::vector<DBuf<int64_t>> hub_offs;
    std::vector<int64_t> hub_total;
    // Batch of disjoint graphs (null for one graph): graph of each vertex, per-graph
    // "still iterating" flags and per-graph 1/2m.
    const int32_t* gid = nullptr;
    const unsigned char* gact = nullptr;
    double* gconst = nullptr;
    // Narrow copies read by the group buckets: edges split once per phase, currComm
    // narrowed before every sweep (it changes between sweeps).
    DBuf<int32_t> tail32, curr32;
    DBuf<double> wgt;
    DBuf<VRec> rec;  // community records, rebuilt before every sweep when the tuning asks
    // A batch narrows currComm once and keeps curr32 current in its rotation (narrow_once);
    // its records are packed for the live list only (pack_live, pack_n) when set.
    bool narrow_once = false, narrowed = false;
    const int32_t* pack_live = nullptr;
    int64_t pack_n = 0;

    Phase(vt_ctx* c, const vt_graph* gr) : ctx(c), g(gr), nv(gr->nv) {}

    void prepare_narrow()
    {
        if constexpr (!kSplitEdges) return;
        if (!tail32.p) {
            const size_t ne1 = (size_t)std::max<int64_t>(g->ne, 1);
            tail32 = DBuf<int32_t>(ctx, ne1);
            wgt = DBuf<double>(ctx, ne1);
            curr32 = DBuf<int32_t>(ctx, (size_t)std::max<int64_t>(nv, 1));
            if (g->ne)
                VT_LAUNCH(ctx, "split_edges", k_split_edges, grid_for(g->ne, 256, 16384), 256, 0,
                          (const DEdge*)g->edges, g->ne, tail32.p, wgt.p);
        }
        if (nv && !(narrow_once && narrowed))
            VT_LAUNCH(ctx, "narrow_curr", k_narrow_ids, grid_for(nv, 256, 8192), 256, 0, (const int64_t*)curr.p, nv, curr32.p);
        narrowed = true;
        if (ctx->tuning.vertex_records && nv) {
            if (!rec.p) rec = DBuf<VRec>(ctx, (size_t)nv);
            const int64_t n = pack_live ? pack_n : nv;
            if (n)
                VT_LAUNCH(ctx, "pack_records", k_pack_records, grid_for(n, 256, 8192), 256, 0, (const int64_t*)curr.p,
                          (const int64_t*)csize.p, (const double*)cdeg.p, n, pack_live, rec.p);
        }
    }

    BucketBounds bounds() const
    {
        BucketBounds bb;
        for (int b = 0; b < 6; b++) bb.mx[b] = ctx->tuning.bucket_max_degree[b];
        bb.wide[0] = ctx->tuning.wide_table_max_degree[0];
        bb.wide[1] = ctx->tuning.wide_table_max_degree[1];
        return bb;
    }

    double* sums() { return scal.p; }
    double* constant() { return scal.p + 2; }
    double* q() { return scal.p + 3; }

    void alloc(bool with_et)
    {
        et = with_et;
        const size_t n = (size_t)std::max<int64_t>(nv, 1);
        past = DBuf<int64_t>(ctx, n);
        curr = DBuf<int64_t>(ctx, n);
        target = DBuf<int64_t>(ctx, n);
        csize = DBuf<int64_t>(ctx, n);
        vdeg = DBuf<double>(ctx, n);
        cdeg = DBuf<double>(ctx, n);
        cw = DBuf<double>(ctx, n);
        if (et) {
            frozen = DBuf<double>(ctx, n);
            pa = DBuf<double>(ctx, n);
            pb = DBuf<double>(ctx, n);
            active = DBuf<unsigned char>(ctx, n);
        }
        scal = DBuf<double>(ctx, 8);
        iscal = DBuf<int64_t>(ctx, 4);
        ucount = DBuf<unsigned long long>(ctx, 4);
    }

    void init_state(bool with_et)
    {
        if (kSplitEdges && !tail32.p && nv && g->ne <= 48 * nv && g->max_degree <= kSfTile) {
            // short rows (a row's fold is one thread's, so no long ones): the split copy the
            // local move reads and the degrees in one pass
            const size_t ne1 = (size_t)std::max<int64_t>(g->ne, 1);
            tail32 = DBuf<int32_t>(ctx, ne1);
            wgt = DBuf<double>(ctx, ne1);
            curr32 = DBuf<int32_t>(ctx, (size_t)nv);
            VT_LAUNCH(ctx, "split_fold", k_split_fold, grid_for(nv, kSfRows, 65536), 256, 0, g->offs, g->edges, nv,
                      tail32.p, wgt.p, vdeg.p);
        } else if (nv && nv < 65536 && g->ne > 48 * nv) {  // few, long rows: a thread per row would leave the GPU idle
            VT_LAUNCH(ctx, "row_fold", k_row_fold_wave, grid_for(nv, kFoldWaveRows, 65536), 64 * kFoldWaveRows, 0, g->offs,
                      g->edges, nv, vdeg.p, (int64_t)0);
        } else if (nv) {
            // a thread per row, and a wave per row for the rows of kFoldLongRow or more
            const bool split = g->max_degree >= kFoldLongRow;
            VT_LAUNCH(ctx, "row_fold", k_row_fold, grid_for(nv, kFoldRows, 16384), kFoldRows, 0, g->offs, g->edges, nv,
                      vdeg.p, split ? kFoldLongRow : INT64_MAX);
            if (split)
                VT_LAUNCH(ctx, "row_fold_long", k_row_fold_wave, grid_for(nv, kFoldWaveRows, 16384), 64 * kFoldWaveRows, 0,
                          g->offs, g->edges, nv, vdeg.p, kFoldLongRow);
        }
        VT_LAUNCH(ctx, "init", k_init, grid_for(nv, 256, 8192), 256, 0, nv, (const double*)vdeg.p, cdeg.p, csize.p, past.p,
                  curr.p, target.p, with_et ? active.p : nullptr, with_et ? frozen.p : nullptr,
                  with_et ? pa.p : nullptr, with_et ? pb.p : nullptr);
    }

    void init(int flags)
    {
        init_state(et);
        if (flags & VT_FLAG_SEQUENTIAL_SUMS) seq_sum2(ctx, vdeg.p, nullptr, nv, sums());
        else tree_sum2(ctx, vdeg.p, nullptr, nv, sums());
        VT_LAUNCH(ctx, "constant", k_constant, 1, 64, 0, (const double*)sums(), constant());
    }

    // Sorts vertices into (color, degree bucket) lists and, with colors, the per-color event
    // order; per color, plans one block range for buckets 0-3 (run as two launches: bucket 2
    // alone, buckets 0, 1 and 3 together).
    void build_lists(const int64_t* d_colors, int64_t num_color, bool want_corder)
    {
        const int64_t ncol = d_colors && !want_corder ? num_color : 1;
        const size_t n1 = (size_t)std::max<int64_t>(nv, 1);
        lists = DBuf<int32_t>(ctx, n1);
        uint32_t* k0 = (uint32_t*)scratch(ctx, "lk0", 4 * n1);
        uint32_t* k1 = (uint32_t*)scratch(ctx, "lk1", 4 * n1);
        uint32_t* v1 = (uint32_t*)scratch(ctx, "lv1", 4 * n1);
        const int64_t nkeys = ncol * kNumBuckets;
        int64_t* hist = (int64_t*)scratch(ctx, "lhist", 8 * (size_t)(nkeys + 1));
        std::vector<int64_t> h(nkeys, 0);
        if (nv) {
            VT_LAUNCH(ctx, "list_keys", k_list_keys, grid_for(nv, 256, 8192), 256, 0, g->offs, nv,
                      ncol > 1 ? d_colors : nullptr, num_color, gid ? 0 : lpv_max_degree(), bounds(), gid, gact,
                      (uint32_t)nkeys, k0, (uint32_t*)lists.p);
            dzero(ctx, hist, 8 * nkeys);
            key_histogram(k0, nv, hist, (int)nkeys);
            bool alt = false;
            sort_pairs_u32(ctx, k0, (uint32_t*)lists.p, nv, (int)bits_for(nkeys), k1, v1, &alt);
            if (alt) d2d(ctx, lists.p, v1, 4 * (size_t)nv);
            d2h(ctx, h.data(), hist, 8 * nkeys);
        }
        sweeps.assign(ncol, ColorSweep{});
        int64_t off = 0, max_hub = 0;
        constexpr int kGroups[kFusedBuckets] = {256 / 4, 256 / 8, 256 / 16, 256 / 32};
        for (int64_t c = 0; c < ncol; c++) {
            ColorSweep& cs = sweeps[c];
            int64_t blocks = 0;
            for (int b = 0; b < kNumBuckets; b++) {
                const int64_t cnt = h[c * kNumBuckets + b];
                if (b < kFusedBuckets) {
                    cs.plan.bbeg[b] = blocks;
                    cs.plan.off[b] = off;
                    cs.plan.n[b] = cnt;
                    const int64_t nb = std::min<int64_t>((cnt + kGroups[b] - 1) / kGroups[b],
                                                         ctx->tuning.blocks_per_bucket);
                    blocks += (nb + kXcds - 1) / kXcds * kXcds;
                } else if (b == kFusedBuckets) {
                    cs.mid_off = off;
                    cs.mid_n = cnt;
                } else if (b == kWide512Bucket || b == kWide1024Bucket || b == kHubBucket - 1) {
                    const int w = b == kWide512Bucket ? 0 : b == kWide1024Bucket ? 1 : 2;
                    cs.wide_off[w] = off;
                    cs.wide_n[w] = cnt;
                } else if (b == kLpvBucket) {
                    cs.lpv_off = off;
                    cs.lpv_n = cnt;
                } else {
                    cs.hub_off = off;
                    cs.hub_n = cnt;
                    max_hub = std::max(max_hub, cnt);
                }
                off += cnt;
            }
            cs.plan.bbeg[kFusedBuckets] = blocks;
            cs.fused_blocks = blocks;
        }
        for (int64_t c = 0; c < ncol; c++) sweeps[c].index = (int32_t)c;
        // Hub rows: by sorting, unless the sweep must stay free of library sorts (a batch
        // replays its sweeps from captured graphs) or edge indexes exceed 32 bits.
        // the sorted path keeps run positions in 32 bits: hub entries (at most ne) below 2^31
        hub_sort = hub_sort && !gid && g->ne < (int64_t(1) << 31);
        hub_offs.clear();
        hub_total.assign(ncol, 0);
        if (hub_sort) {
            for (int64_t c = 0; c < ncol; c++) {
                const ColorSweep& cs = sweeps[c];
                hub_offs.emplace_back(ctx, (size_t)cs.hub_n + 1);
                if (!cs.hub_n) continue;
                DBuf<int64_t> deg(ctx, (size_t)cs.hub_n + 1);
                VT_LAUNCH(ctx, "hub_degrees", k_hub_degrees, grid_for(cs.hub_n + 1, 256, 1024), 256, 0,
                          (const int32_t*)(lists.p + cs.hub_off), (int32_t)cs.hub_n, (const int64_t*)g->offs, deg.p);
                exclusive_scan_i64(ctx, deg.p, hub_offs.back().p, cs.hub_n + 1);
                d2h(ctx, &hub_total[c], hub_offs.back().p + cs.hub_n, 8);
            }
        }
        // Hub tables: one HBM table per resident wave, sized for the largest row.
        hub_cap = 0;
        if (max_hub > 0 && !hub_sort) {
            int64_t cap = 1;
            while (cap < 2 * (g->max_degree + 1)) cap <<= 1;
            hub_cap = cap;
            hub_blocks = (int)std::min<int64_t>(max_hub, 1024);
            const size_t tn = (size_t)hub_cap * hub_blocks;
            int32_t* hk = (int32_t*)scratch(ctx, "hub_keys", 4 * tn);
            double* hv = (double*)scratch(ctx, "hub_vals", 8 * tn);
            scratch(ctx, "hub_occ", 4 * tn);
            VT_LAUNCH(ctx, "hub_init", k_hbm_table_init, grid_for((int64_t)tn, 256, 8192), 256, 0, hk, hv, (int64_t)tn);
        }
        if (d_colors) {
            // colorIndex: vertices by color, ascending inside a color (louvain.cpp:662-677 run
            // by one thread).
            corder = DBuf<int32_t>(ctx, n1);
            int64_t* ch = (int64_t*)scratch(ctx, "chist", 8 * (size_t)(num_color + 1));
            std::vector<int64_t> hc(num_color, 0);
            if (nv) {
                VT_LAUNCH(ctx, "color_keys", k_color_keys, grid_for(nv, 256, 8192), 256, 0, nv, d_colors, num_color, k0,
                          (uint32_t*)corder.p);
                dzero(ctx, ch, 8 * num_color);
                key_histogram(k0, nv, ch, (int)num_color);
                bool alt = false;
                sort_pairs_u32(ctx, k0, (uint32_t*)corder.p, nv, (int)bits_for(num_color), k1, v1, &alt);
                if (alt) d2d(ctx, corder.p, v1, 4 * (size_t)nv);
                d2h(ctx, hc.data(), ch, 8 * num_color);
            }
            color_ptr.assign(num_color + 1, 0);
            for (int64_t c = 0; c < num_color; c++) color_ptr[c + 1] = color_ptr[c] + hc[c];
        }
    }

    void key_histogram(const uint32_t* keys, int64_t n, int64_t* hist, int nbins)
    {
        if (nbins > 16384) fail(ctx, VT_ERR_LIMIT, "too many sweep lists (%d)", nbins);
        VT_LAUNCH(ctx, "key_hist", k_key_histogram, grid_for(n, 256, 2048), 256, sizeof(unsigned) * nbins, keys, n, hist,
                  nbins);
    }

    MoveArgs move_args()
    {
        MoveArgs a;
        a.offs = g->offs;
        a.edges = g->edges;
        a.curr = curr.p;
        a.vdeg = vdeg.p;
        a.cdeg = cdeg.p;
        a.csize = csize.p;
        a.active = et ? active.p : nullptr;
        a.frozen = et ? frozen.p : nullptr;
        a.target = target.p;
        a.cw = cw.p;
        a.constant = gid ? gconst : constant();
        a.inactive_count = ucount.p;
        a.gid = gid;
        a.gact = gact;
        a.tail32 = tail32.p;
        a.wgt = wgt.p;
        a.curr32 = curr32.p;
        a.rec = (ctx->tuning.vertex_records && rec.p) ? rec.p : nullptr;
        return a;
    }

    // Hub rows: a key per hub edge, (hub index << cbits | neighbour community) with the edge
    // index as the value, one stable sort of all of them, the run bounds, then a workgroup per
    // hub.  Keys of at most 32 bits (the usual case: community ids below 2^22 and a few hundred
    // hubs) take the device sort and scan, counted by hoffs[nh] (the hub entry total); wider
    // keys take the 64-bit library sort.
    template <class K>
    void hub_sweep(const MoveArgs& a, const ColorSweep& cs, int64_t total, int cbits, int kbits)
    {
        static_assert(kSplitEdges, "the hub keys read the 32-bit tail and community copies");
        const int32_t nh = (int32_t)cs.hub_n;
        // many hubs, none holding much of the work (the coarse levels): a workgroup per hub
        // sorts its segment and moves it; otherwise (BA's level 0: a few hundred or thousand
        // hubs, the largest with 36 K of 0.6-1.5 M entries) one sort of all the entries, and
        // the flat move launches
        const int32_t hmin = ctx->tuning.hub_block_min;
        const bool many = hmin == 0 || (nh >= hmin && 64 * g->max_degree <= total);
        const int32_t* hubs = (const int32_t*)(lists.p + cs.hub_off);
        const int64_t* hoffs = hub_offs[cs.index].p;
        const size_t t1 = (size_t)std::max<int64_t>(total, 1);
        K* k0 = (K*)scratch(ctx, "hs_k0", 8 * t1);
        K* k1 = (K*)scratch(ctx, "hs_k1", 8 * t1);
        uint32_t* v0 = (uint32_t*)scratch(ctx, "hs_v0", 4 * t1);
        uint32_t* v1 = (uint32_t*)scratch(ctx, "hs_v1", 4 * t1);
        int32_t* head = (int32_t*)scratch(ctx, "hs_head", 4 * t1);
        int32_t* pos = (int32_t*)scratch(ctx, "hs_pos", 4 * t1);
        int32_t* starts = (int32_t*)scratch(ctx, "hs_starts", 4 * t1 + 4);
        double* own = (double*)scratch(ctx, "hs_own", 16 * (size_t)nh);
        unsigned long long* best = (unsigned long long*)scratch(ctx, "hs_best", 8 * (size_t)nh);
        int32_t* bestc = (int32_t*)scratch(ctx, "hs_bestc", 4 * (size_t)nh);
        double* gains = (double*)scratch(ctx, "hs_gains", 8 * t1);
        const K* ks;
        const uint32_t* vs;
        {
            TimedScope ts(ctx, "local_move_hub_sort");
            const char* kname = sizeof(K) == 4 ? "hub_keys32" : "hub_keys64";
            if (many)
                VT_LAUNCH(ctx, kname, k_hub_keys_rows<K>, (unsigned)std::min<int64_t>(nh, 16384), 256, 0, hubs, nh, hoffs,
                          (const int64_t*)g->offs, (const int32_t*)tail32.p, (const int32_t*)curr32.p, cbits, k0, v0);
            else
                VT_LAUNCH(ctx, kname, k_hub_keys<K>, grid_for(total, 256, 16384), 256, 0, hubs, nh, hoffs,
                          (const int64_t*)g->offs, (const int32_t*)tail32.p, (const int32_t*)curr32.p, cbits, k0, v0);
            bool alt = false;
            if (many) {
                // each hub's entries are already together: sort each hub's segment by the
                // community bits alone (the hub bits are the same throughout a segment)
                rocprim::double_buffer<K> kb(k0, k1);
                rocprim::double_buffer<uint32_t> vb(v0, v1);
                size_t tb = 0;
                VT_HIP(ctx, rocprim::segmented_radix_sort_pairs(nullptr, tb, kb, vb, (unsigned)total, (unsigned)nh, hoffs,
                                                                 hoffs + 1, 0u, (unsigned)cbits, ctx->stream));
                void* tmp = scratch(ctx, "hs_segsort", tb);
                VT_HIP(ctx, rocprim::segmented_radix_sort_pairs(tmp, tb, kb, vb, (unsigned)total, (unsigned)nh, hoffs,
                                                                 hoffs + 1, 0u, (unsigned)cbits, ctx->stream));
                alt = kb.current() == k1;
            } else if constexpr (sizeof(K) == 4) {
                alt = dev_sort_pairs_u32(ctx, k0, v0, k1, v1, hoffs + nh, 1, total, kbits);
            } else {
                sort_pairs_u64(ctx, (uint64_t*)k0, v0, total, kbits, (uint64_t*)k1, v1, &alt);
            }
            ks = alt ? k1 : k0;
            vs = alt ? v1 : v0;
            VT_LAUNCH(ctx, "hub_heads", k_hub_heads<K>, grid_for(total, 256, 16384), 256, 0, ks, total, head);
            if constexpr (sizeof(K) == 4) dev_scan_i32(ctx, head, pos, hoffs + nh, total, nullptr, "hs_part");
            else exclusive_scan_i32(ctx, head, pos, total);
            VT_LAUNCH(ctx, "hub_starts", k_hub_starts, grid_for(total, 256, 16384), 256, 0, (const int32_t*)head,
                      (const int32_t*)pos, total, starts);
        }
        TimedScope ts(ctx, "local_move_hub");
        if (many) {
            VT_LAUNCH(ctx, "hub_move_blocks", k_hub_move<K>, (unsigned)std::min<int64_t>(nh, 8192), kHubBlock, 0, a, hubs,
                      nh, hoffs, ks, vs, (const int32_t*)pos, (const int32_t*)head, (const int32_t*)starts, cbits);
            return;
        }
        VT_LAUNCH(ctx, "hub_own", k_hub_own<K>, grid_for(nh, 64, 4096), 64, 0, a, hubs, nh, hoffs, ks, vs,
                  (const int32_t*)pos, (const int32_t*)head, (const int32_t*)starts, cbits, own, best, bestc);
        VT_LAUNCH(ctx, "hub_runs", k_hub_runs<K>, grid_for(total, 256, 16384), 256, 0, a, hubs, nh, hoffs, ks, vs,
                  (const int32_t*)pos, (const int32_t*)head, (const int32_t*)starts, cbits, (const double*)own, gains, best);
        VT_LAUNCH(ctx, "hub_pick", k_hub_pick<K>, grid_for(total, 256, 16384), 256, 0, (const int32_t*)pos,
                  (const int32_t*)head, hoffs, nh, ks, (const int32_t*)starts, cbits, (const double*)gains,
                  (const unsigned long long*)best, bestc);
        VT_LAUNCH(ctx, "hub_final", k_hub_final, grid_for(nh, 64, 4096), 64, 0, a, hubs, nh, (const double*)own,
                  (const unsigned long long*)best, (const int32_t*)bestc);
    }

    void sweep(const ColorSweep& cs)
    {
        prepare_narrow();
        const MoveArgs a = move_args();
        if (cs.lpv_n) {
            const int64_t nblk = std::min<int64_t>((cs.lpv_n + 63) / 64, 16384);
            const int64_t chunk = ((cs.lpv_n + nblk - 1) / nblk + 63) / 64 * 64;
            auto kern = k_local_move_lpv<kLpvMaxDeg, kLpvArena, 4>;
            VT_LAUNCH(ctx, "local_move_lpv", kern, (unsigned)nblk, 64, 0, a, (const int32_t*)(lists.p + cs.lpv_off), cs.lpv_n,
                      chunk);
        }
        const bool batch = a.gid != nullptr;
        if (cs.fused_blocks) {
            const int64_t n16 = cs.plan.bbeg[3] - cs.plan.bbeg[2];
            if (cs.fused_blocks > n16) {
                auto kern = a.rec ? (batch ? k_local_move<true, kPartRest, true> : k_local_move<false, kPartRest, true>)
                                  : (batch ? k_local_move<true, kPartRest, false> : k_local_move<false, kPartRest, false>);
                VT_LAUNCH(ctx, "local_move_rest", kern, (unsigned)(cs.fused_blocks - n16), kFusedBlock, 0, a,
                          (const int32_t*)lists.p, cs.plan);
            }
            if (n16) {
                auto kern = a.rec ? (batch ? k_local_move<true, kPartG16, true> : k_local_move<false, kPartG16, true>)
                                  : (batch ? k_local_move<true, kPartG16, false> : k_local_move<false, kPartG16, false>);
                VT_LAUNCH(ctx, "local_move", kern, (unsigned)n16, kFusedBlock, 0, a, (const int32_t*)lists.p, cs.plan);
            }
        }
        if (cs.mid_n) {
            auto kern = a.rec ? (batch ? k_local_move_mid<true, true> : k_local_move_mid<false, true>)
                              : (batch ? k_local_move_mid<true, false> : k_local_move_mid<false, false>);
            const int64_t nb = std::min<int64_t>((cs.mid_n + 3) / 4, ctx->tuning.blocks_per_bucket);
            VT_LAUNCH(ctx, "local_move_mid", kern, (unsigned)((nb + kXcds - 1) / kXcds * kXcds), 256, 0, a,
                      (const int32_t*)(lists.p + cs.mid_off), cs.mid_n);
        }
        for (int w = 0; w < 3; w++) {
            if (!cs.wide_n[w]) continue;
            auto kern = w == 0 ? k_local_move_wide<512> : w == 1 ? k_local_move_wide<1024> : k_local_move_wide<2048>;
            VT_LAUNCH(ctx, "local_move_wide", kern, (unsigned)std::min<int64_t>(cs.wide_n[w], 16384), 64, 0, a,
                      (const int32_t*)(lists.p + cs.wide_off[w]), cs.wide_n[w]);
        }
        if (cs.hub_n && hub_sort) {
            const int64_t total = hub_total[cs.index];
            const int cbits = (int)bits_for(nv > 1 ? nv - 1 : 1);  // community ids < nv
            const int kbits = cbits + (int)bits_for(cs.hub_n);
            if (kbits <= 32 && total < (int64_t(1) << 31)) hub_sweep<uint32_t>(a, cs, total, cbits, kbits);
            else hub_sweep<unsigned long long>(a, cs, total, cbits, kbits);
        } else if (cs.hub_n) {
            const int blocks = (int)std::min<int64_t>(cs.hub_n, hub_blocks);
            VT_LAUNCH(ctx, "local_move_hub", k_move_hbm, (unsigned)blocks, 64, 0, a, (const int32_t*)(lists.p + cs.hub_off),
                      cs.hub_n, (int32_t*)scratch(ctx, "hub_keys", 0), (double*)scratch(ctx, "hub_vals", 0),
                      (int32_t*)scratch(ctx, "hub_occ", 0), (uint32_t)hub_cap);
        }
    }

    // Folds the sweep's moves into localCinfo in processing order, with no host read: the
    // mover count stays on the device (the scan's total sizes the event sort and folds) and
    // is added to iscal[1], the iteration's movers, which the host reads with Q.
    void apply_moves(const int32_t* order, int64_t n, double* dst_cdeg, int64_t* dst_csize)
    {
        if (n == 0) return;
        int32_t* flags = (int32_t*)scratch(ctx, "mv_flags", 4 * (size_t)n);
        int32_t* pos = (int32_t*)scratch(ctx, "mv_pos", 4 * (size_t)n);
        int64_t* d_movers = iscal.p;
        int64_t* d_n = iscal.p + 2;
        VT_LAUNCH(ctx, "store_count", k_store_i64, 1, 64, 0, d_n, n);
        VT_LAUNCH(ctx, "mover_flags", k_mover_flags, grid_for(n, 256, 8192), 256, 0, order, n, target.p, curr.p,
                  et ? active.p : nullptr, flags);
        dev_scan_i32(ctx, flags, pos, d_n, n, d_movers, "mv_part");
        const int64_t ne = 2 * n;  // events: at most two per vertex
        uint32_t* k0 = (uint32_t*)scratch(ctx, "ev_k0", 4 * (size_t)ne);
        uint32_t* v0 = (uint32_t*)scratch(ctx, "ev_v0", 4 * (size_t)ne);
        uint32_t* k1 = (uint32_t*)scratch(ctx, "ev_k1", 4 * (size_t)ne);
        uint32_t* v1 = (uint32_t*)scratch(ctx, "ev_v1", 4 * (size_t)ne);
        VT_LAUNCH(ctx, "emit_events", k_emit_events, grid_for(n, 256, 8192), 256, 0, order, n, (const int32_t*)flags,
                  (const int32_t*)pos, target.p, curr.p, k0, v0, ucount.p + 1);
        const bool alt = dev_sort_pairs_u32(ctx, k0, v0, k1, v1, d_movers, 2, ne, (int)bits_for(nv > 1 ? nv - 1 : 1));
        fold_events(alt ? k1 : k0, alt ? v1 : v0, ne, 0xFFFFFFFFu, dst_cdeg, dst_csize, d_movers);
        VT_LAUNCH(ctx, "count_movers", k_add_i64, 1, 64, 0, iscal.p + 1, (const int64_t*)d_movers);
    }

    // The iteration's mover count: read with Q by modularity(), else from the device.
    int64_t movers_read = -1;
    int64_t iteration_movers()
    {
        if (movers_read >= 0) return movers_read;
        int64_t m = 0;
        d2h(ctx, &m, iscal.p + 1, 8);
        return m;
    }

    // localCinfo += the sorted events' folds, per community in processing order.
    // ne: the event count, or its upper bound when d_movers (the mover count on the device)
    // is given.
    // The long-run flag (ucount[1]) was cleared by the emission of these events.
    void fold_events(const uint32_t* ks, const uint32_t* vs, int64_t ne, uint32_t none, double* dst_cdeg,
                     int64_t* dst_csize, const int64_t* d_movers = nullptr)
    {
        VT_LAUNCH(ctx, "fold_events", k_fold_events, grid_for(ne, 256, 16384), 256, 0, ks, vs, (const int64_t*)target.p,
                  (const double*)vdeg.p, ne, d_movers, dst_cdeg, dst_csize, none, ucount.p + 1);
        if (ne >= kFoldLongRun)
            VT_LAUNCH(ctx, "fold_events_long", k_fold_events_long, grid_for(ne, 256, 4096), 256, 0, ks, vs,
                      (const int64_t*)target.p, (const double*)vdeg.p, ne, d_movers, dst_cdeg, dst_csize, none,
                      (const unsigned long long*)(ucount.p + 1));
    }

    double modularity(int flags)
    {
        if (flags & VT_FLAG_SEQUENTIAL_SUMS) seq_sum2(ctx, cw.p, cdeg.p, nv, sums());
        else tree_sum2(ctx, cw.p, cdeg.p, nv, sums());
        VT_LAUNCH(ctx, "modularity", k_modularity, 1, 64, 0, (const double*)sums(), (const double*)constant(), q(),
                  (const int64_t*)(iscal.p + 1), (int64_t*)(scal.p + 4));
        double qm[2];  // Q and the movers' bits, one read
        d2h(ctx, qm, q(), 16);
        std::memcpy(&movers_read, &qm[1], 8);
        return qm[0];
    }
};

}  // namespace vt

// A phase in progress: device-resident state between iterations (the locals of
// distLouvainMethod*, louvain.cpp:14-50).
struct vt_phase {
    vt::Phase ph;
    vt_graph* g;
    vt_phase_params p;
    int mode = VT_MODE_PLAIN, et = 0;
    double prevMod = 0.0, currMod = -1.0;
    int numIters = 0;
    bool done = false;
    vt::DBuf<int64_t> dcolors;
    vt_phase(vt_ctx* c, vt_graph* gr) : ph(c, gr), g(gr) {}
};

namespace vt {

vt_phase* phase_begin(vt_ctx* ctx, vt_graph* g, const vt_phase_params* p)
{
    int mode = p->mode, et = p->et_type;
    if (mode < 0 || mode > 2) fail(ctx, VT_ERR_ARG, "vt_louvain_phase: bad mode %d", mode);
    if (et < 0 || et > 4) fail(ctx, VT_ERR_ARG, "vt_louvain_phase: et_type must be 0..4 (got %d)", et);
    if (mode != VT_MODE_PLAIN) {
        if (p->num_color < 1) fail(ctx, VT_ERR_ARG, "vt_louvain_phase: num_color must be >= 1");
        if (p->num_color == 1) { mode = VT_MODE_PLAIN; et = 0; }  // louvain.cpp:601-605
        else if (!p->colors) fail(ctx, VT_ERR_ARG, "vt_louvain_phase: colors required for mode %d", mode);
    }
    const int64_t nv = g->nv;
    vt_phase* st = new vt_phase(ctx, g);
    try {
        st->p = *p;
        st->p.colors = nullptr;
        st->mode = mode;
        st->et = et;
        st->prevMod = p->lower;
        Phase& ph = st->ph;
        ph.alloc(et != 0);
        ph.init(p->flags);
        if (mode != VT_MODE_PLAIN) {
            for (int64_t i = 0; i < nv; i++)
                if (p->colors[i] > p->num_color - 1 || p->colors[i] < -1)
                    fail(ctx, VT_ERR_ARG, "vt_louvain_phase: color %lld of vertex %lld outside [-1, num_color-1]",
                         (long long)p->colors[i], (long long)i);
            st->dcolors = DBuf<int64_t>(ctx, (size_t)std::max<int64_t>(nv, 1));
            h2d(ctx, st->dcolors.p, p->colors, 8 * (size_t)nv);
        }
        // Sweep lists per color for COLOR; one Jacobi sweep (event order = colorIndex) for VORDER.
        ph.build_lists(mode != VT_MODE_PLAIN ? st->dcolors.p : nullptr, p->num_color, mode == VT_MODE_VORDER);
    } catch (...) {
        delete st;
        throw;
    }
    ctx->trace_q.clear();
    ctx->trace_moves.clear();
    return st;
}

// One iteration of the reference loop body (louvain.cpp:473-577 and siblings); sets done
// when the reference would leave the loop.
void phase_step(vt_ctx* ctx, vt_phase* st)
{
    if (st->done) return;
    Phase& ph = st->ph;
    const vt_phase_params& p = st->p;
    const int64_t nv = ph.nv;
    st->numIters++;
    if (ph.et) dzero(ctx, ph.ucount.p, 8);
    dzero(ctx, ph.iscal.p + 1, 8);  // the iteration's movers
    ph.movers_read = -1;
    if (st->mode == VT_MODE_COLOR) {
        for (int64_t ci = 0; ci < p.num_color; ci++) {
            ph.sweep(ph.sweeps[ci]);
            const int64_t a0 = ph.color_ptr[ci], a1 = ph.color_ptr[ci + 1];
            ph.apply_moves(ph.corder.p + a0, a1 - a0, ph.cdeg.p, ph.csize.p);
        }
    } else {
        ph.sweep(ph.sweeps[0]);
    }
    if (st->et == 3 || st->et == 4) {
        unsigned long long vc = 0;
        d2h(ctx, &vc, ph.ucount.p, 8);
        if (vc >= 90) {  // ET_CUTOFF compared with a count (louvain.cpp:116-121)
            st->currMod = -1.0;
            ctx->trace_q.push_back(NAN);
            ctx->trace_moves.push_back(ph.iteration_movers());
            st->done = true;
            return;
        }
    }
    if (st->mode == VT_MODE_PLAIN) ph.apply_moves(nullptr, nv, ph.cdeg.p, ph.csize.p);
    else if (st->mode == VT_MODE_VORDER) ph.apply_moves(ph.corder.p, nv, ph.cdeg.p, ph.csize.p);

    st->currMod = ph.modularity(p.flags);
    ctx->trace_q.push_back(st->currMod);
    ctx->trace_moves.push_back(ph.iteration_movers());
    if ((st->currMod - st->prevMod) < p.thresh) { st->done = true; return; }
    st->prevMod = st->currMod;
    if (st->prevMod < p.lower) st->prevMod = p.lower;
    if (p.max_iters > 0 && st->numIters >= p.max_iters) { st->done = true; return; }

    if (st->et == 0) {
        // past <- curr <- target <- old past (louvain.cpp:565-570), by pointer rotation.
        std::swap(ph.past.p, ph.curr.p);
        std::swap(ph.curr.p, ph.target.p);
    } else if (st->et == 1 || st->et == 3) {
        VT_LAUNCH(ctx, "rotate_et", k_rotate_et1, grid_for(nv, 256, 8192), 256, 0, nv, st->numIters, ph.past.p,
                  ph.curr.p, ph.target.p, ph.active.p);
    } else {
        VT_LAUNCH(ctx, "rotate_et", k_rotate_et2, grid_for(nv, 256, 8192), 256, 0, nv, st->numIters,
                  1.0 - p.et_delta, ph.past.p, ph.curr.p, ph.target.p, ph.active.p, ph.pa.p, (const double*)ph.pb.p);
        std::swap(ph.pa.p, ph.pb.p);  // std::swap(p_curr, p_prev)
    }
}

// cvect = pastComm, iters = numIters, return prevMod (louvain.cpp:579-590); a device copy
// of the partition stays on the graph for the rebuild.
double phase_end(vt_ctx* ctx, vt_phase* st, int64_t* cvect_out, int32_t* iters_out)
{
    Phase& ph = st->ph;
    vt_graph* g = st->g;
    const int64_t nv = g->nv;
    if (!g->part) g->part = (int64_t*)dalloc(ctx, 8 * (size_t)std::max<int64_t>(nv, 1));
    d2d(ctx, g->part, ph.past.p, 8 * (size_t)nv);
    if (cvect_out) d2h(ctx, cvect_out, ph.past.p, 8 * (size_t)nv);
    sync(ctx);
    if (iters_out) *iters_out = st->numIters;
    return st->prevMod;
}

double run_phase(vt_ctx* ctx, vt_graph* g, const vt_phase_params* p, int64_t* cvect_out, int32_t* iters_out)
{
    vt_phase* st = phase_begin(ctx, g, p);
    try {
        while (!st->done) phase_step(ctx, st);
        const double q = phase_end(ctx, st, cvect_out, iters_out);
        delete st;
        return q;
    } catch (...) {
        delete st;
        throw;
    }
}

// One Louvain phase (plain mode) of every graph of a batch at once: all graphs iterate in
// lockstep, a graph that stops improving leaves the loop and stays frozen while the others
// go on.  Each graph sees exactly the arithmetic it would see alone (the local move reads
// only its own vertices and communities and its own 1/2m; the delta folds are per
// community; sums use the per-graph trees of seg_tree_sum2), so per graph the return value,
// iteration count and partition equal those of run_phase on that graph alone.  The
// partition is left on `u` (u->part) for the rebuild.
//
// With `use_graph`, every round is replayed from one captured HIP graph (sweep, movers' scan,
// device sort of the events, folds, sums, decisions, rotation), except round 0 and the round
// after each relist, which run directly while the host captures the next one.
void batch_phase(vt_ctx* ctx, vt_graph* u, const int32_t* d_gid, const SegSums& seg, const std::vector<double>& lower,
                 const std::vector<double>& thresh, int flags, bool use_graph, std::vector<double>& q_out,
                 std::vector<int32_t>& iters_out, int* rounds_out, int* relists_out)
{
    const int32_t ng = seg.ng;
    const int64_t nv = u->nv;
    Phase ph(ctx, u);
    ph.alloc(false);
    DBuf<double> gconst(ctx, ng), gsums(ctx, 2 * (size_t)ng), gprev(ctx, ng), gthr(ctx, ng), glow(ctx, ng);
    DBuf<unsigned char> gact(ctx, ng), grot(ctx, ng);
    DBuf<int32_t> giters(ctx, ng);
    // [0] graphs still iterating, [1] their vertices (set by each round's decisions);
    // [2] the live list's length (the movers' flag scan runs over it); [3] movers (that
    // scan's total); [4] nv
    DBuf<unsigned long long> status(ctx, 5);
    DBuf<int32_t> live(ctx, (size_t)std::max<int64_t>(nv, 1));
    ph.gid = d_gid;
    ph.gact = gact.p;
    ph.gconst = gconst.p;
    ph.hub_sort = false;  // the sweep is replayed from a captured graph: no library sort inside
    const bool seq = (flags & VT_FLAG_SEQUENTIAL_SUMS) != 0;

    ph.init_state(false);
    if (seq) seg_seq_sum2(ctx, seg, ph.vdeg.p, nullptr, gsums.p);
    else seg_tree_sum2(ctx, seg, ph.vdeg.p, nullptr, gsums.p);
    VT_LAUNCH(ctx, "seg_constant", k_seg_constant, grid_for(ng, 256, 64), 256, 0, (const double*)gsums.p, ng, gconst.p);
    h2d(ctx, gthr.p, thresh.data(), 8 * (size_t)ng);
    h2d(ctx, glow.p, lower.data(), 8 * (size_t)ng);
    d2d(ctx, gprev.p, glow.p, 8 * (size_t)ng);
    dzero(ctx, giters.p, 4 * (size_t)ng);
    {
        std::vector<unsigned char> ones((size_t)ng, 1);
        h2d(ctx, gact.p, ones.data(), (size_t)ng);
        const unsigned long long st0[5] = {(unsigned long long)ng, (unsigned long long)nv, (unsigned long long)nv, 0,
                                           (unsigned long long)nv};
        h2d(ctx, status.p, st0, sizeof(st0));
    }
    ph.build_lists(nullptr, 1, false);

    const int64_t nev = 2 * std::max<int64_t>(nv, 1);
    uint32_t* k0 = (uint32_t*)scratch(ctx, "bev_k0", 4 * (size_t)nev);
    uint32_t* v0 = (uint32_t*)scratch(ctx, "bev_v0", 4 * (size_t)nev);
    uint32_t* k1 = (uint32_t*)scratch(ctx, "bev_k1", 4 * (size_t)nev);
    uint32_t* v1 = (uint32_t*)scratch(ctx, "bev_v1", 4 * (size_t)nev);
    int32_t* mflags = (int32_t*)scratch(ctx, "bmv_flags", 4 * (size_t)std::max<int64_t>(nv, 1));
    int32_t* mpos = (int32_t*)scratch(ctx, "bmv_pos", 4 * (size_t)std::max<int64_t>(nv, 1));
    const uint32_t none = (uint32_t)nv;
    const int bits = (int)bits_for(std::max<int64_t>(nv, 1));
    int64_t* d_nlive = (int64_t*)(status.p + 2);
    int64_t* d_movers = (int64_t*)(status.p + 3);
    const int64_t* d_nv = (const int64_t*)(status.p + 4);
    // The live list from the graphs' flags as they stand once the queued work has run; its
    // length is read back (the rounds captured after it take it as their launch size).
    int64_t nlive = nv;
    auto build_live = [&]() {
        if (!nv) {
            nlive = 0;
            return;
        }
        VT_LAUNCH(ctx, "live_flags", k_live_flags, grid_for(nv, 256, 8192), 256, 0, nv, d_gid,
                  (const unsigned char*)gact.p, mflags);
        dev_scan_i32(ctx, mflags, mpos, d_nv, nv, d_nlive, "blive_part");
        VT_LAUNCH(ctx, "live_scatter", k_live_scatter, grid_for(nv, 256, 8192), 256, 0, nv, (const int32_t*)mflags,
                  (const int32_t*)mpos, live.p);
        d2h(ctx, &nlive, d_nlive, 8);
        sync(ctx);
    };
    build_live();
    ph.narrow_once = true;  // k_batch_rotate keeps curr32 equal to curr from here on
    ph.pack_live = live.p;
    ph.pack_n = nlive;

    // One round of the reference loop body (louvain.cpp:498-577) for every graph still
    // iterating, with no host read: the sweep; the movers' flags and their scan (whose total
    // is the mover count, left on the device); two events per mover, sorted stably by
    // community with the device-count sort; the per-community delta folds; the per-graph
    // sums, decisions and rotation.  A graph that has left the phase is skipped by every
    // step, so a round after the last graph has left changes nothing.
    auto round = [&]() {
        ph.sweep(ph.sweeps[0]);
        if (nlive) {
            VT_LAUNCH(ctx, "batch_flags", k_batch_flags, grid_for(nlive, 256, 8192), 256, 0, nlive,
                      (const int32_t*)live.p, (const int64_t*)ph.target.p, (const int64_t*)ph.curr.p, d_gid,
                      (const unsigned char*)gact.p, mflags);
            dev_scan_i32(ctx, mflags, mpos, d_nlive, nlive, d_movers, "bmv_part");
            VT_LAUNCH(ctx, "emit_events", k_emit_events, grid_for(nlive, 256, 8192), 256, 0, (const int32_t*)live.p,
                      nlive, (const int32_t*)mflags, (const int32_t*)mpos, (const int64_t*)ph.target.p,
                      (const int64_t*)ph.curr.p, k0, v0, ph.ucount.p + 1);
        } else {
            dzero(ctx, d_movers, 8);  // every graph has left: no events
        }
        const bool alt = dev_sort_pairs_u32(ctx, k0, v0, k1, v1, d_movers, 2, nev, bits);
        ph.fold_events(alt ? k1 : k0, alt ? v1 : v0, 2 * nv, none, ph.cdeg.p, ph.csize.p, d_movers);
        if (seq) seg_seq_sum2(ctx, seg, ph.cw.p, ph.cdeg.p, gsums.p, (const unsigned char*)gact.p);
        else seg_tree_sum2(ctx, seg, ph.cw.p, ph.cdeg.p, gsums.p, (const unsigned char*)gact.p);
        dzero(ctx, status.p, 16);
        VT_LAUNCH(ctx, "batch_decide", k_batch_decide, grid_for(ng, 256, 64), 256, 0, ng, (const double*)gsums.p,
                  (const double*)gconst.p, (const double*)gthr.p, (const double*)glow.p, gact.p, grot.p, gprev.p,
                  giters.p, 0, (const int64_t*)seg.voff.p, status.p);
        if (nlive)
            VT_LAUNCH(ctx, "batch_rotate", k_batch_rotate, grid_for(nlive, 256, 8192), 256, 0, nlive,
                      (const int32_t*)live.p, d_gid, (const unsigned char*)grot.p, ph.past.p, ph.curr.p,
                      (const int64_t*)ph.target.p, ph.curr32.p);
    };

    hipGraphExec_t gexec = nullptr;
    auto capture = [&]() {
        const bool prof = ctx->profiling;
        ctx->profiling = false;  // no timing events inside a capture
        hipGraph_t graph = nullptr;
        VT_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        try {
            round();
        } catch (...) {
            (void)hipStreamEndCapture(ctx->stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            ctx->profiling = prof;
            throw;
        }
        VT_HIP(ctx, hipStreamEndCapture(ctx->stream, &graph));
        ctx->profiling = prof;
        const hipError_t e = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        VT_HIP(ctx, e);
    };
    // The host reads each round's status from pinned memory one round late: round r is
    // already queued when the host waits for round r-1's copy, so the device never waits for
    // the host.  The round after the last graph has left runs as a no-op.
    unsigned long long* pin = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    auto release = [&]() {
        (void)hipStreamSynchronize(ctx->stream);  // the last queued round has finished
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (pin) (void)hipHostFree(pin);
        pin = nullptr;
    };
    int rounds = 0, relists = 0;
    try {
        VT_HIP(ctx, hipHostMalloc((void**)&pin, 4 * sizeof(unsigned long long), hipHostMallocDefault));
        for (auto& e : ev) VT_HIP(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        int64_t listed = nv;
        for (int r = 0;; r++) {
            // A round without a captured graph (round 0, which allocates the scratch the
            // captured rounds reuse, and the round after a relist) runs directly, and the
            // next round is captured while it runs on the device.
            const bool direct = !gexec;
            if (gexec) VT_HIP(ctx, hipGraphLaunch(gexec, ctx->stream));
            else round();
            unsigned long long* slot = pin + 2 * (r & 1);
            VT_HIP(ctx, hipMemcpyAsync(slot, status.p, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                       ctx->stream));
            VT_HIP(ctx, hipEventRecord(ev[r & 1], ctx->stream));
            if (use_graph && direct) capture();
            if (r == 0) continue;
            VT_HIP(ctx, hipEventSynchronize(ev[(r - 1) & 1]));
            const unsigned long long* prev = pin + 2 * ((r - 1) & 1);
            const int64_t active = (int64_t)prev[0], live_nv = (int64_t)prev[1];
            if (active == 0) {
                rounds = r;  // rounds 0 .. r-1 had graphs iterating; round r was a no-op
                break;
            }
            if (2 * live_nv < listed) {
                // drop the finished graphs' vertices from the sweep lists (rebuilt after the
                // queued round, from the flags as they stand then); the round is captured anew
                listed = live_nv;
                relists++;
                ph.build_lists(nullptr, 1, false);
                build_live();  // ends with a sync: the queued round, which replayed the old capture, has finished
                ph.pack_n = nlive;
                if (gexec) (void)hipGraphExecDestroy(gexec);
                gexec = nullptr;
            }
        }
    } catch (...) {
        release();
        throw;
    }
    release();
    if (!u->part) u->part = (int64_t*)dalloc(ctx, 8 * (size_t)std::max<int64_t>(nv, 1));
    d2d(ctx, u->part, ph.past.p, 8 * (size_t)nv);
    q_out.assign((size_t)ng, 0.0);
    iters_out.assign((size_t)ng, 0);
    d2h(ctx, q_out.data(), gprev.p, 8 * (size_t)ng);
    d2h(ctx, iters_out.data(), giters.p, 4 * (size_t)ng);
    if (rounds_out) *rounds_out = rounds;
    if (relists_out) *relists_out = relists;
}

vt_tuning default_tuning()
{
    vt_tuning t;
    for (int b = 0; b < 6; b++) t.bucket_max_degree[b] = kBucketCap[b];
    // rows of degree 512-1023 go to the hub path: on the coarse levels, where thousands of
    // rows are that long, a workgroup each sorting and moving them beats one wave each with a
    // 2048-slot table (time to final Q on RGG n = 2^24: 0.1045-0.1049 s against 0.1059-0.1062)
    t.bucket_max_degree[5] = 511;
    // the largest cap measured best over whole threshold-cycling runs (time to final Q on RGG
    // n = 2^24: 0.1035-0.1038 s against 0.1045-0.1047 s at 16384 and 0.106 s at 2048)
    t.blocks_per_bucket = 65536;
    t.wide_table_max_degree[0] = 255;  // tables at most half full
    t.wide_table_max_degree[1] = 511;
    t.vertex_records = 0;
    t.hub_block_min = 1024;
    return t;
}

}  // namespace vt

using namespace vt;

extern "C" {

int vt_louvain_phase(vt_ctx* ctx, vt_graph* g, const vt_phase_params* p, int64_t* cvect_out, int32_t* iters_out,
                     double* q_out)
{
    if (!ctx || !g || !p) return VT_ERR_ARG;
    return guarded(ctx, [&] {
        VT_HIP(ctx, hipSetDevice(ctx->device));
        int32_t it = 0;
        const double q = run_phase(ctx, g, p, cvect_out, &it);
        if (iters_out) *iters_out = it;
        if (q_out) *q_out = q;
    });
}

int vt_phase_begin(vt_ctx* ctx, vt_graph* g, const vt_phase_params* p, vt_phase** out)
{
    if (!ctx || !g || !p || !out) return VT_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        VT_HIP(ctx, hipSetDevice(ctx->device));
        *out = phase_begin(ctx, g, p);
        sync(ctx);
    });
}

int vt_phase_step(vt_ctx* ctx, vt_phase* st, double* q_out, int32_t* done_out)
{
    if (!ctx || !st) return VT_ERR_ARG;
    return guarded(ctx, [&] {
        VT_HIP(ctx, hipSetDevice(ctx->device));
        phase_step(ctx, st);
        sync(ctx);
        if (q_out) *q_out = st->currMod;
        if (done_out) *done_out = st->done ? 1 : 0;
    });
}

int vt_phase_sweep(vt_ctx* ctx, vt_phase* st)
{
    if (!ctx || !st) return VT_ERR_ARG;
    return guarded(ctx, [&] {
        VT_HIP(ctx, hipSetDevice(ctx->device));
        for (const auto& cs : st->ph.sweeps) st->ph.sweep(cs);
        sync(ctx);
    });
}

static void check_tuning(vt_ctx* ctx, const vt_tuning* t)
{
    // a row in bucket b has degree <= bound[b]; the kernel of bucket b handles up to cap[b]
    const int32_t* cap = kBucketCap;
    for (int b = 0; b < 6; b++)
        if (t->bucket_max_degree[b] < -1 || t->bucket_max_degree[b] > cap[b])
            fail(ctx, VT_ERR_ARG, "vt_tuning: bucket_max_degree[%d] = %d outside [-1, %d]", b, t->bucket_max_degree[b],
                 cap[b]);
    if (t->blocks_per_bucket < 8 || t->blocks_per_bucket > (1 << 20))
        fail(ctx, VT_ERR_ARG, "vt_tuning: blocks_per_bucket = %d outside [8, 2^20]", t->blocks_per_bucket);
    // a row has at most degree + 1 distinct communities; keep each table at most 3/4 full
    const int32_t w0 = t->wide_table_max_degree[0], w1 = t->wide_table_max_degree[1];
    if (w0 < -1 || w0 > 3 * 512 / 4 - 1 || w1 < -1 || w1 > 3 * 1024 / 4 - 1 || w0 > w1)
        fail(ctx, VT_ERR_ARG, "vt_tuning: wide_table_max_degree = {%d, %d}: need -1 <= [0] <= [1], [0] <= 383, [1] <= 767",
             w0, w1);
    if (t->vertex_records != 0 && t->vertex_records != 1)
        fail(ctx, VT_ERR_ARG, "vt_tuning: vertex_records = %d, need 0 or 1", t->vertex_records);
    if (t->hub_block_min < 0 || t->hub_block_min > (1 << 30))
        fail(ctx, VT_ERR_ARG, "vt_tuning: hub_block_min = %d outside [0, 2^30]", t->hub_block_min);
}

int vt_ctx_get_tuning(const vt_ctx* ctx, vt_tuning* out)
{
    if (!ctx || !out) return VT_ERR_ARG;
    *out = ctx->tuning;
    return VT_OK;
}

int vt_ctx_set_tuning(vt_ctx* ctx, const vt_tuning* t)
{
    if (!ctx || !t) return VT_ERR_ARG;
    return guarded(ctx, [&] {
        check_tuning(ctx, t);
        ctx->tuning = *t;
    });
}

// Coordinate descent over the community-record layout, two whole bound profiles, the bucket
// bounds (one chunk or up to two per row at each group size), the per-bucket block cap (then the bounds again when the
// cap changed) and the wide-row table sizes,
// each candidate timed with HIP events over `reps` sweeps of the same state; a candidate
// must win by 1%, timed twice against the incumbent timed twice, to replace it.
int vt_autotune(vt_ctx* ctx, vt_graph* g, int32_t prep, int32_t reps, vt_tuning* out)
{
    if (!ctx || !g || prep < 0 || reps < 1) return VT_ERR_ARG;
    return guarded(ctx, [&] {
        VT_HIP(ctx, hipSetDevice(ctx->device));
        vt_phase_params p;
        std::memset(&p, 0, sizeof(p));
        p.lower = -1.0;
        p.thresh = 1.0E-3;
        const vt_tuning start = ctx->tuning;
        vt_phase* st = phase_begin(ctx, g, &p);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        try {
            for (int i = 0; i < prep && !st->done; i++) phase_step(ctx, st);
            VT_HIP(ctx, hipEventCreate(&e0));
            VT_HIP(ctx, hipEventCreate(&e1));
            const bool prof = ctx->profiling;
            ctx->profiling = false;
            auto time_it = [&](const vt_tuning& t) {
                ctx->tuning = t;
                st->ph.build_lists(nullptr, 1, false);
                st->ph.sweep(st->ph.sweeps[0]);
                VT_HIP(ctx, hipEventRecord(e0, ctx->stream));
                for (int r = 0; r < reps; r++) st->ph.sweep(st->ph.sweeps[0]);
                VT_HIP(ctx, hipEventRecord(e1, ctx->stream));
                VT_HIP(ctx, hipEventSynchronize(e1));
                float ms = 0.f;
                VT_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
                return (double)ms / reps;
            };
            vt_tuning best = start;
            double tb = time_it(best);
            // a candidate that times 1% faster is timed again together with the incumbent,
            // and replaces it only if its mean still wins (one timing can win by noise)
            // VITE_AMD_AUTOTUNE_LOG=1 prints every candidate's timing to stderr
            const bool log = std::getenv("VITE_AMD_AUTOTUNE_LOG") != nullptr;
            auto show = [&](const char* what, const vt_tuning& t, double ms) {
                if (!log) return;
                std::fprintf(stderr, "autotune %s: bounds {%d,%d,%d,%d,%d,%d} blocks %d wide {%d,%d} records %d: %.4f ms\n",
                             what, t.bucket_max_degree[0], t.bucket_max_degree[1], t.bucket_max_degree[2],
                             t.bucket_max_degree[3], t.bucket_max_degree[4], t.bucket_max_degree[5],
                             t.blocks_per_bucket, t.wide_table_max_degree[0], t.wide_table_max_degree[1],
                             t.vertex_records, ms);
            };
            show("start", best, tb);
            auto consider = [&](const vt_tuning& t) {
                const double tt = time_it(t);
                show("candidate", t, tt);
                if (tt >= 0.99 * tb) return;
                const double mt = 0.5 * (tt + time_it(t));
                const double mb = 0.5 * (tb + time_it(best));
                if (mt < 0.99 * mb) {
                    best = t;
                    tb = mt;
                    show("accepted", t, mt);
                } else {
                    tb = mb;
                }
            };
            // community records against separate arrays: first, since it changes which
            // bucket bounds and caps are best, and again at the end
            auto try_records = [&]() {
                vt_tuning t = best;
                t.vertex_records = 1 - best.vertex_records;
                consider(t);
            };
            try_records();
            // whole bound profiles first, which single-bucket moves reach only through
            // intermediate settings that are slower on their own: the widest group that a
            // row fills in one chunk (random graphs, where every row's reads are one round
            // trip whatever the group), and one chunk per row at each group size
            // each with and without the community records: on a skewed graph the two gain
            // only together (BA n = 2^22: 2.29 ms against 2.7 for either alone)
            auto try_profiles = [&]() {
                const vt_tuning before = best;
                for (const std::array<int32_t, 6>& prof : {std::array<int32_t, 6>{7, 7, 15, 31, 191, 1023},
                                                            std::array<int32_t, 6>{3, 7, 15, 31, 63, 1023}}) {
                    for (int rec = 0; rec < 2; rec++) {
                        vt_tuning t = best;
                        for (int b = 0; b < 6; b++) t.bucket_max_degree[b] = prof[b];
                        t.vertex_records = rec;
                        consider(t);
                    }
                }
                return std::memcmp(before.bucket_max_degree, best.bucket_max_degree, sizeof(before.bucket_max_degree)) != 0;
            };
            try_profiles();
            auto try_bounds = [&]() {
                for (int b = 0; b < 5; b++) {
                    const int G = 4 << b;
                    std::vector<int> cands;
                    for (int h = 2; h <= 2 * kNch[b]; h++) cands.push_back(h * G / 2 - 1);  // 1, 1.5, 2, ... chunks
                    for (int cand : cands) {
                        if (cand == best.bucket_max_degree[b]) continue;
                        vt_tuning t = best;
                        t.bucket_max_degree[b] = cand;
                        consider(t);
                    }
                }
            };
            try_bounds();
            auto try_caps = [&]() {
                const int32_t before = best.blocks_per_bucket;
                for (int cap : {2048, 4096, 8192, 16384, 65536}) {
                    if (cap == best.blocks_per_bucket) continue;
                    vt_tuning t = best;
                    t.blocks_per_bucket = cap;
                    consider(t);
                }
                return best.blocks_per_bucket != before;
            };
            // a new cap moves the best bounds: the profiles and one more pass over the bounds
            if (try_caps()) {
                try_profiles();
                try_bounds();
            }
            // wide-row table sizes: fuller small tables (more waves per CU) against emptier
            // large ones (shorter probes), when the graph has wide rows at all
            const ColorSweep& cs0 = st->ph.sweeps[0];
            if (cs0.wide_n[0] + cs0.wide_n[1] + cs0.wide_n[2] > 0) {
                for (int k = 0; k < 2; k++) {
                    const std::vector<int> cands = k == 0 ? std::vector<int>{-1, 255, 319, 383}
                                                          : std::vector<int>{511, 639, 767};
                    for (int cand : cands) {
                        if (cand == best.wide_table_max_degree[k]) continue;
                        vt_tuning t = best;
                        t.wide_table_max_degree[k] = cand;
                        if (t.wide_table_max_degree[0] > t.wide_table_max_degree[1]) continue;
                        consider(t);
                    }
                }
            }
            // the profiles once more against the final choice (the cap and tables they are
            // judged with; a profile that wins can want another cap), then the record layout
            if (try_profiles()) try_caps();
            try_records();
            ctx->profiling = prof;
            ctx->tuning = best;
        } catch (...) {
            ctx->tuning = start;
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            delete st;
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        delete st;
        sync(ctx);
        if (out) *out = ctx->tuning;
    });
}

int vt_phase_get_state(vt_ctx* ctx, const vt_phase* st, int64_t* curr, int64_t* csize, double* cdeg, double* vdeg,
                       double* constant, int64_t* target, double* cw)
{
    if (!ctx || !st) return VT_ERR_ARG;
    return guarded(ctx, [&] {
        VT_HIP(ctx, hipSetDevice(ctx->device));
        const Phase& ph = st->ph;
        const size_t b8 = 8 * (size_t)ph.nv;
        if (curr) d2h(ctx, curr, ph.curr.p, b8);
        if (csize) d2h(ctx, csize, ph.csize.p, b8);
        if (cdeg) d2h(ctx, cdeg, ph.cdeg.p, b8);
        if (vdeg) d2h(ctx, vdeg, ph.vdeg.p, b8);
        if (target) d2h(ctx, target, ph.target.p, b8);
        if (cw) d2h(ctx, cw, ph.cw.p, b8);
        if (constant) d2h(ctx, constant, ph.scal.p + 2, 8);
    });
}

int vt_phase_end(vt_ctx* ctx, vt_phase* st, int64_t* cvect_out, int32_t* iters_out, double* q_out)
{
    if (!ctx || !st) return VT_ERR_ARG;
    const int rc = guarded(ctx, [&] {
        VT_HIP(ctx, hipSetDevice(ctx->device));
        const double q = phase_end(ctx, st, cvect_out, iters_out);
        if (q_out) *q_out = q;
    });
    delete st;
    return rc;
}

int vt_last_phase_trace(const vt_ctx* ctx, int32_t cap, double* q, int64_t* moves, int32_t* n)
{
    if (!ctx) return VT_ERR_ARG;
    const int32_t m = (int32_t)std::min<size_t>(ctx->trace_q.size(), (size_t)std::max(cap, 0));
    for (int32_t i = 0; i < m; i++) {
        if (q) q[i] = ctx->trace_q[i];
        if (moves) moves[i] = ctx->trace_moves[i];
    }
    if (n) *n = (int32_t)ctx->trace_q.size();
    return VT_OK;
}

int vt_init_degrees(vt_ctx* ctx, const vt_graph* g, int32_t flags, double* vdeg_out, double* constant_out)
{
    if (!ctx || !g) return VT_ERR_ARG;
    return guarded(ctx, [&] {
        VT_HIP(ctx, hipSetDevice(ctx->device));
        Phase ph(ctx, g);
        ph.alloc(false);
        ph.init(flags);
        if (vdeg_out) d2h(ctx, vdeg_out, ph.vdeg.p, 8 * (size_t)g->nv);
        if (constant_out) d2h(ctx, constant_out, ph.constant(), 8);
        sync(ctx);
    });
}

int vt_sweep(vt_ctx* ctx, const vt_graph* g, const int64_t* curr, const int64_t* csize, const double* cdeg,
             const double* vdeg, double constant, int64_t* target_out, double* cw_out, int64_t* dsize_out,
             double* ddeg_out)
{
    if (!ctx || !g || !curr || !csize || !cdeg || !vdeg) return VT_ERR_ARG;
    return guarded(ctx, [&] {
        VT_HIP(ctx, hipSetDevice(ctx->device));
        const int64_t nv = g->nv;
        for (int64_t i = 0; i < nv; i++)
            if (curr[i] < 0 || curr[i] >= nv) fail(ctx, VT_ERR_ARG, "vt_sweep: community of vertex %lld out of range", (long long)i);
        Phase ph(ctx, g);
        ph.alloc(false);
        const size_t b8 = 8 * (size_t)nv;
        h2d(ctx, ph.curr.p, curr, b8);
        h2d(ctx, ph.csize.p, csize, b8);
        h2d(ctx, ph.cdeg.p, cdeg, b8);
        h2d(ctx, ph.vdeg.p, vdeg, b8);
        h2d(ctx, ph.constant(), &constant, 8);
        ph.build_lists(nullptr, 1, false);
        ph.sweep(ph.sweeps[0]);
        if (target_out) d2h(ctx, target_out, ph.target.p, b8);
        if (cw_out) d2h(ctx, cw_out, ph.cw.p, b8);
        if (dsize_out || ddeg_out) {
            DBuf<double> dd(ctx, (size_t)std::max<int64_t>(nv, 1));
            DBuf<int64_t> ds(ctx, (size_t)std::max<int64_t>(nv, 1));
            dzero(ctx, dd.p, b8);
            dzero(ctx, ds.p, b8);
            ph.apply_moves(nullptr, nv, dd.p, ds.p);
            if (ddeg_out) d2h(ctx, ddeg_out, dd.p, b8);
            if (dsize_out) d2h(ctx, dsize_out, ds.p, b8);
        }
        sync(ctx);
    });
}

int vt_modularity(vt_ctx* ctx, int64_t nv, const double* cw, const double* cdeg, double constant, int32_t flags,
                  double* q_out)
{
    if (!ctx || !cw || !cdeg || nv < 0) return VT_ERR_ARG;
    return guarded(ctx, [&] {
        VT_HIP(ctx, hipSetDevice(ctx->device));
        DBuf<double> a(ctx, (size_t)std::max<int64_t>(nv, 1)), b(ctx, (size_t)std::max<int64_t>(nv, 1)), s(ctx, 8);
        h2d(ctx, a.p, cw, 8 * (size_t)nv);
        h2d(ctx, b.p, cdeg, 8 * (size_t)nv);
        h2d(ctx, s.p + 2, &constant, 8);
        if (flags & VT_FLAG_SEQUENTIAL_SUMS) seq_sum2(ctx, a.p, b.p, nv, s.p);
        else tree_sum2(ctx, a.p, b.p, nv, s.p);
        VT_LAUNCH(ctx, "modularity", k_modularity, 1, 64, 0, (const double*)s.p, (const double*)(s.p + 2), s.p + 3,
                  (const int64_t*)nullptr, (int64_t*)nullptr);
        double q = 0;
        d2h(ctx, &q, s.p + 3, 8);
        if (q_out) *q_out = q;
        sync(ctx);
    });
}

}  // extern "C"
