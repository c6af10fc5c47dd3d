// Traversal kernels: closest-hit rays, closest / kNN / radius point queries, over
// binary (32 B) or 4-wide (128 B) nodes, for triangle, point and Gaussian leaves.
//
// One lane per ray or query, depth-first.  The pruning predicates are the
// reference's (src/gpu/raycast.cu:75-245, src/gpu/query.cu:52-637): padded slab
// test against the current best t for rays, box lower bound <= bound for
// closest / kNN (so equal-distance primitives with a smaller index are still
// reached), box lower bound <= r^2 for radius.  Every predicate only prunes
// subtrees that cannot hold a better (dist2, index) or a smaller t, so the result
// is the brute-force answer whatever the tree, node width or visiting order -
// which is what makes binary, wide, LBVH, SAH and reordered runs bit-identical.
//
// Layout for gfx950:
//   * the per-lane DFS stack lives in LDS ([depth][lane], 32 entries = 32 KB per
//     256-lane workgroup); deeper entries spill to scratch, which only very deep
//     or very wide traversals touch;
//   * a node's children are ordered with a fixed compare-exchange network in
//     registers (no dynamically indexed private arrays);
//   * leaf children are tested as soon as their parent is opened, nearest first,
//     against the live bound, and only interior children go on the stack;
//   * closest is kNN with k = 1 (identical record and tie rule).
//
// Queries may be visited in a permuted order (Morton reordering): lane i works
// on query order[i] and writes its result straight to slot order[i], so no
// gather/scatter copies of queries or results are made.
#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include "qr_traverse.h"

namespace qr {

namespace {

constexpr int kBlock = 256;
constexpr int kLdsDepth = 32;
constexpr int kSpill = 112;  // scratch entries beyond the LDS part (never reached by LBVH trees)

// ---- XCD-aware block numbering ------------------------------------------------------
// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.  The
// traversal kernels renumber their blocks so that XCD x walks one contiguous range
// of the (Morton-ordered) queries: neighbouring queries share most of their nodes,
// and those nodes then stay in one L2 instead of being fetched by all eight.
// Bijective on [0, gridDim.x): XCD x owns q + (x < r) blocks, q = g / 8, r = g % 8.
__device__ __forceinline__ int xcdBlock() {
    const int g = static_cast<int>(gridDim.x), i = static_cast<int>(blockIdx.x);
    const int x = i & 7, j = i >> 3;
    const int q = g >> 3, r = g & 7;
    return x * q + (x < r ? x : r) + j;
}

// Largest v over the wave, one atomic per wave (nothing when dst is null).
__device__ __forceinline__ void waveMaxAtomic(int* dst, int v) {
    if (dst == nullptr) return;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    if ((threadIdx.x & 63) == 0 && v > 0) atomicMax(dst + blockIdx.x % kMaxCopies, v);
}

// ---- stats -------------------------------------------------------------------------
// stats[0] node tests, [1] leaf tests, [2] stack overflows, [3] interior nodes opened.
// Pinned host word that any walk whose stack overflowed sets to 1 (the walk stops
// and its result is incomplete); the host checks it after every query call.  Null
// until armOverflowWord() has run on the device.
__device__ int* gOverflowWord = nullptr;

template <bool kStats>
__device__ __forceinline__ void flushStats(unsigned long long* stats, int nodeTests, int leafTests,
                                           int overflow, int opened) {
    if (overflow) {  // an incomplete walk: tell the host, whatever the build
        int* w = gOverflowWord;
        if (w) __hip_atomic_store(w, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    if (!kStats) return;
    unsigned long long a = static_cast<unsigned long long>(nodeTests);
    unsigned long long b = static_cast<unsigned long long>(leafTests);
    unsigned long long c = static_cast<unsigned long long>(overflow);
    unsigned long long d = static_cast<unsigned long long>(opened);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
        c += __shfl_xor(c, off, 64);
        d += __shfl_xor(d, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        unsigned long long* st = stats + 4 * (blockIdx.x % kStatCopies);
        if (a) atomicAdd(&st[0], a);
        if (b) atomicAdd(&st[1], b);
        if (c) atomicAdd(&st[2], c);
        if (d) atomicAdd(&st[3], d);
    }
}

// ---- per-lane stack: LDS first, scratch beyond ---------------------------------------
// The stack pointer and LDS column are plain registers; only the spill area (the
// part beyond the LDS depth) lives in scratch.  Both areas are accessed through
// volatile pointers: that keeps every LDS access a ds_read/ds_write (which one
// wave executes in order) and stops the compiler from merging the two branches
// of a pop into one generic (flat) load, which is not ordered behind the
// preceding ds_write pushes.
typedef __attribute__((address_space(3))) volatile int LdsInt;      // workgroup LDS
typedef __attribute__((address_space(5))) volatile int PrivateInt;  // per-lane scratch

template <int kS, int kD>
struct LaneStackT {
    static constexpr int kStride = kS;  // lanes per workgroup: distance between a lane's entries
    static constexpr int kDepth = kD;   // entries kept in LDS before spilling to scratch
    LdsInt* base;  // this lane's column of the workgroup's LDS stack
    int sp;
    __device__ __forceinline__ void init(int* lds) {
        base = (LdsInt*)(lds + threadIdx.x);
        sp = 0;
    }
    __device__ __forceinline__ bool empty() const { return sp == 0; }
    __device__ __forceinline__ bool room(int n) const { return sp + n <= kD + kSpill; }
};
using LaneStack = LaneStackT<kBlock, kLdsDepth>;

struct SpillArea {
    int v[kSpill];
    __device__ __forceinline__ PrivateInt* at(int i) {
        return (PrivateInt*)(&v[i]);
    }
};

#define QR_PUSH(st, spill, val)                               \
    do {                                                      \
        const int qrv_ = (val);                               \
        if ((st).sp < (st).kDepth) {                          \
            (st).base[(st).sp * (st).kStride] = qrv_;         \
        } else {                                              \
            *(spill).at((st).sp - (st).kDepth) = qrv_;        \
        }                                                     \
        ++(st).sp;                                            \
    } while (0)

#define QR_POP(st, spill, out)                                \
    do {                                                      \
        --(st).sp;                                            \
        if ((st).sp < (st).kDepth) {                          \
            (out) = (st).base[(st).sp * (st).kStride];        \
        } else {                                              \
            (out) = *(spill).at((st).sp - (st).kDepth);       \
        }                                                     \
    } while (0)

// Stack operations shared by the walks: the LDS/scratch stack above, and a stack
// whose top R entries live in registers (a shift register: static indices only, so
// it stays in VGPRs) with deeper entries in the lane's scratch spill area.
template <typename Stack>
__device__ __forceinline__ void stackPush(Stack& st, SpillArea& spill, int v) {
    QR_PUSH(st, spill, v);
}
template <typename Stack>
__device__ __forceinline__ int stackPop(Stack& st, SpillArea& spill) {
    int v;
    QR_POP(st, spill, v);
    return v;
}

template <int R>
struct RegStack {
    int r[R];  // r[0]: top
    int sp;    // entries; those beyond R sit in the spill area, deepest first
    __device__ __forceinline__ void init(int*) { sp = 0; }
    __device__ __forceinline__ bool empty() const { return sp == 0; }
    __device__ __forceinline__ bool room(int n) const { return sp + n <= R + kSpill; }
};
template <int R>
__device__ __forceinline__ void stackPush(RegStack<R>& st, SpillArea& spill, int v) {
    if (st.sp >= R) *spill.at(st.sp - R) = st.r[R - 1];
#pragma unroll
    for (int k = R - 1; k > 0; --k) st.r[k] = st.r[k - 1];
    st.r[0] = v;
    ++st.sp;
}
template <int R>
__device__ __forceinline__ int stackPop(RegStack<R>& st, SpillArea& spill) {
    const int v = st.r[0];
#pragma unroll
    for (int k = 0; k < R - 1; ++k) st.r[k] = st.r[k + 1];
    --st.sp;
    if (st.sp >= R) st.r[R - 1] = *spill.at(st.sp - R);
    return v;
}
constexpr int kRegStack = 8;

// Compare-exchange on (key, payload) pairs: ascending by key.
__device__ __forceinline__ void cx(float& ka, int& va, float& kb, int& vb) {
    const bool sw = kb < ka;
    const float tk = ka;
    const int tv = va;
    ka = sw ? kb : ka;
    va = sw ? vb : va;
    kb = sw ? tk : kb;
    vb = sw ? tv : vb;
}

// Sorts four (key, child) pairs ascending (5 compare-exchanges).
__device__ __forceinline__ void sort4(float (&k)[4], int (&c)[4]) {
    cx(k[0], c[0], k[1], c[1]);
    cx(k[2], c[2], k[3], c[3]);
    cx(k[0], c[0], k[2], c[2]);
    cx(k[1], c[1], k[3], c[3]);
    cx(k[1], c[1], k[2], c[2]);
}

__device__ __forceinline__ bool closer(float d, int id, float bd, int bid) {
    return d < bd || (d == bd && id < bid);
}

// ---- k best so far ------------------------------------------------------------------
// Capacity-K list sorted by (dist2, index) in registers, for any k <= K.  It keeps
// the K best seen so far and prunes with the K-th: the first k entries are then
// the exact k best (a subtree beyond the K-th best cannot hold one of the k best),
// and every index into the list is static, so it never leaves registers.  A
// candidate is inserted by one ascending compare-exchange sweep; the displaced
// worst drops off the end.  Semantics equal the reference's sorted table
// (src/gpu/query.cu:150-158) and, for k = 1, its closest-point rule
// (src/gpu/query.cu:83-90).
template <int K>
struct KnnList {
    float d[K];
    int id[K];
    int n;
    __device__ __forceinline__ void init() {
        n = 0;
#pragma unroll
        for (int s = 0; s < K; ++s) {
            d[s] = kInf;
            id[s] = 0x7fffffff;
        }
    }
    __device__ __forceinline__ bool full() const { return n >= K; }
    // Pruning test for a subtree or primitive whose lower bound is lb.
    __device__ __forceinline__ bool admits(float lb) const { return !full() || lb <= d[K - 1]; }
    __device__ __forceinline__ void offer(float nd, int nid) {
        if (full() && !closer(nd, nid, d[K - 1], id[K - 1])) return;
        float cd = nd;
        int cid = nid;
#pragma unroll
        for (int s = 0; s < K; ++s) {
            const bool sw = closer(cd, cid, d[s], id[s]);
            const float td = d[s];
            const int tid = id[s];
            d[s] = sw ? cd : td;
            id[s] = sw ? cid : tid;
            cd = sw ? td : cd;
            cid = sw ? tid : cid;
        }
        n = n < K ? n + 1 : K;
    }
};

// ---- rays --------------------------------------------------------------------------
// Each iteration opens one interior node: its children are tested, leaves are
// intersected nearest first, and the nearest surviving interior child becomes the
// next node without a stack round trip; the other survivors are pushed far first.

// One ray-triangle leaf test.  With kSave the hit's scene index is kept too
// (equal t: the smaller scene index), for the differentiable mode; t is the same.
template <bool kSave>
__device__ __forceinline__ void rayLeaf(const Ray& ray, const Tri* __restrict__ tris,
                                        const int* __restrict__ primIds, int slot, float& best,
                                        bool& any, int& bestId) {
    float t;
    if (!rayTri(ray, tris[slot], t)) return;
    if (t < best) {
        best = t;
        any = true;
        if (kSave) bestId = primIds[slot];
    } else if (kSave && any && t == best) {
        const int sid = primIds[slot];
        if (sid < bestId) bestId = sid;
    }
}

template <bool kStats, bool kSave>
__global__ __launch_bounds__(kBlock) void rayBinaryKernel(const Node2* __restrict__ nodes,
                                                          const Tri* __restrict__ tris,
                                                          const Ray* __restrict__ rays,
                                                          const uint32_t* __restrict__ order,
                                                          int count, uint64_t* __restrict__ out,
                                                          unsigned long long* __restrict__ stats,
                                                          const int* __restrict__ primIds,
                                                          int* __restrict__ hitPrim) {
    __shared__ int lds[kLdsDepth * kBlock];
    const int i = xcdBlock() * kBlock + threadIdx.x;
    int nodeTests = 0, leafTests = 0, overflow = 0, opened = 0;
    if (i < count) {
        const int q = order ? static_cast<int>(order[i]) : i;
        const Ray ray = rays[q];
        float best = ray.tMax;
        bool any = false;
        int bestId = -1;
        LaneStack st;
        SpillArea spill;
        st.init(lds);
        float tEnter;
        const Node2 root = nodes[0];
        int id = -1;
        if (rayBox(ray, nodeBox(root), best, tEnter)) {
            ++nodeTests;
            id = 0;
        }
        while (id >= 0) {
            if (kStats) ++opened;
            const Node2 nd = nodes[id];
            const int kid[2] = {nd.left, nd.right};
            float t[2];
            int c[2];
            bool leaf[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const Node2 cn = nodes[kid[s]];
                leaf[s] = cn.left < 0;
                c[s] = leaf[s] ? -cn.left - 1 : kid[s];
                t[s] = kInf;
                if (rayBox(ray, nodeBox(cn), best, t[s])) {
                    ++nodeTests;
                } else {
                    c[s] = kWideEmpty;
                    t[s] = kInf;
                }
            }
            const bool swap = t[1] < t[0];
            const int c0 = swap ? c[1] : c[0], c1 = swap ? c[0] : c[1];
            const bool l0 = swap ? leaf[1] : leaf[0], l1 = swap ? leaf[0] : leaf[1];
            if (c0 != kWideEmpty && l0) {
                ++leafTests;
                rayLeaf<kSave>(ray, tris, primIds, c0, best, any, bestId);
            }
            if (c1 != kWideEmpty && l1) {
                ++leafTests;
                rayLeaf<kSave>(ray, tris, primIds, c1, best, any, bestId);
            }
            const bool in0 = c0 != kWideEmpty && !l0;
            const bool in1 = c1 != kWideEmpty && !l1;
            if (in0) {
                if (in1) {
                    if (!st.room(1)) {
                        overflow = 1;
                        break;
                    }
                    QR_PUSH(st, spill, c1);
                }
                id = c0;
            } else if (in1) {
                id = c1;
            } else if (!st.empty()) {
                QR_POP(st, spill, id);
            } else {
                id = -1;
            }
        }
        out[q] = any ? ((uint64_t{1} << 32) | floatBits(best)) : 0ull;
        if (kSave) hitPrim[q] = any ? bestId : -1;
    }
    flushStats<kStats>(stats, nodeTests, leafTests, overflow, opened);
}

__device__ __forceinline__ void loadWide(const Wide4* __restrict__ wide, int id, WideEntry (&e)[4]) {
    const Wide4& wn = wide[id];
#pragma unroll
    for (int s = 0; s < 4; ++s) e[s] = wn.c[s];
}

// After the four children are sorted: the first interior survivor continues the
// descent, the rest are pushed far first.  Returns the next node, or pops one
// (-1 when the traversal is complete).  ok[s] says whether slot s survives.
__device__ __forceinline__ int descend(const int (&c)[4], const bool (&ok)[4], LaneStack& st,
                                       SpillArea& spill, int& overflow) {
    int next = -1;
    int first = -1;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (ok[s] && c[s] >= 0 && first < 0) first = s;
    }
    if (!st.room(3)) {
        overflow = 1;
        return -1;
    }
#pragma unroll
    for (int s = 3; s >= 0; --s) {
        if (ok[s] && c[s] >= 0 && s != first) QR_PUSH(st, spill, c[s]);
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
        if (s == first) next = c[s];
    if (first < 0) {
        if (!st.empty()) {
            QR_POP(st, spill, next);
        }
    }
    return next;
}

template <bool kStats, bool kSave>
__global__ __launch_bounds__(kBlock) void rayWideKernel(const Wide4* __restrict__ wide,
                                                        const Tri* __restrict__ tris,
                                                        const Ray* __restrict__ rays,
                                                        const uint32_t* __restrict__ order,
                                                        int count, uint64_t* __restrict__ out,
                                                        unsigned long long* __restrict__ stats,
                                                        const int* __restrict__ primIds,
                                                        int* __restrict__ hitPrim) {
    __shared__ int lds[kLdsDepth * kBlock];
    const int i = xcdBlock() * kBlock + threadIdx.x;
    int nodeTests = 0, leafTests = 0, overflow = 0, opened = 0;
    if (i < count) {
        const int q = order ? static_cast<int>(order[i]) : i;
        const Ray ray = rays[q];
        float best = ray.tMax;
        bool any = false;
        int bestId = -1;
        LaneStack st;
        SpillArea spill;
        st.init(lds);
        int id = 0;
        while (id >= 0) {
            if (kStats) ++opened;
            WideEntry e[4];
            loadWide(wide, id, e);
            float key[4];
            int c[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                c[s] = e[s].child;
                key[s] = kInf;
                float t;
                if (c[s] != kWideEmpty && rayBox(ray, entryBox(e[s]), best, t)) {
                    key[s] = t;
                    ++nodeTests;
                } else {
                    c[s] = kWideEmpty;
                }
            }
            sort4(key, c);
            bool ok[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                ok[s] = c[s] != kWideEmpty;
                if (ok[s] && c[s] < 0) {
                    ++leafTests;
                    rayLeaf<kSave>(ray, tris, primIds, -c[s] - 1, best, any, bestId);
                }
            }
            id = descend(c, ok, st, spill, overflow);
        }
        out[q] = any ? ((uint64_t{1} << 32) | floatBits(best)) : 0ull;
        if (kSave) hitPrim[q] = any ? bestId : -1;
    }
    flushStats<kStats>(stats, nodeTests, leafTests, overflow, opened);
}

// ---- closest / kNN ------------------------------------------------------------------

template <typename L, int K>
__device__ __forceinline__ void writeList(const KnnList<K>& list, int k, ClosestResult* dst) {
#pragma unroll
    for (int s = 0; s < K; ++s) {
        if (s < k) {
            ClosestResult r{0u, -1};
            if (s < list.n) r = {floatBits(list.d[s]), list.id[s]};
            dst[s] = r;
        }
    }
}

// Stack entries of the closest / kNN walks carry the child's box distance above its id,
// truncated to the bits the id leaves free: the leading bits of a non-negative
// float are a lower bound of it.  A popped node whose bound already exceeds the
// k-th distance is dropped without being opened -- opening it would find every
// child beyond the k-th distance too (a child's box lies inside its parent's), so
// the walk and its results are unchanged.  idBits = 32 turns the keys off (trees
// whose wide ids need more than 23 bits).
// Bits of a stack entry taken by a node id below `idCount` (the rest hold the
// distance key); 32 = no key, when ids need more than 23 bits (fewer than 9 would
// be left, too few for the float's exponent and a mantissa bit).
__host__ __device__ __forceinline__ int stackIdBitsDev(int idCount) {
    int bits = 1;
    while (bits < 32 && (1LL << bits) < static_cast<long long>(idCount)) ++bits;
    return bits <= 23 ? bits : 32;
}

__device__ __forceinline__ int packEntry(int id, float key, int idBits) {
    if (idBits >= 32) return id;
    const uint32_t q = floatBits(key) >> (idBits - 1);
    return static_cast<int>((q << idBits) | static_cast<uint32_t>(id));
}
__device__ __forceinline__ int entryId(int v, int idBits) {
    return idBits >= 32 ? v : static_cast<int>(static_cast<uint32_t>(v) & ((1u << idBits) - 1u));
}
__device__ __forceinline__ float entryKey(int v, int idBits) {
    return idBits >= 32 ? 0.0f : bitsFloat((static_cast<uint32_t>(v) >> idBits) << (idBits - 1));
}

// Closest / kNN walk of one query over a binary tree whose node ids, leaf slots
// and prim ids are local to `nodes`, `prims` and `primIds` (a plain tree, or one
// tree of a batch addressed through offset pointers).
template <typename L, int K, bool kStats, typename Stack>
__device__ __forceinline__ void nearestBinaryWalk(const Node2* __restrict__ nodes,
                                                  const typename L::Item* __restrict__ prims,
                                                  const int* __restrict__ primIds, V3 p,
                                                  KnnList<K>& list, Stack& st, SpillArea& spill,
                                                  int& nodeTests, int& leafTests, int& overflow,
                                                  int& opened, int idBits) {
    int id = 0;
    while (id >= 0) {
        if (kStats) ++opened;
        const Node2 nd = nodes[id];
        ++nodeTests;
        const int kid[2] = {nd.left, nd.right};
        float key[2];
        int c[2];
        bool leaf[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const Node2 cn = nodes[kid[s]];
            leaf[s] = cn.left < 0;
            c[s] = leaf[s] ? -cn.left - 1 : kid[s];
            key[s] = boxDist2(p, nodeBox(cn));
        }
        const bool swap = key[1] < key[0];
        const int c0 = swap ? c[1] : c[0], c1 = swap ? c[0] : c[1];
        const bool l0 = swap ? leaf[1] : leaf[0], l1 = swap ? leaf[0] : leaf[1];
        const float k0 = swap ? key[1] : key[0], k1 = swap ? key[0] : key[1];
        if (l0 && list.admits(k0)) {
            ++leafTests;
            list.offer(L::dist2(p, prims[c0]), primIds[c0]);
        }
        if (l1 && list.admits(k1)) {
            ++leafTests;
            list.offer(L::dist2(p, prims[c1]), primIds[c1]);
        }
        const bool in0 = !l0 && list.admits(k0);
        const bool in1 = !l1 && list.admits(k1);
        if (in0) {
            if (in1) {
                if (!st.room(1)) {
                    overflow = 1;
                    break;
                }
                stackPush(st, spill, packEntry(c1, k1, idBits));
            }
            id = c0;
        } else if (in1) {
            id = c1;
        } else {
            id = -1;
            while (!st.empty()) {
                const int v = stackPop(st, spill);
                if (list.admits(entryKey(v, idBits))) {
                    id = entryId(v, idBits);
                    break;
                }
            }
        }
    }
}

template <typename L, int K, bool kStats>
__global__ __launch_bounds__(kBlock) void nearestBinaryKernel(
    const Node2* __restrict__ nodes, const typename L::Item* __restrict__ prims,
    const int* __restrict__ primIds, const V3* __restrict__ points,
    const uint32_t* __restrict__ order, int count, int k, int outStride,
    ClosestResult* __restrict__ out, unsigned long long* __restrict__ stats, int idBits) {
    __shared__ int lds[kLdsDepth * kBlock];
    const int i = xcdBlock() * kBlock + threadIdx.x;
    int nodeTests = 0, leafTests = 0, overflow = 0, opened = 0;
    if (i < count) {
        const int q = order ? static_cast<int>(order[i]) : i;
        KnnList<K> list;
        list.init();
        LaneStack st;
        SpillArea spill;
        st.init(lds);
        nearestBinaryWalk<L, K, kStats>(nodes, prims, primIds, points[q], list, st, spill, nodeTests,
                                        leafTests, overflow, opened, idBits);
        writeList<L, K>(list, k, out + static_cast<size_t>(q) * outStride);
    }
    flushStats<kStats>(stats, nodeTests, leafTests, overflow, opened);
}

// ---- batched scenes: persistent work-queue traversal ------------------------------------
// Queries of all scenes form one queue (scene-major, optionally Morton-ordered inside
// each scene).  Each wave takes the next 64 queries with one atomic and walks, per
// lane, the tree of that query's scene (scene s: items from primOff[s]; binary nodes
// from 2 primOff[s] - s, or 4-wide nodes from wideRoot[s]); the grid is sized to the
// resident occupancy and every wave leaves when the queue is exhausted.  Stack
// placement: kD entries per lane in LDS with the rest in scratch (kD = 0: all in
// scratch), or, with kReg, the top kRegStack entries in registers.
template <typename L, int K, int kBT, int kD, bool kWide, bool kReg>
__global__ __launch_bounds__(kBT) void batchNearestKernel(
    const Node2* __restrict__ nodes, const Wide4* __restrict__ wide, const int* __restrict__ wideRoot,
    const typename L::Item* __restrict__ prims, const int* __restrict__ primIds, const int* __restrict__ primOff,
    const V3* __restrict__ points, const int* __restrict__ qScene, const uint32_t* __restrict__ order, int count,
    int k, int outStride, ClosestResult* __restrict__ out, int* __restrict__ counter) {
    __shared__ int lds[(kD > 0 && !kReg ? kD : 1) * kBT];
    const int lane = threadIdx.x & 63;
    while (true) {
        int base = 0;
        if (lane == 0) base = atomicAdd(counter, 64);
        base = __shfl(base, 0, 64);
        if (base >= count) break;
        const int i = base + lane;
        if (i < count) {
            const int q = order ? static_cast<int>(order[i]) : i;
            const int s = qScene[q];
            const int ps = primOff[s];
            KnnList<K> list;
            list.init();
            typename std::conditional<kReg, RegStack<kRegStack>, LaneStackT<kBT, kD>>::type st;
            SpillArea spill;
            st.init(lds);
            int a = 0, b = 0, c = 0, d = 0;
            if (kWide) {
                const int wr = wideRoot[s];
                // wide ids of the scene's tree stay below its wide-node count
                nearestWideWalk<L, K, false>(wide + wr, prims + ps, primIds + ps, points[q], list, st, spill, a, b,
                                             c, d, stackIdBitsDev(wideRoot[s + 1] - wr));
            } else {
                // node ids of the scene's tree stay below 2 n_s - 1
                nearestBinaryWalk<L, K, false>(nodes + (2 * ps - s), prims + ps, primIds + ps, points[q], list, st,
                                               spill, a, b, c, d, stackIdBitsDev(2 * (primOff[s + 1] - ps) - 1));
            }
            if (c) {
                int* w = gOverflowWord;
                if (w) __hip_atomic_store(w, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
            writeList<L, K>(list, k, out + static_cast<size_t>(q) * outStride);
        }
    }
}

__global__ void batchResetKernel(int* __restrict__ counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] = 0;
}

// Closest / kNN walk of one query over a 4-wide tree whose node ids, leaf slots and
// prim ids are local to `wide`, `prims` and `primIds` (root at wide[0]).
template <typename L, int K, bool kStats, typename Stack>
__device__ __forceinline__ void nearestWideWalk(const Wide4* __restrict__ wide,
                                                const typename L::Item* __restrict__ prims,
                                                const int* __restrict__ primIds, V3 p, KnnList<K>& list,
                                                Stack& st, SpillArea& spill, int& nodeTests, int& leafTests,
                                                int& overflow, int& opened, int idBits) {
    int id = 0;
    while (id >= 0) {
        if (kStats) ++opened;
        WideEntry e[4];
        loadWide(wide, id, e);
        float key[4];
        int c[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            c[s] = e[s].child;
            key[s] = kInf;
            if (c[s] != kWideEmpty) {
                const float d = boxDist2(p, entryBox(e[s]));
                if (list.admits(d)) {
                    key[s] = d;
                    ++nodeTests;
                } else {
                    c[s] = kWideEmpty;
                }
            }
        }
        sort4(key, c);
        // leaves now, nearest first, against the live bound
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (c[s] != kWideEmpty && c[s] < 0 && list.admits(key[s])) {
                const int slot = -c[s] - 1;
                ++leafTests;
                list.offer(L::dist2(p, prims[slot]), primIds[slot]);
            }
        }
        // nearest surviving interior child next; the others pushed far first with
        // their distance keys; pops skip entries the k-th distance has overtaken
        int first = -1;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (c[s] >= 0 && list.admits(key[s]) && first < 0) first = s;
        }
        if (!st.room(3)) {
            overflow = 1;
            break;
        }
#pragma unroll
        for (int s = 3; s >= 0; --s) {
            if (c[s] >= 0 && list.admits(key[s]) && s != first) stackPush(st, spill, packEntry(c[s], key[s], idBits));
        }
        id = -1;
#pragma unroll
        for (int s = 0; s < 4; ++s)
            if (s == first) id = c[s];
        if (first < 0) {
            while (!st.empty()) {
                const int v = stackPop(st, spill);
                if (list.admits(entryKey(v, idBits))) {
                    id = entryId(v, idBits);
                    break;
                }
            }
        }
    }
}

template <typename L, int K, bool kStats>
__global__ __launch_bounds__(kBlock) void nearestWideKernel(
    const Wide4* __restrict__ wide, const typename L::Item* __restrict__ prims,
    const int* __restrict__ primIds, const V3* __restrict__ points,
    const uint32_t* __restrict__ order, int count, int k, int outStride,
    ClosestResult* __restrict__ out, unsigned long long* __restrict__ stats, int idBits) {
    __shared__ int lds[kLdsDepth * kBlock];
    const int i = xcdBlock() * kBlock + threadIdx.x;
    int nodeTests = 0, leafTests = 0, overflow = 0, opened = 0;
    if (i < count) {
        const int q = order ? static_cast<int>(order[i]) : i;
        KnnList<K> list;
        list.init();
        LaneStack st;
        SpillArea spill;
        st.init(lds);
        nearestWideWalk<L, K, kStats>(wide, prims, primIds, points[q], list, st, spill, nodeTests, leafTests,
                                      overflow, opened, idBits);
        writeList<L, K>(list, k, out + static_cast<size_t>(q) * outStride);
    }
    flushStats<kStats>(stats, nodeTests, leafTests, overflow, opened);
}

// ---- four lanes per query -------------------------------------------------------------
// For query sets too small to fill the GPU one lane per query, a quad of lanes
// walks one query: lane s takes entry s of each 4-wide node (one 128 B load per
// quad), tests its box and, if it is a leaf, its primitive -- in parallel.  Every
// quantity that steers the walk (the k-best list, the stack pointer, the next
// node) is kept identical in the four lanes: candidates are exchanged within the
// quad (DPP quad broadcasts) and offered in slot order by all four lanes.  The
// quad's stack is a private LDS column written by the pushing lanes.

template <int S>
__device__ __forceinline__ int quadGet(int v) {
    return __builtin_amdgcn_mov_dpp(v, S | (S << 2) | (S << 4) | (S << 6), 0xF, 0xF, false);
}
template <int S>
__device__ __forceinline__ float quadGetF(float v) {
    return __int_as_float(quadGet<S>(__float_as_int(v)));
}

constexpr int kQuadDepth = 96;  // 3 x the deepest 4-wide level a tree here can have

// DPP quad permutation: lane s of each quad receives lane P[s] (2 bits per lane).
template <int P>
__device__ __forceinline__ int quadPerm(int v) {
    return __builtin_amdgcn_mov_dpp(v, P, 0xF, 0xF, false);
}
constexpr int kQuadXor1 = 1 | (0 << 2) | (3 << 4) | (2 << 6);
constexpr int kQuadXor2 = 2 | (3 << 2) | (0 << 4) | (1 << 6);
constexpr int kQuadPrev = 0 | (0 << 2) | (1 << 4) | (2 << 6);

// The k-best list of a quad, spread over its four lanes: lane s holds entries
// [s E, s E + E) of the sorted list (E = K / 4).  Same contents and order as
// KnnList<K> after the same offers: an offer (the same candidate in all four lanes)
// goes to position pos = number of entries not farther than it (a quad sum of the
// lanes' counts); entries at pos and after move one place down, the one crossing a
// lane boundary coming from the previous lane.  Every lane keeps the list's last
// entry (kd, kid) for pruning.  Per offer this is E compares and 2 E selects per
// lane instead of K of each in every lane.
template <int K>
struct QuadKnnList {
    static constexpr int E = K / 4;
    static_assert(K % 4 == 0, "K must be a multiple of 4");
    float d[E];
    int id[E];
    int n;
    float kd;
    int kid;
    __device__ __forceinline__ void init() {
        n = 0;
#pragma unroll
        for (int j = 0; j < E; ++j) {
            d[j] = kInf;
            id[j] = 0x7fffffff;
        }
        kd = kInf;
        kid = 0x7fffffff;
    }
    __device__ __forceinline__ bool admits(float lb) const { return n < K || lb <= kd; }
    __device__ __forceinline__ void offer(float nd, int nid, int slot) {
        if (n >= K && !closer(nd, nid, kd, kid)) return;
        int before = 0;
#pragma unroll
        for (int j = 0; j < E; ++j) before += closer(nd, nid, d[j], id[j]) ? 0 : 1;
        int pos = before + quadPerm<kQuadXor1>(before);
        pos += quadPerm<kQuadXor2>(pos);
        const float pd = __int_as_float(quadPerm<kQuadPrev>(__float_as_int(d[E - 1])));
        const int pi = quadPerm<kQuadPrev>(id[E - 1]);
#pragma unroll
        for (int j = E - 1; j >= 0; --j) {
            const int g = slot * E + j;
            const float od = j > 0 ? d[j - 1] : pd;
            const int oi = j > 0 ? id[j - 1] : pi;
            d[j] = g > pos ? od : (g == pos ? nd : d[j]);
            id[j] = g > pos ? oi : (g == pos ? nid : id[j]);
        }
        n = n < K ? n + 1 : K;
        kd = quadGetF<3>(d[E - 1]);
        kid = quadGet<3>(id[E - 1]);
    }
    __device__ __forceinline__ void store(ClosestResult* dst, int k, int slot) const {
#pragma unroll
        for (int j = 0; j < E; ++j) {
            const int g = slot * E + j;
            if (g < k) {
                ClosestResult r{0u, -1};
                if (g < n) r = {floatBits(d[j]), id[j]};
                dst[g] = r;
            }
        }
    }
};

// k = 1: one entry, kept in every lane.
template <>
struct QuadKnnList<1> {
    float kd;
    int kid;
    int n;
    __device__ __forceinline__ void init() {
        kd = kInf;
        kid = 0x7fffffff;
        n = 0;
    }
    __device__ __forceinline__ bool admits(float lb) const { return n == 0 || lb <= kd; }
    __device__ __forceinline__ void offer(float nd, int nid, int) {
        if (n == 0 || closer(nd, nid, kd, kid)) {
            kd = nd;
            kid = nid;
        }
        n = 1;
    }
    __device__ __forceinline__ void store(ClosestResult* dst, int k, int slot) const {
        if (slot == 0 && k > 0) dst[0] = n ? ClosestResult{floatBits(kd), kid} : ClosestResult{0u, -1};
    }
};

template <typename L, int K, bool kStats>
__global__ __launch_bounds__(kBlock) void nearestQuadKernel(
    const Wide4* __restrict__ wide, const typename L::Item* __restrict__ prims,
    const int* __restrict__ primIds, const V3* __restrict__ points,
    const uint32_t* __restrict__ order, int count, int k, int outStride,
    ClosestResult* __restrict__ out, unsigned long long* __restrict__ stats, int idBits) {
    __shared__ int lds[kQuadDepth * (kBlock / 4)];
    const int t = xcdBlock() * kBlock + threadIdx.x;
    const int qi = t >> 2;
    const int slot = threadIdx.x & 3;
    int nodeTests = 0, leafTests = 0, overflow = 0, opened = 0;
    if (qi < count) {  // the same for the four lanes of a quad
        const int q = order ? static_cast<int>(order[qi]) : qi;
        const V3 p = points[q];
        QuadKnnList<K> list;
        list.init();
        LdsInt* stk = (LdsInt*)(lds + (threadIdx.x >> 2) * kQuadDepth);
        int sp = 0;
        // The successor is requested before this node's leaves are tested: offers can
        // only shrink the k-th distance, so the walk continues either with the nearest
        // child that survives the box tests now or, if none survives the offers,
        // with the top of the stack -- that one is requested speculatively and the
        // entry is reloaded only when the offers pruned it.
        WideEntry e = wide[0].c[slot];
        while (true) {
            if (kStats && slot == 0) ++opened;
            const int ch = e.child;
            const bool leaf = ch < 0 && ch != kWideEmpty;
            typename L::Item it;
            int pid = 0;
            if (leaf) {
                it = prims[-ch - 1];
                pid = primIds[-ch - 1];
            }
            float d = kInf;
            bool adm = false;
            if (ch != kWideEmpty) {
                d = boxDist2(p, entryBox(e));
                adm = list.admits(d);
            }
            if (adm) ++nodeTests;
            int spec = -1;
            {
                const bool in0 = adm && !leaf;
                const float key0 = in0 ? d : kInf;
                const int child0 = in0 ? ch : -1;
                const float k0 = quadGetF<0>(key0), k1 = quadGetF<1>(key0), k2 = quadGetF<2>(key0), k3 = quadGetF<3>(key0);
                const int c0 = quadGet<0>(child0), c1 = quadGet<1>(child0), c2 = quadGet<2>(child0), c3 = quadGet<3>(child0);
                float bk = kInf;
                if (c3 >= 0) { spec = c3; bk = k3; }
                if (c2 >= 0 && !(bk < k2)) { spec = c2; bk = k2; }
                if (c1 >= 0 && !(bk < k1)) { spec = c1; bk = k1; }
                if (c0 >= 0 && !(bk < k0)) { spec = c0; bk = k0; }
                for (int j = sp - 1; spec < 0 && j >= 0; --j) {
                    const int v = stk[j];
                    if (list.admits(entryKey(v, idBits))) spec = entryId(v, idBits);
                }
            }
            // (the root when there is no successor: a cached, harmless read)
            const WideEntry se = wide[spec >= 0 ? spec : 0].c[slot];
            float cd = kInf;
            int cid = 0x7fffffff;
            int cand = 0;
            if (adm && leaf) {
                ++leafTests;
                cd = L::dist2(p, it);
                cid = pid;
                cand = 1;
            }
            {
                const float d0 = quadGetF<0>(cd), d1 = quadGetF<1>(cd), d2 = quadGetF<2>(cd), d3 = quadGetF<3>(cd);
                const int i0 = quadGet<0>(cid), i1 = quadGet<1>(cid), i2 = quadGet<2>(cid), i3 = quadGet<3>(cid);
                const int o0 = quadGet<0>(cand), o1 = quadGet<1>(cand), o2 = quadGet<2>(cand), o3 = quadGet<3>(cand);
                if (o0) list.offer(d0, i0, slot);
                if (o1) list.offer(d1, i1, slot);
                if (o2) list.offer(d2, i2, slot);
                if (o3) list.offer(d3, i3, slot);
            }
            const bool in = adm && !leaf && list.admits(d);
            const float key = in ? d : kInf;
            const int child = in ? ch : -1;
            const float k0 = quadGetF<0>(key), k1 = quadGetF<1>(key), k2 = quadGetF<2>(key), k3 = quadGetF<3>(key);
            const int c0 = quadGet<0>(child), c1 = quadGet<1>(child), c2 = quadGet<2>(child), c3 = quadGet<3>(child);
            // rank of this lane's child among the quad's surviving children (ties by slot)
            int rank = 0;
            rank += (slot != 0) && (k0 < key || (k0 == key && 0 < slot));
            rank += (slot != 1) && (k1 < key || (k1 == key && 1 < slot));
            rank += (slot != 2) && (k2 < key || (k2 == key && 2 < slot));
            rank += (slot != 3) && (k3 < key || (k3 == key && 3 < slot));
            const int m = (c0 >= 0) + (c1 >= 0) + (c2 >= 0) + (c3 >= 0);
            if (m > 0) {
                if (sp + m - 1 > kQuadDepth) {
                    if (slot == 0) overflow = 1;
                    break;
                }
                if (in && rank > 0) stk[sp + (m - 1 - rank)] = packEntry(ch, d, idBits);
                sp += m - 1;
                // the nearest survivor: smallest key, then smallest slot
                int next = -1;
                float bk = kInf;
                if (c3 >= 0) { next = c3; bk = k3; }
                if (c2 >= 0 && !(bk < k2)) { next = c2; bk = k2; }
                if (c1 >= 0 && !(bk < k1)) { next = c1; bk = k1; }
                if (c0 >= 0 && !(bk < k0)) { next = c0; bk = k0; }
                if (next == spec) e = se;
                else e = wide[next].c[slot];
            } else {
                while (sp > 0 && !list.admits(entryKey(stk[sp - 1], idBits))) --sp;
                if (sp == 0) break;
                --sp;
                const int next = entryId(stk[sp], idBits);
                if (next == spec) e = se;
                else e = wide[next].c[slot];
            }
        }
        list.store(out + static_cast<size_t>(q) * outStride, k, slot);
    }
    flushStats<kStats>(stats, nodeTests, leafTests, overflow, opened);
}

// Radius with four lanes per query: lane s tests entry s; hits of a node are
// counted, or written at the quad's cursor plus the number of hits in lower slots.
// kMode: kRadiusCount counts; kRadiusFill writes into the query's CSR segment
// (bounded by the next offset), skipping queries whose hits the capture pass already
// holds (counts[q] <= cap); kRadiusCaptureMode counts and writes the first `cap` hits
// into the query's fixed slot of `keys`, so that one walk serves every query with at
// most that many hits.
enum RadiusMode : int { kRadiusCount = 0, kRadiusFill = 1, kRadiusCaptureMode = 2 };
template <typename L, int kMode, bool kStats>
__global__ __launch_bounds__(kBlock) void radiusQuadKernel(
    const Wide4* __restrict__ wide, const typename L::Item* __restrict__ prims,
    const int* __restrict__ primIds, const V3* __restrict__ points, const uint32_t* __restrict__ order,
    int count, float r2, int* __restrict__ counts, const int* __restrict__ offsets,
    uint64_t* __restrict__ keys, unsigned long long* __restrict__ stats, int* __restrict__ maxCount, int cap) {
    __shared__ int lds[kQuadDepth * (kBlock / 4)];
    const int t = xcdBlock() * kBlock + threadIdx.x;
    const int qi = t >> 2;
    const int slot = threadIdx.x & 3;
    int nodeTests = 0, leafTests = 0, overflow = 0, opened = 0;
    int queryHits = 0;
    constexpr bool kFill = kMode != kRadiusCount;  // hits are written somewhere
    bool live = qi < count;
    int q = 0;
    if (live) {
        q = order ? static_cast<int>(order[qi]) : qi;
        if (kMode == kRadiusFill && counts[q] <= cap) live = false;  // captured already
    }
    if (live) {
        const V3 p = points[q];
        int found = 0;
        const long long segBase = kMode == kRadiusCaptureMode ? static_cast<long long>(q) * cap : 0;
        int write = kMode == kRadiusFill ? offsets[q] : 0;
        const int end = kMode == kRadiusFill ? offsets[q + 1] : (kMode == kRadiusCaptureMode ? cap : 0);
        LdsInt* stk = (LdsInt*)(lds + (threadIdx.x >> 2) * kQuadDepth);
        int sp = 0;
        // The next node depends on the box tests only, so its entry is requested
        // before this node's leaves are tested: the leaf loads and tests overlap the
        // next node's load.
        WideEntry e = wide[0].c[slot];
        while (true) {
            if (kStats && slot == 0) ++opened;
            const int ch = e.child;
            const bool leaf = ch < 0 && ch != kWideEmpty;
            typename L::Item it;
            int pid = 0;
            if (leaf) {
                it = prims[-ch - 1];
                if (kFill) pid = primIds[-ch - 1];
            }
            const bool adm = ch != kWideEmpty && boxDist2(p, entryBox(e)) <= r2;
            if (adm) ++nodeTests;
            const int child = (adm && !leaf) ? ch : -1;
            const int c0 = quadGet<0>(child), c1 = quadGet<1>(child), c2 = quadGet<2>(child), c3 = quadGet<3>(child);
            const int m = (c0 >= 0) + (c1 >= 0) + (c2 >= 0) + (c3 >= 0);
            int next = -1;
            if (m > 0) {
                if (sp + m - 1 > kQuadDepth) {
                    if (slot == 0) overflow = 1;
                    break;
                }
                // the lowest surviving slot continues; the others are pushed in slot order
                const int first = c0 >= 0 ? 0 : (c1 >= 0 ? 1 : (c2 >= 0 ? 2 : 3));
                const int below = (slot > 0 && c0 >= 0) + (slot > 1 && c1 >= 0) + (slot > 2 && c2 >= 0);
                if (child >= 0 && slot != first) stk[sp + below - 1] = child;
                sp += m - 1;
                next = first == 0 ? c0 : (first == 1 ? c1 : (first == 2 ? c2 : c3));
            } else if (sp > 0) {
                --sp;
                next = stk[sp];
            }
            WideEntry ne;
            if (next >= 0) ne = wide[next].c[slot];
            int hit = 0;
            float d2 = 0.0f;
            if (adm && leaf) {
                ++leafTests;
                d2 = L::dist2(p, it);
                hit = d2 <= r2 ? 1 : 0;
            }
            const int h0 = quadGet<0>(hit), h1 = quadGet<1>(hit), h2 = quadGet<2>(hit), h3 = quadGet<3>(hit);
            if (kFill) {
                const int before = (slot > 0 ? h0 : 0) + (slot > 1 ? h1 : 0) + (slot > 2 ? h2 : 0);
                const int at = write + before;
                if (hit && at < end)
                    keys[segBase + at] = (static_cast<uint64_t>(floatBits(d2)) << 32) | static_cast<uint32_t>(pid);
                write += h0 + h1 + h2 + h3;
            } else {
                found += hit;
            }
            if (next < 0) break;
            e = ne;
        }
        if (kMode == kRadiusCount) {
            const int total = quadGet<0>(found) + quadGet<1>(found) + quadGet<2>(found) + quadGet<3>(found);
            if (slot == 0) counts[q] = total;
            queryHits = total;
        } else if (kMode == kRadiusCaptureMode) {
            if (slot == 0) counts[q] = write;
            queryHits = write;
        }
    }
    if (kMode != kRadiusFill) waveMaxAtomic(maxCount, queryHits);
    flushStats<kStats>(stats, nodeTests, leafTests, overflow, opened);
}

template <typename V>
__device__ __forceinline__ V waveXor(V v, int m) {
    return __shfl_xor(v, m, 64);
}

// Bitonic sort of one (key, value) pair per lane, ascending by (key, value), over
// aligned groups of `width` lanes (a power of two, at most 64; the default sorts
// the whole wave).
__device__ __forceinline__ void waveSortPairs(float& k, int& v, int lane, int width = 64) {
    for (int size = 2; size <= width; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const float ok = waveXor(k, stride);
            const int ov = waveXor(v, stride);
            const bool up = (lane & size) == 0 || size == width;
            const bool low = (lane & stride) == 0;
            const bool otherFirst = closer(ok, ov, k, v);
            // the lower lane of an ascending pair keeps the smaller, the upper the larger
            const bool take = (low == up) ? otherFirst : !otherFirst;
            if (take) {
                k = ok;
                v = ov;
            }
        }
    }
}

// Bitonic merge of a bitonic sequence (one pair per lane, aligned groups of `width`
// lanes) into ascending order.
__device__ __forceinline__ void waveMergePairs(float& k, int& v, int lane, int width = 64) {
    for (int stride = width >> 1; stride > 0; stride >>= 1) {
        const float ok = waveXor(k, stride);
        const int ov = waveXor(v, stride);
        const bool low = (lane & stride) == 0;
        const bool otherFirst = closer(ok, ov, k, v);
        if (low == otherFirst) {
            k = ok;
            v = ov;
        }
    }
}

// Radius capture with one wave per query, for query sets too small to fill the
// device four lanes per query.  The wave's 16 quads expand up to 16 nodes of the
// query's walk per round (quad j takes the j-th entry from the top of a stack shared
// in LDS, lane s tests entry s of it); admitted interior children are pushed in lane
// order, hits are written at the query's cursor in lane order.  The walk order does
// not matter: the segment sort orders the hits by (dist2, index).  Stack bound: a
// round pops w entries and pushes at most 4w, and w is held to
// (kWaveStack - kWaveHeadroom - sp) / 3 (at least 1), so sp stays below
// kWaveStack - kWaveHeadroom except while popping one entry per round, which adds at
// most 3 per 4-wide level -- within the headroom for trees up to 32 levels deep
// (the quad kernels' stack assumes the same depth).
constexpr int kWaveStack = 1024;
constexpr int kWaveHeadroom = 3 * 32 + 4;
constexpr int kInsertMax = 16;  // a round's candidates inserted one by one up to this many (measured: 4 -> 0.473 ms, 16 -> 0.425, 64 -> 0.422 for C2 kNN16; the sort and merge 0.493)

// Bitonic sort by (dist2, index) of up to 64 pairs held two per lane by a group of
// 32 lanes (lane l holds elements l and l + 32; empty elements are (inf, INT_MAX)).
__device__ __forceinline__ void groupSort64x2(float& ak, int& av, float& bk, int& bv, int gl) {
    for (int size = 2; size <= 64; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            if (stride == 32) {  // (size 64) elements l and l + 32 meet in the lane
                if (closer(bk, bv, ak, av)) {
                    const float tk = ak;
                    const int tv = av;
                    ak = bk;
                    av = bv;
                    bk = tk;
                    bv = tv;
                }
                continue;
            }
            const float oak = waveXor(ak, stride), obk = waveXor(bk, stride);
            const int oav = waveXor(av, stride), obv = waveXor(bv, stride);
            const bool low = (gl & stride) == 0;
            const bool upA = (gl & size) == 0 || size == 64;
            const bool upB = ((gl + 32) & size) == 0 || size == 64;
            const bool aOther = closer(oak, oav, ak, av), bOther = closer(obk, obv, bk, bv);
            if ((low == upA) ? aOther : !aOther) {
                ak = oak;
                av = oav;
            }
            if ((low == upB) ? bOther : !bOther) {
                bk = obk;
                bv = obv;
            }
        }
    }
}

// G lanes per query (G = 64: one wave; G = 32: two queries per wave, their groups'
// ballots and shuffles masked to the group, the LDS stack split between them): the
// group's quads expand up to G / 4 nodes of the walk per round.  A captured segment of
// at most 64 hits is left sorted (over the group's lanes, two per lane above G = 32).
template <typename L, bool kStats, int G>
__global__ __launch_bounds__(kBlock) void radiusWaveKernel(
    const Wide4* __restrict__ wide, const typename L::Item* __restrict__ prims,
    const int* __restrict__ primIds, const V3* __restrict__ points, const uint32_t* __restrict__ order,
    int count, float r2, int* __restrict__ counts, uint64_t* __restrict__ capture,
    unsigned long long* __restrict__ stats, int* __restrict__ maxCount, int cap) {
    static_assert(G == 32 || G == 64, "lanes per query");
    constexpr int kPerWave = 64 / G;
    constexpr int kQuads = G / 4;
    constexpr int kStack = kWaveStack / kPerWave;
    __shared__ int lds[kWaveStack * (kBlock / 64)];
    const int lane = threadIdx.x & 63;
    const int gl = lane & (G - 1);
    const int gb = lane - gl;
    const unsigned long long gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull) << gb;
    const int quad = gl >> 2;
    const int slot = lane & 3;
    const int qi = (xcdBlock() * (kBlock / 64) + (threadIdx.x >> 6)) * kPerWave + lane / G;
    const unsigned long long below = ((1ull << lane) - 1ull) & gmask;
    int nodeTests = 0, leafTests = 0, overflow = 0, opened = 0;
    int queryHits = 0;
    if (qi < count) {  // the same in the whole group
        const int q = order ? static_cast<int>(order[qi]) : qi;
        const V3 p = points[q];
        LdsInt* stk = (LdsInt*)(lds + (threadIdx.x >> 6) * kWaveStack + (lane / G) * kStack);
        uint64_t* seg = capture + static_cast<long long>(q) * cap;
        if (gl == 0) stk[0] = 0;
        int sp = 1;
        int found = 0;
        while (sp > 0) {
            int width = (kStack - kWaveHeadroom - sp) / 3;
            width = width < 1 ? 1 : (width > kQuads ? kQuads : width);
            const int take = width < sp ? width : sp;
            const int node = quad < take ? stk[sp - 1 - quad] : -1;
            sp -= take;
            if (kStats && slot == 0 && node >= 0) ++opened;
            WideEntry e{};
            int ch = kWideEmpty;
            if (node >= 0) {
                e = wide[node].c[slot];
                ch = e.child;
            }
            const bool leaf = ch < 0 && ch != kWideEmpty;
            typename L::Item it;
            int pid = 0;
            if (leaf) {
                it = prims[-ch - 1];
                pid = primIds[-ch - 1];
            }
            const bool adm = ch != kWideEmpty && boxDist2(p, entryBox(e)) <= r2;
            if (adm) ++nodeTests;
            const bool push = adm && !leaf;
            const unsigned long long pm = __ballot(push) & gmask;
            const int pushes = __popcll(pm);
            if (sp + pushes > kStack) {  // deeper than the bound above allows
                if (gl == 0) overflow = 1;
                break;
            }
            if (push) stk[sp + __popcll(pm & below)] = ch;
            sp += pushes;
            bool hit = false;
            float d2 = 0.0f;
            if (adm && leaf) {
                ++leafTests;
                d2 = L::dist2(p, it);
                hit = d2 <= r2;
            }
            const unsigned long long hm = __ballot(hit) & gmask;
            if (hit) {
                const int at = found + __popcll(hm & below);
                if (at < cap) seg[at] = (static_cast<uint64_t>(floatBits(d2)) << 32) | static_cast<uint32_t>(pid);
            }
            found += __popcll(hm);
        }
        if (found > 1 && found <= G && found <= cap) {
            // a segment of at most G hits leaves sorted by (dist2, index): the order
            // of the 64-bit keys, dist2 being non-negative
            int w = 2;
            while (w < found) w <<= 1;
            float sk = kInf;
            int sv = 0x7fffffff;
            if (gl < found) {
                const uint64_t key = seg[gl];
                sk = bitsFloat(static_cast<uint32_t>(key >> 32));
                sv = static_cast<int>(static_cast<uint32_t>(key));
            }
            waveSortPairs(sk, sv, lane, w);
            if (gl < found) seg[gl] = (static_cast<uint64_t>(floatBits(sk)) << 32) | static_cast<uint32_t>(sv);
        } else if (G == 32 && found > G && found <= 64 && found <= cap) {
            float ak = kInf, bk = kInf;
            int av = 0x7fffffff, bv = 0x7fffffff;
            {
                const uint64_t key = seg[gl];
                ak = bitsFloat(static_cast<uint32_t>(key >> 32));
                av = static_cast<int>(static_cast<uint32_t>(key));
            }
            if (gl + 32 < found) {
                const uint64_t key = seg[gl + 32];
                bk = bitsFloat(static_cast<uint32_t>(key >> 32));
                bv = static_cast<int>(static_cast<uint32_t>(key));
            }
            groupSort64x2(ak, av, bk, bv, gl);
            seg[gl] = (static_cast<uint64_t>(floatBits(ak)) << 32) | static_cast<uint32_t>(av);
            if (gl + 32 < found) seg[gl + 32] = (static_cast<uint64_t>(floatBits(bk)) << 32) | static_cast<uint32_t>(bv);
        }
        if (gl == 0) counts[q] = found;
        queryHits = found;
    }
    waveMaxAtomic(maxCount, queryHits);
    flushStats<kStats>(stats, nodeTests, leafTests, overflow, opened);
}

// Closest / kNN with one wave per query (the radius capture's layout: 16 quads expand
// up to 16 nodes of one walk per round from a shared LDS stack, lane s of a quad
// tests entry s).  The k-best list is spread one entry per lane (lanes 0..K-1,
// sorted by (dist2, index), (inf, INT_MAX) when empty); a round's leaf candidates
// that beat the list's K-th are sorted across the wave (bitonic network) and merged
// in (the elementwise smaller of the list and the reversed candidates holds the K
// best, a bitonic merge sorts it).  Interior children are pushed nearest-parent
// first and, within a parent, nearest on top, so the next round expands about the
// nearest 16 without a sort (for Gaussians: sorted by box distance across the wave).  The admission rule is the lane kernels' (n < K or
// bound <= K-th), so the list ends as the K best by (dist2, index) whatever the
// walk order.
template <typename L, int K, int G, bool kStats>
__global__ __launch_bounds__(kBlock) void nearestWaveKernel(
    const Wide4* __restrict__ wide, const typename L::Item* __restrict__ prims,
    const int* __restrict__ primIds, const V3* __restrict__ points, const uint32_t* __restrict__ order,
    int count, int k, int outStride, ClosestResult* __restrict__ out, unsigned long long* __restrict__ stats,
    int idBits) {
    // Gaussian leaves cost the most to test, so their walks pay for a sorted push
    // order (fewer nodes opened); triangles and points push quad by quad.
    constexpr bool kSortPushes = L::kKind == 2;
    static_assert(G == 16 || G == 32 || G == 64, "lanes per query");
    static_assert(K <= G, "the k-best list takes one lane per entry");
    constexpr int kPerWave = 64 / G;              // queries per wave
    constexpr int kQuads = G / 4;                 // nodes a query expands per round
    constexpr int kStack = kWaveStack / kPerWave;  // stack entries per query
    __shared__ int lds[kWaveStack * (kBlock / 64)];
    __shared__ int ldsCmp[2 * 64 * (kBlock / 64)];  // per wave: candidate keys, values
    const int lane = threadIdx.x & 63;
    const int gl = lane & (G - 1);   // lane within the query's group
    const int gb = lane - gl;        // the group's first lane
    const unsigned long long gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull) << gb;
    LdsInt* cmpK = (LdsInt*)(ldsCmp + (threadIdx.x >> 6) * 128) + gb;
    LdsInt* cmpV = cmpK + 64;
    const int quad = gl >> 2;
    const int slot = lane & 3;
    const int qi = (xcdBlock() * (kBlock / 64) + (threadIdx.x >> 6)) * kPerWave + lane / G;
    const unsigned long long below = ((1ull << lane) - 1ull) & gmask;
    int nodeTests = 0, leafTests = 0, overflow = 0, opened = 0;
    if (qi < count) {  // the same in the whole group
        const int q = order ? static_cast<int>(order[qi]) : qi;
        const V3 p = points[q];
        LdsInt* stk = (LdsInt*)(lds + (threadIdx.x >> 6) * kWaveStack + (lane / G) * kStack);
        float ld = kInf;  // this lane's list entry (group lanes 0..K-1)
        int lid = 0x7fffffff;
        float kd = kInf;  // the list's K-th (the same in every lane)
        int kid = 0x7fffffff;
        int n = 0;
        auto admits = [&](float lb) { return n < K || lb <= kd; };
        if (gl == 0) stk[0] = packEntry(0, 0.0f, idBits);
        int sp = 1;
        while (sp > 0) {
            int width = (kStack - kWaveHeadroom - sp) / 3;
            width = width < 1 ? 1 : (width > kQuads ? kQuads : width);
            int take = width < sp ? width : sp;
            int node = -1;
            if (quad < take) {
                const int v = stk[sp - 1 - quad];
                if (admits(entryKey(v, idBits))) node = entryId(v, idBits);
            }
            if ((__ballot(quad < take && node < 0) & gmask) != 0) {
                // the bound has overtaken entries the quads would take: drop those of
                // the top G (order kept) and take again, so no quad idles on them
                const int look = sp < G ? sp : G;
                int v = 0;
                bool keep = false;
                if (gl < look) {
                    v = stk[sp - look + gl];
                    keep = admits(entryKey(v, idBits));
                }
                const unsigned long long km = __ballot(keep) & gmask;
                if (keep) stk[sp - look + __popcll(km & below)] = v;
                sp = sp - look + __popcll(km);
                if (sp == 0) break;
                width = (kStack - kWaveHeadroom - sp) / 3;
                width = width < 1 ? 1 : (width > kQuads ? kQuads : width);
                take = width < sp ? width : sp;
                node = quad < take ? entryId(stk[sp - 1 - quad], idBits) : -1;
            }
            sp -= take;
            if (kStats && slot == 0 && node >= 0) ++opened;
            WideEntry e{};
            int ch = kWideEmpty;
            if (node >= 0) {
                e = wide[node].c[slot];
                ch = e.child;
            }
            const bool leaf = ch < 0 && ch != kWideEmpty;
            typename L::Item it;
            int pid = 0;
            if (leaf) {
                it = prims[-ch - 1];
                pid = primIds[-ch - 1];
            }
            float d = kInf;
            bool adm = false;
            if (ch != kWideEmpty) {
                d = boxDist2(p, entryBox(e));
                adm = admits(d);
            }
            if (adm) ++nodeTests;
            float cd = kInf;
            int cid = 0x7fffffff;
            bool better = false;
            if (adm && leaf) {
                ++leafTests;
                cd = L::dist2(p, it);
                cid = pid;
                better = n < K || closer(cd, cid, kd, kid);
            }
            const unsigned long long bm = __ballot(better) & gmask;
            if (K == 1 && bm != 0) {  // the group's smallest candidate replaces the entry
                float ck = better ? cd : kInf;
                int cv = better ? cid : 0x7fffffff;
#pragma unroll
                for (int off = G / 2; off > 0; off >>= 1) {
                    const float ok = waveXor(ck, off);
                    const int ov = waveXor(cv, off);
                    if (closer(ok, ov, ck, cv)) {
                        ck = ok;
                        cv = ov;
                    }
                }
                kd = ck;
                kid = cv;
                ld = gl == 0 ? ck : kInf;
                lid = gl == 0 ? cv : 0x7fffffff;
                n = 1;
            } else if (bm != 0 && __popcll(bm) <= kInsertMax) {
                // a few candidates: each in turn is broadcast to the group, finds its
                // place by one ballot (the list entries closer than it) and the entries
                // from there shift up one lane -- the same K best by (dist2, index)
                // as the sort and merge below, for far fewer instructions
                unsigned long long rest = bm;
                const int nb = __popcll(bm);
                for (int i = 0; i < nb; ++i) {
                    const int src = __ffsll(static_cast<long long>(rest)) - 1;
                    rest &= rest - 1ull;
                    const float c = __shfl(cd, src, 64);
                    const int ci = __shfl(cid, src, 64);
                    const int pos = __popcll(__ballot(gl < K && closer(ld, lid, c, ci)) & gmask);
                    const float pk = __shfl(ld, lane - 1, 64);
                    const int pv = __shfl(lid, lane - 1, 64);
                    if (gl < K && gl >= pos) {
                        ld = gl == pos ? c : pk;
                        lid = gl == pos ? ci : pv;
                    }
                }
                n = n + nb;
                n = n < K ? n : K;
                kd = __shfl(ld, gb + K - 1, 64);
                kid = __shfl(lid, gb + K - 1, 64);
            } else if (bm != 0) {
                // the nb candidates move to lanes 0..nb-1 (through LDS), are sorted
                // over the smallest power-of-two width m >= max(nb, K), and merged:
                // the list (lanes >= K hold (inf, INT_MAX)) against the reversed
                // candidates, elementwise, is a bitonic sequence holding the K best
                const int nb = __popcll(bm);
                int m = K;
                while (m < nb) m <<= 1;
                if (better) {
                    const int at = __popcll(bm & below);
                    cmpK[at] = __float_as_int(cd);
                    cmpV[at] = cid;
                }
                float ck = gl < nb ? __int_as_float(cmpK[gl]) : kInf;
                int cv = gl < nb ? cmpV[gl] : 0x7fffffff;
                waveSortPairs(ck, cv, lane, m);
                const float rk = __shfl(ck, gb + m - 1 - (gl & (m - 1)), 64);
                const int rv = __shfl(cv, gb + m - 1 - (gl & (m - 1)), 64);
                float mk = ld;
                int mv = lid;
                if (gl < m && closer(rk, rv, mk, mv)) {
                    mk = rk;
                    mv = rv;
                }
                waveMergePairs(mk, mv, lane, m);
                ld = gl < K ? mk : kInf;
                lid = gl < K ? mv : 0x7fffffff;
                n = n + __popcll(bm);
                n = n < K ? n : K;
                kd = __shfl(ld, gb + K - 1, 64);
                kid = __shfl(lid, gb + K - 1, 64);
            }
            // interior children that survive the updated bound: quad j popped the j-th
            // nearest entry, so its children go above those of quads j+1.., each
            // quad's nearest child on top of its group
            const bool push = adm && !leaf && admits(d);
            const unsigned long long pm = __ballot(push) & gmask;
            const int pushes = __popcll(pm);
            if (pushes > 0) {
                if (sp + pushes > kStack) {  // deeper than the stack bound allows
                    if (gl == 0) overflow = 1;
                    break;
                }
                if (kSortPushes) {  // all of the round's children in box-distance order
                    float pk = push ? d : kInf;
                    int pv = push ? ch : 0x7fffffff;
                    waveSortPairs(pk, pv, lane, G);
                    if (gl < pushes) stk[sp + pushes - 1 - gl] = packEntry(pv, pk, idBits);
                } else {
                    const float key = push ? d : kInf;
                    const float k0 = quadGetF<0>(key), k1 = quadGetF<1>(key), k2 = quadGetF<2>(key), k3 = quadGetF<3>(key);
                    int rank = 0;
                    rank += (slot != 0) && (k0 < key || (k0 == key && 0 < slot));
                    rank += (slot != 1) && (k1 < key || (k1 == key && 1 < slot));
                    rank += (slot != 2) && (k2 < key || (k2 == key && 2 < slot));
                    rank += (slot != 3) && (k3 < key || (k3 == key && 3 < slot));
                    const int mine = __popcll((pm >> (gb + 4 * quad)) & 0xFull);
                    const int deeper = quad < kQuads - 1 ? __popcll(pm >> (gb + 4 * (quad + 1))) : 0;
                    if (push) stk[sp + deeper + (mine - 1 - rank)] = packEntry(ch, d, idBits);
                }
                sp += pushes;
            }
        }
        if (gl < k) {
            ClosestResult r{0u, -1};
            if (gl < n) r = {floatBits(ld), lid};
            out[static_cast<size_t>(q) * outStride + gl] = r;
        }
    }
    flushStats<kStats>(stats, nodeTests, leafTests, overflow, opened);
}

// Closest-hit rays with one wave per ray (the layout of the wave kernels above: 16
// quads expand up to 16 stacked nodes of one walk per round, lane s of a quad tests
// entry s).  A round's hits reduce to the wave's smallest (t, then primitive index --
// the lane kernels' rule for equal t), boxes are tested against the best t so far
// with the lane kernels' slack, and popped entries whose entry distance (kept in the
// stack entry, truncated: a lower bound) is beyond it are dropped.  Interior children
// go on the stack nearest-parent first, each parent's nearest on top.
template <bool kStats, bool kSave>
__global__ __launch_bounds__(kBlock) void rayWaveKernel(const Wide4* __restrict__ wide,
                                                        const Tri* __restrict__ tris,
                                                        const Ray* __restrict__ rays,
                                                        const uint32_t* __restrict__ order, int count,
                                                        uint64_t* __restrict__ out,
                                                        unsigned long long* __restrict__ stats,
                                                        const int* __restrict__ primIds,
                                                        int* __restrict__ hitPrim, int idBits) {
    __shared__ int lds[kWaveStack * (kBlock / 64)];
    const int lane = threadIdx.x & 63;
    const int quad = lane >> 2;
    const int slot = lane & 3;
    const int qi = xcdBlock() * (kBlock / 64) + (threadIdx.x >> 6);
    int nodeTests = 0, leafTests = 0, overflow = 0, opened = 0;
    if (qi < count) {  // the same in the whole wave
        const int q = order ? static_cast<int>(order[qi]) : qi;
        const Ray ray = rays[q];
        float best = ray.tMax;
        bool any = false;
        int bestId = 0x7fffffff;
        LdsInt* stk = (LdsInt*)(lds + (threadIdx.x >> 6) * kWaveStack);
        if (lane == 0) stk[0] = packEntry(0, 0.0f, idBits);
        int sp = 1;
        while (sp > 0) {
            int width = (kWaveStack - kWaveHeadroom - sp) / 3;
            width = width < 1 ? 1 : (width > 16 ? 16 : width);
            const int take = width < sp ? width : sp;
            int node = -1;
            if (quad < take) {
                const int v = stk[sp - 1 - quad];
                if (entryKey(v, idBits) <= best * (1.0f + 1e-6f) + 1e-6f) node = entryId(v, idBits);
            }
            sp -= take;
            if (kStats && slot == 0 && node >= 0) ++opened;
            WideEntry e{};
            int ch = kWideEmpty;
            if (node >= 0) {
                e = wide[node].c[slot];
                ch = e.child;
            }
            float tEnter = kInf;
            bool adm = false;
            if (ch != kWideEmpty) adm = rayBox(ray, entryBox(e), best, tEnter);
            if (adm) ++nodeTests;
            const bool leaf = ch < 0 && ch != kWideEmpty;
            float ht = kInf;
            int hid = 0x7fffffff;
            if (adm && leaf) {
                ++leafTests;
                float t;
                // the lane kernels' rule: a nearer hit, or an equal one once there is a hit
                if (rayTri(ray, tris[-ch - 1], t) && (t < best || (any && t == best))) {
                    ht = t;
                    hid = primIds[-ch - 1];
                }
            }
            if (__ballot(ht < kInf) != 0) {
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    const float ot = waveXor(ht, off);
                    const int oi = waveXor(hid, off);
                    if (closer(ot, oi, ht, hid)) {
                        ht = ot;
                        hid = oi;
                    }
                }
                if (ht < best) {
                    best = ht;
                    bestId = hid;
                    any = true;
                } else if (any && ht == best && hid < bestId) {
                    bestId = hid;
                }
            }
            const bool push = adm && !leaf && tEnter <= best * (1.0f + 1e-6f) + 1e-6f;
            const unsigned long long pm = __ballot(push);
            const int pushes = __popcll(pm);
            if (pushes > 0) {
                if (sp + pushes > kWaveStack) {  // deeper than the stack bound allows
                    if (lane == 0) overflow = 1;
                    break;
                }
                const float key = push ? tEnter : kInf;
                const float k0 = quadGetF<0>(key), k1 = quadGetF<1>(key), k2 = quadGetF<2>(key), k3 = quadGetF<3>(key);
                int rank = 0;
                rank += (slot != 0) && (k0 < key || (k0 == key && 0 < slot));
                rank += (slot != 1) && (k1 < key || (k1 == key && 1 < slot));
                rank += (slot != 2) && (k2 < key || (k2 == key && 2 < slot));
                rank += (slot != 3) && (k3 < key || (k3 == key && 3 < slot));
                const int mine = __popcll((pm >> (4 * quad)) & 0xFull);
                const int deeper = quad < 15 ? __popcll(pm >> (4 * (quad + 1))) : 0;
                if (push) stk[sp + deeper + (mine - 1 - rank)] = packEntry(ch, tEnter, idBits);
                sp += pushes;
            }
        }
        if (lane == 0) {
            out[q] = any ? ((uint64_t{1} << 32) | floatBits(best)) : 0ull;
            if (kSave) hitPrim[q] = any ? bestId : -1;
        }
    }
    flushStats<kStats>(stats, nodeTests, leafTests, overflow, opened);
}

// ---- radius: count and fill passes ---------------------------------------------------
// kFill = false counts matches per query; kFill = true writes the 64-bit keys
// (dist2 bits << 32 | index) into the query's CSR segment (bounded by the next
// offset, so a fill can never write past its segment).
template <typename L, bool kWide, bool kFill, bool kStats>
__global__ __launch_bounds__(kBlock) void radiusKernel(
    const Node2* __restrict__ nodes, const Wide4* __restrict__ wide,
    const typename L::Item* __restrict__ prims, const int* __restrict__ primIds,
    const V3* __restrict__ points, const uint32_t* __restrict__ order, int count, float r2,
    int* __restrict__ counts, const int* __restrict__ offsets, uint64_t* __restrict__ keys,
    unsigned long long* __restrict__ stats, int* __restrict__ maxCount) {
    __shared__ int lds[kLdsDepth * kBlock];
    int queryHits = 0;
    const int i = xcdBlock() * kBlock + threadIdx.x;
    int nodeTests = 0, leafTests = 0, overflow = 0, opened = 0;
    if (i < count) {
        const int q = order ? static_cast<int>(order[i]) : i;
        const V3 p = points[q];
        int found = 0;
        int write = kFill ? offsets[q] : 0;
        const int end = kFill ? offsets[q + 1] : 0;
        auto leafTest = [&](int slot) {
            ++leafTests;
            const float d2 = L::dist2(p, prims[slot]);
            if (d2 <= r2) {
                if (kFill) {
                    if (write < end)
                        keys[write++] = (static_cast<uint64_t>(floatBits(d2)) << 32) |
                                        static_cast<uint32_t>(primIds[slot]);
                } else {
                    ++found;
                }
            }
        };
        LaneStack st;
        SpillArea spill;
        st.init(lds);
        int id = -1;
        if (!kWide) {
            const Node2 root = nodes[0];
            if (boxDist2(p, nodeBox(root)) <= r2) {
                ++nodeTests;
                id = 0;
            }
        } else {
            id = 0;
        }
        while (id >= 0) {
            if (kStats) ++opened;
            int c[4] = {kWideEmpty, kWideEmpty, kWideEmpty, kWideEmpty};
            bool ok[4] = {false, false, false, false};
            if (!kWide) {
                const Node2 nd = nodes[id];
                const int kid[2] = {nd.left, nd.right};
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    const Node2 cn = nodes[kid[s]];
                    if (boxDist2(p, nodeBox(cn)) > r2) continue;
                    ++nodeTests;
                    if (cn.left < 0) {
                        leafTest(-cn.left - 1);
                    } else {
                        c[s] = kid[s];
                        ok[s] = true;
                    }
                }
            } else {
                WideEntry e[4];
                loadWide(wide, id, e);
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int ch = e[s].child;
                    if (ch == kWideEmpty || boxDist2(p, entryBox(e[s])) > r2) continue;
                    ++nodeTests;
                    if (ch < 0) {
                        leafTest(-ch - 1);
                    } else {
                        c[s] = ch;
                        ok[s] = true;
                    }
                }
            }
            id = descend(c, ok, st, spill, overflow);
        }
        if (!kFill) counts[q] = found;
        queryHits = found;
    }
    if (!kFill) waveMaxAtomic(maxCount, queryHits);
    flushStats<kStats>(stats, nodeTests, leafTests, overflow, opened);
}

// ---- segment sort ----------------------------------------------------------------
// Sorts each CSR segment ascending.  One wave per segment of up to kWaveCap keys
// (bitonic network in LDS); longer segments are handled by one workgroup each
// with a bitonic network over up to kBlockCap keys.
constexpr int kWaveCap = 256;
constexpr int kBlockCap = 8192;

__device__ __forceinline__ void bitonicLds(uint64_t* buf, int cap, int lane, int lanes,
                                           bool blockSync) {
    for (int size = 2; size <= cap; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = lane; i < cap; i += lanes) {
                const int partner = i ^ stride;
                if (partner > i) {
                    const uint64_t a = buf[i];
                    const uint64_t b = buf[partner];
                    const bool up = (i & size) == 0;
                    if (up ? (a > b) : (a < b)) {
                        buf[i] = b;
                        buf[partner] = a;
                    }
                }
            }
            if (blockSync) {
                __syncthreads();
            } else {
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void segSortWaveKernel(const int* __restrict__ offsets,
                                                            uint64_t* __restrict__ keys,
                                                            int count) {
    __shared__ uint64_t smem[kBlock / 64][kWaveCap];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int seg = blockIdx.x * (kBlock / 64) + w;
    if (seg >= count) return;
    const int begin = offsets[seg];
    const int m = offsets[seg + 1] - begin;
    if (m <= 1 || m > kWaveCap) return;
    int cap = 2;
    while (cap < m) cap <<= 1;
    uint64_t* buf = smem[w];
    for (int i = lane; i < cap; i += 64) buf[i] = i < m ? keys[begin + i] : ~0ull;
    __builtin_amdgcn_wave_barrier();
    bitonicLds(buf, cap, lane, 64, false);
    for (int i = lane; i < m; i += 64) keys[begin + i] = buf[i];
}

// Segments of kWaveCap+1 .. kBlockCap keys, listed by segClassifyKernel (info[0] of
// them); each workgroup takes list entries blockIdx.x, blockIdx.x + gridDim.x, ...
__global__ __launch_bounds__(kBlock) void segSortBlockKernel(const int* __restrict__ offsets,
                                                             const int* __restrict__ segList,
                                                             const int* __restrict__ info,
                                                             uint64_t* __restrict__ keys) {
    extern __shared__ uint64_t dyn[];
    const int listed = info[0];
    for (int e = blockIdx.x; e < listed; e += gridDim.x) {
        const int seg = segList[e];
        const int begin = offsets[seg];
        const int m = offsets[seg + 1] - begin;
        int cap = 2;
        while (cap < m) cap <<= 1;
        for (int i = threadIdx.x; i < cap; i += kBlock) dyn[i] = i < m ? keys[begin + i] : ~0ull;
        __syncthreads();
        bitonicLds(dyn, cap, threadIdx.x, kBlock, true);
        for (int i = threadIdx.x; i < m; i += kBlock) keys[begin + i] = dyn[i];
        __syncthreads();
    }
}

// Collects segments longer than the wave cap; info[0] = count of block-sized ones,
// info[1] = count of segments beyond the block cap, info[2] = max segment length.
__global__ void segClassifyKernel(const int* __restrict__ offsets, int count,
                                  int* __restrict__ blockList, int* __restrict__ hugeList,
                                  int* __restrict__ info) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= count) return;
    const int m = offsets[s + 1] - offsets[s];
    if (m > kWaveCap && m <= kBlockCap) blockList[atomicAdd(&info[0], 1)] = s;
    if (m > kBlockCap) hugeList[atomicAdd(&info[1], 1)] = s;
}

__global__ void zeroIntsKernel(int* __restrict__ p, int n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void splitKeysKernel(const uint64_t* __restrict__ keys, int m, uint32_t* __restrict__ lo,
                                uint32_t* __restrict__ hi) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= m) return;
    const uint64_t k = keys[i];
    lo[i] = static_cast<uint32_t>(k);
    hi[i] = static_cast<uint32_t>(k >> 32);
}

__global__ void gatherU32Kernel(const uint32_t* __restrict__ src, const uint32_t* __restrict__ idx, int m,
                                uint32_t* __restrict__ dst) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < m) dst[i] = src[idx[i]];
}

// keys[i] = hiSorted[i] << 32 | lo[ord[i]]
__global__ void joinKeysKernel(const uint32_t* __restrict__ hiSorted, const uint32_t* __restrict__ lo,
                               const uint32_t* __restrict__ ord, int m, uint64_t* __restrict__ keys) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < m) keys[i] = (static_cast<uint64_t>(hiSorted[i]) << 32) | lo[ord[i]];
}

// ---- Morton reordering of queries / rays ------------------------------------------------
// quantize10 over the query (or ray-origin) bounds (src/gpu/morton.h:8-29).
__device__ __forceinline__ uint32_t quant10(float v, float lo, float hi) {
    if (!(hi > lo)) return 0u;
    float f = (v - lo) / (hi - lo);
    f = fminf(fmaxf(f, 0.0f), 1.0f);
    return static_cast<uint32_t>(f * 1023.0f + 0.5f);
}
__device__ __forceinline__ uint32_t spread10b(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void pointBoundsInitKernel(uint32_t* __restrict__ bb) {
    if (threadIdx.x < 3) bb[threadIdx.x] = orderedBits(kInf);
    if (threadIdx.x >= 3 && threadIdx.x < 6) bb[threadIdx.x] = orderedBits(-kInf);
}

// Bounds of points laid out with `stride` floats per element (3 for points, 8 for
// rays: origin first, 8 + 3 for ray directions).
__global__ void pointBoundsKernel(const float* __restrict__ base, int stride, int count,
                                  uint32_t* __restrict__ bb) {
    float lo0 = kInf, lo1 = kInf, lo2 = kInf, hi0 = -kInf, hi1 = -kInf, hi2 = -kInf;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < count; i += kBlock * gridDim.x) {
        const float* p = base + static_cast<size_t>(i) * stride;
        lo0 = fminf(lo0, p[0]);
        lo1 = fminf(lo1, p[1]);
        lo2 = fminf(lo2, p[2]);
        hi0 = fmaxf(hi0, p[0]);
        hi1 = fmaxf(hi1, p[1]);
        hi2 = fmaxf(hi2, p[2]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lo0 = fminf(lo0, __shfl_xor(lo0, off, 64));
        lo1 = fminf(lo1, __shfl_xor(lo1, off, 64));
        lo2 = fminf(lo2, __shfl_xor(lo2, off, 64));
        hi0 = fmaxf(hi0, __shfl_xor(hi0, off, 64));
        hi1 = fmaxf(hi1, __shfl_xor(hi1, off, 64));
        hi2 = fmaxf(hi2, __shfl_xor(hi2, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&bb[0], orderedBits(lo0));
        atomicMin(&bb[1], orderedBits(lo1));
        atomicMin(&bb[2], orderedBits(lo2));
        atomicMax(&bb[3], orderedBits(hi0));
        atomicMax(&bb[4], orderedBits(hi1));
        atomicMax(&bb[5], orderedBits(hi2));
    }
}

// fixedBox: use [-1,1]^3 (ray directions) instead of the measured bounds.
__global__ void mortonKeysKernel(const float* __restrict__ base, int stride, int count,
                                 const uint32_t* __restrict__ bb, int fixedBox, int drop,
                                 uint32_t* __restrict__ keys, uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[4][256];
    for (int i = threadIdx.x; i < 4 * 256; i += kBlock) (&h[0][0])[i] = 0u;
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = fixedBox ? -1.0f : fromOrderedBits(bb[a]);
        hi[a] = fixedBox ? 1.0f : fromOrderedBits(bb[3 + a]);
    }
    __syncthreads();
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < count; i += kBlock * gridDim.x) {
        const float* p = base + static_cast<size_t>(i) * stride;
        const uint32_t code = (spread10b(quant10(p[0], lo[0], hi[0])) |
                              (spread10b(quant10(p[1], lo[1], hi[1])) << 1) |
                              (spread10b(quant10(p[2], lo[2], hi[2])) << 2)) >> drop;
        keys[i] = code;
        atomicAdd(&h[0][code & 255u], 1u);
        atomicAdd(&h[1][(code >> 8) & 255u], 1u);
        atomicAdd(&h[2][(code >> 16) & 255u], 1u);
        atomicAdd(&h[3][code >> 24], 1u);
    }
    __syncthreads();
    flushHistogram(&h[0][0], 4 * 256, hist);
}

__global__ void zeroU64Kernel(unsigned long long* __restrict__ p, int n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

}  // namespace

// ---------------------------------------------------------------------------------------
// Host launchers.

void zeroInts(int* p, int n, hipStream_t s) {
    zeroIntsKernel<<<blocksFor(n, kBlock), kBlock, 0, s>>>(p, n);
}

void zeroStats(unsigned long long* stats, hipStream_t s) {
    zeroU64Kernel<<<blocksFor(kStatWords, kBlock), kBlock, 0, s>>>(stats, kStatWords);
}

void mortonOrder(const float* base, int stride, int count, int boxMode, SortWorkspace& ws,
                 uint32_t* keys, uint32_t* keysSorted, uint32_t* orderOut, uint32_t* bb,
                 hipStream_t s) {
    // The order only steers locality (results do not depend on it), so the code is
    // cut to whole 8-bit digits: the top 16 bits up to 262,144 elements (at most four
    // per cell; the 100k-query terrain walks measured no slower than on the full
    // code), 24 up to 16M, all 30 above -- one or two sort passes fewer.
    int passes = 4, drop = 0;
    if (count <= (1 << 18)) {
        passes = 2;
        drop = 14;
    } else if (count <= (1 << 24)) {
        passes = 3;
        drop = 6;
    }
    ws.reserve(static_cast<size_t>(count));
    sortClear(ws, static_cast<size_t>(count), passes, s);
    int grid = blocksFor(count, kBlock * 4);
    if (grid < 1) grid = 1;
    if (grid > 1024) grid = 1024;
    if (boxMode == kBoxMeasure) {
        pointBoundsInitKernel<<<1, 64, 0, s>>>(bb);
        pointBoundsKernel<<<grid, kBlock, 0, s>>>(base, stride, count, bb);
    }
    mortonKeysKernel<<<grid, kBlock, 0, s>>>(base, stride, count, bb, boxMode == kBoxUnit ? 1 : 0, drop,
                                             keys, ws.hist.p);
    hipCheck(hipGetLastError(), "morton keys");
    sortPairs(ws, keys, nullptr, keysSorted, orderOut, static_cast<size_t>(count), passes, s);
}

// 0 automatic, 1 one lane per query, 2 four lanes per query, 3 one wave per query,
// 4 closest / kNN and rays tuned whatever the query count
int gTraversalMode = 0;

// Shapes of a walk over 4-wide nodes.
// (Group16 / Group32: 16 or 32 lanes per query, four or two queries per wave.)
enum NearestShape : int { kShapeLane = 0, kShapeQuad = 1, kShapeWave = 2, kShapeGroup16 = 3, kShapeGroup32 = 4 };

// Tuned shapes.  The best shape depends on how long and how uneven the walks are,
// which the query count alone does not tell: on the terrain a million closest
// queries run fastest four lanes per query (3.2 ms, one wave per query 4.0), while
// 262,144 queries in and around an icosphere -- long walks of very different
// lengths -- run 5.5x faster one wave per query (1.6 ms against 8.6 one lane per
// query).  From kTuneFrom queries on, the first call for a (tree, k class, query
// count octave) times the three shapes on its own queries and keeps the fastest;
// every shape writes the same bits, so the trials' output is the call's output.
constexpr int kTuneFrom = 65536;

struct TuneKey {
    const void* tree;
    int wideCount, kind, K, octave;
    bool operator<(const TuneKey& o) const {
        if (tree != o.tree) return tree < o.tree;
        if (wideCount != o.wideCount) return wideCount < o.wideCount;
        if (kind != o.kind) return kind < o.kind;
        if (K != o.K) return K < o.K;
        return octave < o.octave;
    }
};
static std::mutex gTuneMutex;
static std::map<TuneKey, int> gTuned;

template <bool kSave>
static void launchRayShape(int shape, const TraverseArgs& a, const Ray* rays, int count, uint64_t* out,
                           int* hitPrim, unsigned long long* stats, hipStream_t s) {
    const Tri* tris = static_cast<const Tri*>(a.items);
    if (shape == kShapeWave) {
        const int grid = blocksFor(count, kBlock / 64);
        const int idBits = stackIdBitsDev(a.wideCount);
        if (stats)
            rayWaveKernel<true, kSave><<<grid, kBlock, 0, s>>>(a.wideNodes, tris, rays, a.order, count, out, stats,
                                                              a.primIds, hitPrim, idBits);
        else
            rayWaveKernel<false, kSave><<<grid, kBlock, 0, s>>>(a.wideNodes, tris, rays, a.order, count, out, nullptr,
                                                               a.primIds, hitPrim, idBits);
    } else {
        const int grid = blocksFor(count, kBlock);
        if (stats)
            rayWideKernel<true, kSave><<<grid, kBlock, 0, s>>>(a.wideNodes, tris, rays, a.order, count, out, stats,
                                                              a.primIds, hitPrim);
        else
            rayWideKernel<false, kSave><<<grid, kBlock, 0, s>>>(a.wideNodes, tris, rays, a.order, count, out,
                                                               nullptr, a.primIds, hitPrim);
    }
}

// Rays over 4-wide nodes run one lane or one wave per ray (same bits).  One lane per
// ray unless mode 3 forces one wave or mode 4 tunes (the first call for a tree, save
// flag and ray-count octave times both and keeps the faster): closest-hit walks are
// short and prune well one lane at a time -- one wave per ray measured 1.3-5x slower
// on the down rays, the city and the icosphere (0.44 vs 2.17 ms, 0.75 vs 2.52, 1.06
// vs 3.04 for 1M rays).
template <bool kSave>
static int rayShape(const TraverseArgs& a, const Ray* rays, int count, uint64_t* out, int* hitPrim,
                    hipStream_t s) {
    if (gTraversalMode == 3) return kShapeWave;
    if (gTraversalMode != 4) return kShapeLane;
    int octave = 0;
    while ((2LL << octave) <= count) ++octave;
    const TuneKey key{a.wideNodes, a.wideCount, 100 + (kSave ? 1 : 0), 0, octave};
    {
        std::lock_guard<std::mutex> lock(gTuneMutex);
        const auto it = gTuned.find(key);
        if (it != gTuned.end()) return it->second;
    }
    EventSet<4> ev;
    ev.record(0, s);
    launchRayShape<kSave>(kShapeLane, a, rays, count, out, hitPrim, nullptr, s);
    ev.record(1, s);
    ev.record(2, s);
    launchRayShape<kSave>(kShapeWave, a, rays, count, out, hitPrim, nullptr, s);
    ev.record(3, s);
    hipCheck(hipGetLastError(), "ray tuning");
    ev.sync(3);
    const int best = ev.ms(2, 3) < ev.ms(0, 1) ? kShapeWave : kShapeLane;
    std::lock_guard<std::mutex> lock(gTuneMutex);
    if (gTuned.size() >= 4096) gTuned.clear();
    gTuned[key] = best;
    return best;
}

template <bool kSave>
static void launchRaysT(const TraverseArgs& a, const Ray* rays, int count, uint64_t* out, int* hitPrim,
                        hipStream_t s) {
    const int grid = blocksFor(count, kBlock);
    if (grid == 0) return;
    const Tri* tris = static_cast<const Tri*>(a.items);
    if (a.wide) {
        launchRayShape<kSave>(rayShape<kSave>(a, rays, count, out, hitPrim, s), a, rays, count, out, hitPrim,
                              a.stats, s);
    } else {
        if (a.stats)
            rayBinaryKernel<true, kSave><<<grid, kBlock, 0, s>>>(a.nodes, tris, rays, a.order, count, out, a.stats,
                                                                a.primIds, hitPrim);
        else
            rayBinaryKernel<false, kSave><<<grid, kBlock, 0, s>>>(a.nodes, tris, rays, a.order, count, out,
                                                                 nullptr, a.primIds, hitPrim);
    }
    hipCheck(hipGetLastError(), "raycast");
}

void launchRays(const TraverseArgs& a, const Ray* rays, int count, uint64_t* out, hipStream_t s) {
    launchRaysT<false>(a, rays, count, out, nullptr, s);
}

void launchRaysSave(const TraverseArgs& a, const Ray* rays, int count, uint64_t* out, int* hitPrim,
                    hipStream_t s) {
    launchRaysT<true>(a, rays, count, out, hitPrim, s);
}

namespace {

}

void setTraversalMode(int mode) { gTraversalMode = mode; }

namespace {
constexpr int kOverflowDevices = 64;
std::mutex gOverflowMutex;
int* gOverflowHost[kOverflowDevices] = {};

int* overflowHost() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kOverflowDevices) return nullptr;
    std::lock_guard<std::mutex> lock(gOverflowMutex);
    return gOverflowHost[dev];
}
}  // namespace

void armOverflowWord() {
    int dev = 0;
    hipCheck(hipGetDevice(&dev), "get device");
    if (dev < 0 || dev >= kOverflowDevices) return;
    std::lock_guard<std::mutex> lock(gOverflowMutex);
    if (gOverflowHost[dev]) return;
    int* h = nullptr;
    hipCheck(hipHostMalloc(reinterpret_cast<void**>(&h), 256, hipHostMallocDefault), "overflow word");
    *h = 0;
    int* d = nullptr;
    hipCheck(hipHostGetDevicePointer(reinterpret_cast<void**>(&d), h, 0), "overflow word address");
    hipCheck(hipMemcpyToSymbol(HIP_SYMBOL(gOverflowWord), &d, sizeof(d)), "overflow word symbol");
    gOverflowHost[dev] = h;
}

void checkOverflow() {
    int* h = overflowHost();
    if (!h || __atomic_load_n(h, __ATOMIC_ACQUIRE) == 0) return;
    __atomic_store_n(h, 0, __ATOMIC_RELEASE);
    throw DeviceError("traversal stack overflow: the tree is deeper than the traversal stack holds, "
                      "so some results are incomplete");
}

// kNN (k > 1): four lanes per query at every size -- the quad's spread k-best list
// makes each offer a quarter of the per-lane work, and it measured faster from 20k
// to 1M queries on the 2M-triangle terrain.  Closest (k = 1): four lanes per query
// while the query set leaves most of the GPU idle one lane per query (fewer than 4
// waves per SIMD: 256 CUs x 4 SIMDs x 4 waves x 64 lanes).
// Radius capture one wave per query below this many queries, four lanes per query
// above (traversal mode 2 forces four lanes per query).
int gRadiusWaveBelow = 1 << 17;

// Closest / kNN one wave per query below a leaf- and k-dependent count (automatic
// mode, under kTuneFrom queries).  A small query set fills the device only with more
// lanes per query, while past that the quad kernel needs fewer instructions per
// query on short, even walks.
// Measured crossovers (kernel time against the quad kernel, same bits): the
// 2M-triangle terrain on a SAH tree -- closest about 200k queries, kNN16 about 150k
// (100k: closest 0.45 vs 0.57 ms, kNN16 0.72 vs 0.80); 1M points -- closest about
// 35k (20k: 0.08 vs 0.11 ms), kNN8 about 45k; 1M Gaussians -- closest about 45k,
// kNN8 beyond 80k (80k: 0.81 vs 1.08 ms).
template <typename L, int K>
static bool useWave(int count) {
    int below;
    if (L::kKind == 0) below = K > 1 ? 131072 : 196608;
    else if (L::kKind == 1) below = K > 1 ? 45056 : 32768;
    else below = K > 1 ? 131072 : 45056;
    return count < below;
}

// G lanes per query (64: one wave per query).
template <typename L, int K, int G>
static void launchGroup(const TraverseArgs& a, const V3* pts, int count, int k, int outStride, ClosestResult* out,
                        unsigned long long* stats, hipStream_t s) {
    const auto* prims = static_cast<const typename L::Item*>(a.items);
    const int idBits = stackIdBitsDev(a.wideCount);
    const int grid = blocksFor(count, (kBlock / 64) * (64 / G));
    if (stats)
        nearestWaveKernel<L, K, G, true><<<grid, kBlock, 0, s>>>(a.wideNodes, prims, a.primIds, pts, a.order, count, k, outStride, out, stats, idBits);
    else
        nearestWaveKernel<L, K, G, false><<<grid, kBlock, 0, s>>>(a.wideNodes, prims, a.primIds, pts, a.order, count, k, outStride, out, nullptr, idBits);
}

template <typename L, int K>
static void launchShape(int shape, const TraverseArgs& a, const V3* pts, int count, int k, int outStride,
                        ClosestResult* out, unsigned long long* stats, hipStream_t s) {
    const auto* prims = static_cast<const typename L::Item*>(a.items);
    const int idBits = stackIdBitsDev(a.wideCount);
    if (shape == kShapeWave) {
        launchGroup<L, K, 64>(a, pts, count, k, outStride, out, stats, s);
    } else if (shape == kShapeGroup16 || shape == kShapeGroup32) {
        if constexpr (K <= 16) {
            if (shape == kShapeGroup16)
                launchGroup<L, K, 16>(a, pts, count, k, outStride, out, stats, s);
            else
                launchGroup<L, K, 32>(a, pts, count, k, outStride, out, stats, s);
        } else {
            launchGroup<L, K, 32>(a, pts, count, k, outStride, out, stats, s);  // 32 entries need 32 lanes
        }
    } else if (shape == kShapeQuad) {
        const int grid = blocksFor(4LL * count, kBlock);
        if (stats)
            nearestQuadKernel<L, K, true><<<grid, kBlock, 0, s>>>(a.wideNodes, prims, a.primIds, pts, a.order, count, k, outStride, out, stats, idBits);
        else
            nearestQuadKernel<L, K, false><<<grid, kBlock, 0, s>>>(a.wideNodes, prims, a.primIds, pts, a.order, count, k, outStride, out, nullptr, idBits);
    } else {
        const int grid = blocksFor(count, kBlock);
        if (stats)
            nearestWideKernel<L, K, true><<<grid, kBlock, 0, s>>>(a.wideNodes, prims, a.primIds, pts, a.order, count, k, outStride, out, stats, idBits);
        else
            nearestWideKernel<L, K, false><<<grid, kBlock, 0, s>>>(a.wideNodes, prims, a.primIds, pts, a.order, count, k, outStride, out, nullptr, idBits);
    }
}

// The shape for a set of `count` queries by the measured crossovers (above).
template <typename L, int K>
static int defaultShape(int count) {
    if (useWave<L, K>(count)) return kShapeWave;
    return (K > 1 || count < 256 * 4 * 4 * 64) ? kShapeQuad : kShapeLane;
}

template <typename L, int K>
static int tunedShape(const TraverseArgs& a, const V3* pts, int count, int k, int outStride, ClosestResult* out,
                      hipStream_t s) {
    int octave = 0;
    while ((2LL << octave) <= count) ++octave;
    const TuneKey key{a.wideNodes, a.wideCount, L::kKind, K, octave};
    {
        std::lock_guard<std::mutex> lock(gTuneMutex);
        const auto it = gTuned.find(key);
        if (it != gTuned.end()) return it->second;
    }
    // every candidate shape once, after one untimed walk that brings the tree into
    // the caches (so the first timed shape is not the only one reading it cold)
    constexpr int kShapes = K <= 16 ? 5 : 4;
    const int shapes[5] = {kShapeLane, kShapeQuad, kShapeWave, kShapeGroup32, kShapeGroup16};
    EventSet<2 * 5> ev;
    launchShape<L, K>(kShapeWave, a, pts, count, k, outStride, out, nullptr, s);
    for (int i = 0; i < kShapes; ++i) {
        ev.record(2 * i, s);
        launchShape<L, K>(shapes[i], a, pts, count, k, outStride, out, nullptr, s);
        ev.record(2 * i + 1, s);
    }
    hipCheck(hipGetLastError(), "nearest tuning");
    ev.sync(2 * kShapes - 1);
    int best = kShapeLane;
    float bestMs = 0.0f;
    for (int i = 0; i < kShapes; ++i) {
        const float ms = ev.ms(2 * i, 2 * i + 1);
        if (i == 0 || ms < bestMs) {
            best = shapes[i];
            bestMs = ms;
        }
    }
    std::lock_guard<std::mutex> lock(gTuneMutex);
    if (gTuned.size() >= 4096) gTuned.clear();  // trees come and go; keep the table small
    gTuned[key] = best;
    return best;
}

// The shape an earlier tuning chose for this tree and size class, else the measured
// crossovers' (no timing trials).
template <typename L, int K>
static int cachedShape(const TraverseArgs& a, int count) {
    int octave = 0;
    while ((2LL << octave) <= count) ++octave;
    const TuneKey key{a.wideNodes, a.wideCount, L::kKind, K, octave};
    {
        std::lock_guard<std::mutex> lock(gTuneMutex);
        const auto it = gTuned.find(key);
        if (it != gTuned.end()) return it->second;
    }
    return defaultShape<L, K>(count);
}

template <typename L, int K>
static void launchNearestTK(const TraverseArgs& a, const V3* pts, int count, int k, int outStride,
                            ClosestResult* out, hipStream_t s) {
    const auto* prims = static_cast<const typename L::Item*>(a.items);
    if (!a.wide) {
        const int grid = blocksFor(count, kBlock);
        if (a.stats)
            nearestBinaryKernel<L, K, true><<<grid, kBlock, 0, s>>>(a.nodes, prims, a.primIds, pts, a.order, count, k, outStride, out, a.stats, stackIdBitsDev(a.nodeCount));
        else
            nearestBinaryKernel<L, K, false><<<grid, kBlock, 0, s>>>(a.nodes, prims, a.primIds, pts, a.order, count, k, outStride, out, nullptr, stackIdBitsDev(a.nodeCount));
        return;
    }
    int shape;
    switch (gTraversalMode) {
        case 1: shape = kShapeLane; break;
        case 2: shape = kShapeQuad; break;
        case 3: shape = kShapeWave; break;
        case 5: shape = kShapeGroup16; break;
        case 6: shape = kShapeGroup32; break;
        case 4:
            shape = a.tune ? tunedShape<L, K>(a, pts, count, k, outStride, out, s) : cachedShape<L, K>(a, count);
            break;
        default:
            shape = count < kTuneFrom ? defaultShape<L, K>(count)
                    : a.tune          ? tunedShape<L, K>(a, pts, count, k, outStride, out, s)
                                      : cachedShape<L, K>(a, count);
    }
    launchShape<L, K>(shape, a, pts, count, k, outStride, out, a.stats, s);
}

template <typename L>
static void launchNearestT(const TraverseArgs& a, const V3* pts, int count, int k, int outStride,
                           ClosestResult* out, hipStream_t s) {
    if (k == 1) {
        launchNearestTK<L, 1>(a, pts, count, k, outStride, out, s);
    } else if (k <= 4) {
        launchNearestTK<L, 4>(a, pts, count, k, outStride, out, s);
    } else if (k <= 8) {
        launchNearestTK<L, 8>(a, pts, count, k, outStride, out, s);
    } else if (k <= 16) {
        launchNearestTK<L, 16>(a, pts, count, k, outStride, out, s);
    } else {
        launchNearestTK<L, 32>(a, pts, count, k, outStride, out, s);
    }
}

static void launchNearest(const TraverseArgs& a, const V3* pts, int count, int k, int outStride,
                          ClosestResult* out, hipStream_t s) {
    if (count <= 0) return;
    switch (a.kind) {
        case 0: launchNearestT<TriLeaf>(a, pts, count, k, outStride, out, s); break;
        case 1: launchNearestT<PointLeaf>(a, pts, count, k, outStride, out, s); break;
        default: launchNearestT<GaussLeaf>(a, pts, count, k, outStride, out, s); break;
    }
    hipCheck(hipGetLastError(), k == 1 ? "closest" : "knn");
}

void launchClosest(const TraverseArgs& a, const V3* pts, int count, ClosestResult* out,
                   hipStream_t s) {
    launchNearest(a, pts, count, 1, 1, out, s);
}

void launchKnn(const TraverseArgs& a, const V3* pts, int count, int k, int outStride,
               ClosestResult* out, hipStream_t s) {
    launchNearest(a, pts, count, k, outStride, out, s);
}

// Radius on four lanes per query measured faster at every query count from 20k to
// 1M on the 2M-triangle terrain (the lanes of a quad test the four children of a
// node and its leaves together instead of walking them one after another); it
// serves 4-wide trees unless one lane per query is forced.
static bool radiusOnQuads(const TraverseArgs& a) { return a.wide && gTraversalMode != 1; }

// cap > 0: a fill on quads after a capture (only the queries the capture could not
// hold are walked); cap == 0: the per-lane count / fill kernels.
template <typename L, bool kFill>
static void launchRadiusT(const TraverseArgs& a, const V3* pts, int count, float r2, int* counts,
                          const int* offsets, uint64_t* keys, int* maxCount, int cap, hipStream_t s) {
    const int grid = blocksFor(count, kBlock);
    const auto* prims = static_cast<const typename L::Item*>(a.items);
    unsigned long long* st = kFill ? nullptr : a.stats;
    if (kFill && cap > 0) {
        const int qgrid = blocksFor(4LL * count, kBlock);
        radiusQuadKernel<L, kRadiusFill, false><<<qgrid, kBlock, 0, s>>>(a.wideNodes, prims, a.primIds, pts, a.order, count, r2, counts, offsets, keys, nullptr, maxCount, cap);
    } else if (a.wide) {
        if (st)
            radiusKernel<L, true, kFill, true><<<grid, kBlock, 0, s>>>(a.nodes, a.wideNodes, prims, a.primIds, pts, a.order, count, r2, counts, offsets, keys, st, maxCount);
        else
            radiusKernel<L, true, kFill, false><<<grid, kBlock, 0, s>>>(a.nodes, a.wideNodes, prims, a.primIds, pts, a.order, count, r2, counts, offsets, keys, nullptr, maxCount);
    } else {
        if (st)
            radiusKernel<L, false, kFill, true><<<grid, kBlock, 0, s>>>(a.nodes, a.wideNodes, prims, a.primIds, pts, a.order, count, r2, counts, offsets, keys, st, maxCount);
        else
            radiusKernel<L, false, kFill, false><<<grid, kBlock, 0, s>>>(a.nodes, a.wideNodes, prims, a.primIds, pts, a.order, count, r2, counts, offsets, keys, nullptr, maxCount);
    }
}

template <typename L>
static void launchCaptureT(const TraverseArgs& a, const V3* pts, int count, float r2, int* counts,
                           uint64_t* capture, int cap, int* maxCount, hipStream_t s) {
    const int qgrid = blocksFor(4LL * count, kBlock);
    const auto* prims = static_cast<const typename L::Item*>(a.items);
    // below gRadiusWaveBelow two queries per wave (measured ahead of one per wave at
    // every count tried: terrain r = 2 at 20k-115k queries 0.045-0.127 ms against
    // 0.049-0.161, the 1M-point cloud r = 1 at 20k / 65k 0.106 / 0.605 against
    // 0.124 / 0.745; its segments of up to 64 hits leave sorted, so the quads' segment
    // sort is not needed either); mode 3 forces one wave per query
    if ((count < gRadiusWaveBelow && (gTraversalMode == 0 || gTraversalMode == 4 || gTraversalMode == 5)) ||
        gTraversalMode == 6) {
        const int wgrid = blocksFor(count, 2 * (kBlock / 64));
        if (a.stats)
            radiusWaveKernel<L, true, 32><<<wgrid, kBlock, 0, s>>>(a.wideNodes, prims, a.primIds, pts, a.order, count, r2, counts, capture, a.stats, maxCount, cap);
        else
            radiusWaveKernel<L, false, 32><<<wgrid, kBlock, 0, s>>>(a.wideNodes, prims, a.primIds, pts, a.order, count, r2, counts, capture, nullptr, maxCount, cap);
    } else if (gTraversalMode == 3) {
        const int wgrid = blocksFor(count, kBlock / 64);
        if (a.stats)
            radiusWaveKernel<L, true, 64><<<wgrid, kBlock, 0, s>>>(a.wideNodes, prims, a.primIds, pts, a.order, count, r2, counts, capture, a.stats, maxCount, cap);
        else
            radiusWaveKernel<L, false, 64><<<wgrid, kBlock, 0, s>>>(a.wideNodes, prims, a.primIds, pts, a.order, count, r2, counts, capture, nullptr, maxCount, cap);
    } else if (a.stats)
        radiusQuadKernel<L, kRadiusCaptureMode, true><<<qgrid, kBlock, 0, s>>>(a.wideNodes, prims, a.primIds, pts, a.order, count, r2, counts, nullptr, capture, a.stats, maxCount, cap);
    else
        radiusQuadKernel<L, kRadiusCaptureMode, false><<<qgrid, kBlock, 0, s>>>(a.wideNodes, prims, a.primIds, pts, a.order, count, r2, counts, nullptr, capture, nullptr, maxCount, cap);
}

// Hits captured per query by the single walk; a second walk serves only the queries
// with more.  Up to 2,048 per query while the buffer stays within 256 MB (20k queries:
// 1,638, which holds every segment of the 1M-point cloud at r = 1, at most 572), at
// least 64 while 512 MB hold them (the 2M-triangle terrain at r = 2 has at most 43);
// fewer than 8 turns the capture off.
int radiusCaptureCap(int count) {
    if (count <= 0) return 64;
    const long long roomy = (256LL << 20) / static_cast<long long>(sizeof(uint64_t)) / count;
    if (roomy >= 64) return roomy > 2048 ? 2048 : static_cast<int>(roomy);
    const long long tight = (512LL << 20) / static_cast<long long>(sizeof(uint64_t)) / count;
    if (tight >= 64) return 64;
    return tight < 8 ? 0 : static_cast<int>(tight);
}

bool radiusCapturePresorts(const TraverseArgs& a, int count) {
    return radiusOnQuads(a) &&
           ((count < gRadiusWaveBelow && (gTraversalMode == 0 || gTraversalMode == 4 || gTraversalMode == 5)) ||
            gTraversalMode == 3 || gTraversalMode == 6);
}

bool launchRadiusCapture(const TraverseArgs& a, const V3* pts, int count, float r2, int* counts,
                         uint64_t* capture, int cap, int* maxCount, hipStream_t s) {
    if (count <= 0) return false;
    if (!radiusOnQuads(a) || cap <= 0) {
        switch (a.kind) {
            case 0: launchRadiusT<TriLeaf, false>(a, pts, count, r2, counts, nullptr, nullptr, maxCount, 0, s); break;
            case 1: launchRadiusT<PointLeaf, false>(a, pts, count, r2, counts, nullptr, nullptr, maxCount, 0, s); break;
            default: launchRadiusT<GaussLeaf, false>(a, pts, count, r2, counts, nullptr, nullptr, maxCount, 0, s); break;
        }
        hipCheck(hipGetLastError(), "radius count");
        return false;
    }
    switch (a.kind) {
        case 0: launchCaptureT<TriLeaf>(a, pts, count, r2, counts, capture, cap, maxCount, s); break;
        case 1: launchCaptureT<PointLeaf>(a, pts, count, r2, counts, capture, cap, maxCount, s); break;
        default: launchCaptureT<GaussLeaf>(a, pts, count, r2, counts, capture, cap, maxCount, s); break;
    }
    hipCheck(hipGetLastError(), "radius capture");
    return true;
}

void launchRadiusFill(const TraverseArgs& a, const V3* pts, int count, float r2, const int* counts,
                      int cap, const int* offsets, uint64_t* keys, hipStream_t s) {
    if (count <= 0) return;
    int* c = const_cast<int*>(counts);  // read only by the fill pass
    switch (a.kind) {
        case 0: launchRadiusT<TriLeaf, true>(a, pts, count, r2, c, offsets, keys, nullptr, cap, s); break;
        case 1: launchRadiusT<PointLeaf, true>(a, pts, count, r2, c, offsets, keys, nullptr, cap, s); break;
        default: launchRadiusT<GaussLeaf, true>(a, pts, count, r2, c, offsets, keys, nullptr, cap, s); break;
    }
    hipCheck(hipGetLastError(), "radius fill");
}

namespace {
// One wave per query copies its captured hits (queries with more than `cap` hits are
// left to the fill walk); a segment that would end past `capacity` keys is skipped
// (the caller compacts again into a buffer of the exact size).
__global__ void radiusCompactKernel(const int* __restrict__ counts, const int* __restrict__ offsets,
                                    const uint64_t* __restrict__ capture, int count, int cap,
                                    uint64_t* __restrict__ keys, long long capacity) {
    const long long q = static_cast<long long>(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
    if (q >= count) return;
    const int c = counts[q];
    if (c > cap) return;
    const long long off = offsets[q];
    if (off + c > capacity) return;
    const uint64_t* src = capture + q * cap;
    uint64_t* dst = keys + off;
    for (int j = threadIdx.x & 63; j < c; j += 64) dst[j] = src[j];
}

// The radius call's two host-side figures in one step: the total hit count and the
// largest per-query count (the maximum of the kMaxCopies words), stored straight into
// the caller's pinned host words.
__global__ void radiusSummaryKernel(const int* __restrict__ total, const int* __restrict__ maxCopies,
                                    int* __restrict__ host) {
    const int lane = threadIdx.x;
    int m = lane < kMaxCopies ? maxCopies[lane] : 0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    if (lane == 0) {
        __hip_atomic_store(&host[0], *total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&host[1], m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}
}  // namespace

void radiusCompact(const int* counts, const int* offsets, const uint64_t* capture, int count, int cap,
                   uint64_t* keys, long long capacity, hipStream_t s) {
    if (count <= 0 || cap <= 0) return;
    radiusCompactKernel<<<blocksFor(count, kBlock / 64), kBlock, 0, s>>>(counts, offsets, capture, count, cap,
                                                                        keys, capacity);
    hipCheck(hipGetLastError(), "radius compact");
}

void radiusSummary(const int* total, const int* maxCopies, int* hostDev, hipStream_t s) {
    static_assert(kMaxCopies <= 64, "one lane per copy");
    radiusSummaryKernel<<<1, 64, 0, s>>>(total, maxCopies, hostDev);
    hipCheck(hipGetLastError(), "radius summary");
}

void segmentSort(const int* offsets, uint64_t* keys, int count, int maxSeg, int* scratchInts,
                 hipStream_t s) {
    if (count <= 0 || maxSeg <= 1) return;
    const int wavesPerBlock = kBlock / 64;
    segSortWaveKernel<<<blocksFor(count, wavesPerBlock), kBlock, 0, s>>>(offsets, keys, count);
    hipCheck(hipGetLastError(), "segment sort (wave)");
    if (maxSeg <= kWaveCap) return;
    // scratchInts: [info 4][blockList count][hugeList count]
    int* info = scratchInts;
    int* blockList = scratchInts + 4;
    int* hugeList = blockList + count;
    zeroU64Kernel<<<1, kBlock, 0, s>>>(reinterpret_cast<unsigned long long*>(info), 2);
    segClassifyKernel<<<blocksFor(count, kBlock), kBlock, 0, s>>>(offsets, count, blockList,
                                                                  hugeList, info);
    // the block-sized segments: a fixed grid walks the device-side list (no host round trip)
    const int grid = count < 1024 ? count : 1024;
    segSortBlockKernel<<<grid, kBlock, kBlockCap * sizeof(uint64_t), s>>>(offsets, blockList, info, keys);
    hipCheck(hipGetLastError(), "segment sort (block)");
    if (maxSeg <= kBlockCap) return;
    int h[2] = {0, 0};
    hipCheck(hipMemcpyAsync(h, info, sizeof(h), hipMemcpyDeviceToHost, s), "seg info");
    hipCheck(hipStreamSynchronize(s), "seg info sync");
    if (h[1] > 0) {
        // Segments above the workgroup network's capacity: each is sorted by the
        // device radix sort as two stable 32-bit passes (index half, then distance
        // half), which orders the 64-bit keys exactly.
        std::vector<int> huge(static_cast<size_t>(h[1]));
        hipCheck(hipMemcpyAsync(huge.data(), hugeList, sizeof(int) * huge.size(), hipMemcpyDeviceToHost, s),
                 "huge list");
        std::vector<int> offs(static_cast<size_t>(count) + 1);
        hipCheck(hipMemcpyAsync(offs.data(), offsets, sizeof(int) * offs.size(), hipMemcpyDeviceToHost, s),
                 "huge offsets");
        hipCheck(hipStreamSynchronize(s), "huge sync");
        int maxM = 0;
        for (int sg : huge) maxM = std::max(maxM, offs[static_cast<size_t>(sg) + 1] - offs[static_cast<size_t>(sg)]);
        SortWorkspace ws;
        ws.reserve(static_cast<size_t>(maxM));
        DevBuf<uint32_t> lo(maxM), hi(maxM), tmp(maxM), ord1(maxM), ord2(maxM);
        for (int sg : huge) {
            const int off = offs[static_cast<size_t>(sg)];
            const int m = offs[static_cast<size_t>(sg) + 1] - off;
            splitKeysKernel<<<blocksFor(m, kBlock), kBlock, 0, s>>>(keys + off, m, lo.p, hi.p);
            sortClear(ws, static_cast<size_t>(m), 4, s);
            sortHistogram(ws, lo.p, static_cast<size_t>(m), 4, s);
            sortPairs(ws, lo.p, nullptr, tmp.p, ord1.p, static_cast<size_t>(m), 4, s);
            gatherU32Kernel<<<blocksFor(m, kBlock), kBlock, 0, s>>>(hi.p, ord1.p, m, tmp.p);
            sortClear(ws, static_cast<size_t>(m), 4, s);
            sortHistogram(ws, tmp.p, static_cast<size_t>(m), 4, s);
            sortPairs(ws, tmp.p, ord1.p, hi.p, ord2.p, static_cast<size_t>(m), 4, s);
            joinKeysKernel<<<blocksFor(m, kBlock), kBlock, 0, s>>>(hi.p, lo.p, ord2.p, m, keys + off);
            hipCheck(hipGetLastError(), "segment sort (huge)");
        }
        hipCheck(hipStreamSynchronize(s), "huge sort sync");
    }
}


// ---- batched traversal launchers ---------------------------------------------------------

namespace {

int residentCus() {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        hipCheck(hipGetDevice(&dev), "get device");
        hipDeviceProp_t prop{};
        hipCheck(hipGetDeviceProperties(&prop, dev), "device properties");
        cus = prop.multiProcessorCount;
    }
    return cus;
}

template <typename L, int K, int kBT, int kD, bool kWide, bool kReg>
void launchBatchTKV(const BatchQueryArgs& a, hipStream_t s) {
    auto kern = batchNearestKernel<L, K, kBT, kD, kWide, kReg>;
    static int perCu = 0;
    if (perCu == 0) {
        hipCheck(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, reinterpret_cast<const void*>(kern), kBT, 0),
                 "occupancy");
        if (perCu < 1) perCu = 1;
    }
    const int waves = (a.count + 63) / 64;
    int grid = perCu * residentCus();
    const int need = (waves * 64 + kBT - 1) / kBT;
    if (grid > need) grid = need;
    if (grid < 1) grid = 1;
    kern<<<grid, kBT, 0, s>>>(a.nodes, a.wide, a.wideRoot, static_cast<const typename L::Item*>(a.items), a.primIds,
                              a.primOff, a.points, a.qScene, a.order, a.count, a.k, a.outStride, a.out, a.counter);
}

template <typename L, int K>
void launchBatchTK(const BatchQueryArgs& a, int variant, hipStream_t s) {
    switch (variant) {
        case kBatchLds256: launchBatchTKV<L, K, 256, kLdsDepth, false, false>(a, s); break;
        case kBatchLds64: launchBatchTKV<L, K, 64, kLdsDepth, false, false>(a, s); break;
        case kBatchScratch64: launchBatchTKV<L, K, 64, 0, false, false>(a, s); break;
        case kBatchReg64: launchBatchTKV<L, K, 64, 0, false, true>(a, s); break;
        case kBatchWideLds256: launchBatchTKV<L, K, 256, kLdsDepth, true, false>(a, s); break;
        case kBatchWideLds64: launchBatchTKV<L, K, 64, kLdsDepth, true, false>(a, s); break;
        case kBatchWideScratch64: launchBatchTKV<L, K, 64, 0, true, false>(a, s); break;
        default: launchBatchTKV<L, K, 64, 0, true, true>(a, s); break;
    }
}

template <typename L>
void launchBatchT(const BatchQueryArgs& a, int variant, hipStream_t s) {
    if (a.k <= 1) {
        launchBatchTK<L, 1>(a, variant, s);
    } else if (a.k <= 8) {
        launchBatchTK<L, 8>(a, variant, s);
    } else {
        launchBatchTK<L, 32>(a, variant, s);
    }
}

}  // namespace

void launchBatchNearest(const BatchQueryArgs& a, int variant, hipStream_t s) {
    if (a.count <= 0) return;
    if (a.k > 32) throw ArgError("k must be in [1, 32]");
    switch (a.kind) {
        case 0: launchBatchT<TriLeaf>(a, variant, s); break;
        case 1: launchBatchT<PointLeaf>(a, variant, s); break;
        default: launchBatchT<GaussLeaf>(a, variant, s); break;
    }
    hipCheck(hipGetLastError(), "batch traversal");
}

void launchBatchReset(int* counter, hipStream_t s) {
    batchResetKernel<<<1, 64, 0, s>>>(counter);
    hipCheck(hipGetLastError(), "batch reset");
}

}  // namespace qr
