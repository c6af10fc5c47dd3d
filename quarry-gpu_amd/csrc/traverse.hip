// Traversal kernels: closest-hit rays, closest / kNN / radius point queries, over
// binary (32 B) or 4-wide (128 B) nodes, for triangle, point and Gaussian leaves.
//
// One lane per ray or query, depth-first with an explicit per-lane stack.  The
// pruning predicates are the reference's (src/gpu/raycast.cu:75-245,
// src/gpu/query.cu:52-637): padded slab test against the current best t for
// rays, box lower bound <= best for closest/kNN (so equal-distance primitives
// with a smaller index are still reached), box lower bound <= r^2 for radius.
// Because every predicate only ever prunes subtrees that cannot hold a better
// (dist2, index) or a smaller t, the result is the brute-force answer whatever
// the tree, node width or visiting order - which is what makes binary, wide,
// LBVH, SAH and reordered runs bit-identical.
//
// Queries may be visited in a permuted order (Morton reordering): lane i works
// on query order[i] and writes its result straight to slot order[i], so no
// gather/scatter copies of queries or results are made.
#include "qr_traverse.h"

namespace qr {

namespace {

constexpr int kBlock = 256;
constexpr int kStack = 72;       // binary: one net push per level, depth <= 64 for LBVH
constexpr int kStackWide = 144;  // 4-wide: at most 3 net pushes per wide level

// ---- stats -------------------------------------------------------------------------
// stats[0] node tests, [1] leaf tests, [2] stack overflows.
template <bool kStats>
__device__ __forceinline__ void flushStats(unsigned long long* stats, int nodeTests, int leafTests,
                                           int overflow) {
    if (!kStats) return;
    unsigned long long a = static_cast<unsigned long long>(nodeTests);
    unsigned long long b = static_cast<unsigned long long>(leafTests);
    unsigned long long c = static_cast<unsigned long long>(overflow);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
        c += __shfl_xor(c, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (a) atomicAdd(&stats[0], a);
        if (b) atomicAdd(&stats[1], b);
        if (c) atomicAdd(&stats[2], c);
    }
}

// ---- rays --------------------------------------------------------------------------

template <bool kStats>
__global__ __launch_bounds__(kBlock) void rayBinaryKernel(const Node2* __restrict__ nodes,
                                                          const Tri* __restrict__ tris,
                                                          const Ray* __restrict__ rays,
                                                          const uint32_t* __restrict__ order,
                                                          int count, uint64_t* __restrict__ out,
                                                          unsigned long long* __restrict__ stats) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    int nodeTests = 0, leafTests = 0, overflow = 0;
    if (i < count) {
        const int q = order ? static_cast<int>(order[i]) : i;
        const Ray ray = rays[q];
        float best = ray.tMax;
        bool any = false;
        int stack[kStack];
        int sp = 0;
        float tEnter;
        if (rayBox(ray, nodeBox(nodes[0]), best, tEnter)) {
            ++nodeTests;
            stack[sp++] = 0;
            while (sp > 0) {
                const int id = stack[--sp];
                const Node2 nd = nodes[id];
                if (nd.left < 0) {
                    ++leafTests;
                    float t;
                    if (rayTri(ray, tris[-nd.left - 1], t) && t < best) {
                        best = t;
                        any = true;
                    }
                    continue;
                }
                float tl, tr;
                const bool hl = rayBox(ray, nodeBox(nodes[nd.left]), best, tl);
                const bool hr = rayBox(ray, nodeBox(nodes[nd.right]), best, tr);
                nodeTests += static_cast<int>(hl) + static_cast<int>(hr);
                if (sp + 2 > kStack) {
                    overflow = 1;
                    break;
                }
                if (hl && hr) {
                    const bool leftFirst = tl <= tr;
                    stack[sp++] = leftFirst ? nd.right : nd.left;
                    stack[sp++] = leftFirst ? nd.left : nd.right;
                } else if (hl) {
                    stack[sp++] = nd.left;
                } else if (hr) {
                    stack[sp++] = nd.right;
                }
            }
        }
        out[q] = any ? ((uint64_t{1} << 32) | floatBits(best)) : 0ull;
    }
    flushStats<kStats>(stats, nodeTests, leafTests, overflow);
}

template <bool kStats>
__global__ __launch_bounds__(kBlock) void rayWideKernel(const Wide4* __restrict__ wide,
                                                        const Tri* __restrict__ tris,
                                                        const Ray* __restrict__ rays,
                                                        const uint32_t* __restrict__ order,
                                                        int count, uint64_t* __restrict__ out,
                                                        unsigned long long* __restrict__ stats) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    int nodeTests = 0, leafTests = 0, overflow = 0;
    if (i < count) {
        const int q = order ? static_cast<int>(order[i]) : i;
        const Ray ray = rays[q];
        float best = ray.tMax;
        bool any = false;
        int stack[kStackWide];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const int id = stack[--sp];
            if (id < 0) {
                ++leafTests;
                float t;
                if (rayTri(ray, tris[-id - 1], t) && t < best) {
                    best = t;
                    any = true;
                }
                continue;
            }
            const Wide4& wn = wide[id];
            int child[4];
            float key[4];
            int hits = 0;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const WideEntry e = wn.c[s];
                if (e.child == kWideEmpty) continue;
                float t;
                if (rayBox(ray, entryBox(e), best, t)) {
                    ++nodeTests;
                    int j = hits++;
                    while (j > 0 && key[j - 1] > t) {
                        key[j] = key[j - 1];
                        child[j] = child[j - 1];
                        --j;
                    }
                    key[j] = t;
                    child[j] = e.child;
                }
            }
            if (sp + hits > kStackWide) {
                overflow = 1;
                break;
            }
            for (int j = hits - 1; j >= 0; --j) stack[sp++] = child[j];
        }
        out[q] = any ? ((uint64_t{1} << 32) | floatBits(best)) : 0ull;
    }
    flushStats<kStats>(stats, nodeTests, leafTests, overflow);
}

// ---- closest -------------------------------------------------------------------------

__device__ __forceinline__ bool closer(float d, int id, float bd, int bid) {
    return d < bd || (d == bd && id < bid);
}

template <typename L, bool kStats>
__global__ __launch_bounds__(kBlock) void closestBinaryKernel(
    const Node2* __restrict__ nodes, const typename L::Item* __restrict__ prims,
    const int* __restrict__ primIds, const V3* __restrict__ points,
    const uint32_t* __restrict__ order, int count, ClosestResult* __restrict__ out,
    unsigned long long* __restrict__ stats) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    int nodeTests = 0, leafTests = 0, overflow = 0;
    if (i < count) {
        const int q = order ? static_cast<int>(order[i]) : i;
        const V3 p = points[q];
        float best2 = kInf;
        int bestPrim = -1;
        int stack[kStack];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const int id = stack[--sp];
            const Node2 nd = nodes[id];
            ++nodeTests;
            if (nd.left < 0) {
                const int slot = -nd.left - 1;
                ++leafTests;
                const float d2 = L::dist2(p, prims[slot]);
                const int prim = primIds[slot];
                if (bestPrim < 0 || closer(d2, prim, best2, bestPrim)) {
                    best2 = d2;
                    bestPrim = prim;
                }
                continue;
            }
            const float dl = boxDist2(p, nodeBox(nodes[nd.left]));
            const float dr = boxDist2(p, nodeBox(nodes[nd.right]));
            const bool okL = bestPrim < 0 || dl <= best2;
            const bool okR = bestPrim < 0 || dr <= best2;
            if (sp + 2 > kStack) {
                overflow = 1;
                break;
            }
            if (okL && okR) {
                const bool leftFirst = dl <= dr;
                stack[sp++] = leftFirst ? nd.right : nd.left;
                stack[sp++] = leftFirst ? nd.left : nd.right;
            } else if (okL) {
                stack[sp++] = nd.left;
            } else if (okR) {
                stack[sp++] = nd.right;
            }
        }
        ClosestResult r{0u, -1};
        if (bestPrim >= 0) r = {floatBits(best2), bestPrim};
        out[q] = r;
    }
    flushStats<kStats>(stats, nodeTests, leafTests, overflow);
}

template <typename L, bool kStats>
__global__ __launch_bounds__(kBlock) void closestWideKernel(
    const Wide4* __restrict__ wide, const typename L::Item* __restrict__ prims,
    const int* __restrict__ primIds, const V3* __restrict__ points,
    const uint32_t* __restrict__ order, int count, ClosestResult* __restrict__ out,
    unsigned long long* __restrict__ stats) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    int nodeTests = 0, leafTests = 0, overflow = 0;
    if (i < count) {
        const int q = order ? static_cast<int>(order[i]) : i;
        const V3 p = points[q];
        float best2 = kInf;
        int bestPrim = -1;
        int stack[kStackWide];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const int id = stack[--sp];
            if (id < 0) {
                const int slot = -id - 1;
                ++leafTests;
                const float d2 = L::dist2(p, prims[slot]);
                const int prim = primIds[slot];
                if (bestPrim < 0 || closer(d2, prim, best2, bestPrim)) {
                    best2 = d2;
                    bestPrim = prim;
                }
                continue;
            }
            const Wide4& wn = wide[id];
            int child[4];
            float key[4];
            int hits = 0;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const WideEntry e = wn.c[s];
                if (e.child == kWideEmpty) continue;
                const float d = boxDist2(p, entryBox(e));
                if (bestPrim >= 0 && d > best2) continue;
                ++nodeTests;
                int j = hits++;
                while (j > 0 && key[j - 1] > d) {
                    key[j] = key[j - 1];
                    child[j] = child[j - 1];
                    --j;
                }
                key[j] = d;
                child[j] = e.child;
            }
            if (sp + hits > kStackWide) {
                overflow = 1;
                break;
            }
            for (int j = hits - 1; j >= 0; --j) stack[sp++] = child[j];
        }
        ClosestResult r{0u, -1};
        if (bestPrim >= 0) r = {floatBits(best2), bestPrim};
        out[q] = r;
    }
    flushStats<kStats>(stats, nodeTests, leafTests, overflow);
}

// ---- kNN --------------------------------------------------------------------------
// The k best so far live in registers, sorted by (dist2, index).  A candidate is
// inserted by one ascending compare-exchange sweep with static indices (no
// dynamically indexed private array); the displaced worst drops off the end.
// Semantics equal the reference's sorted table (src/gpu/query.cu:150-158).
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
    // Worst kept entry when full (the pruning bound).
    __device__ __forceinline__ bool full(int k) const { return n >= k; }
    __device__ __forceinline__ float worstD(int k) const {
        float w = d[0];
#pragma unroll
        for (int s = 1; s < K; ++s)
            if (s == k - 1) w = d[s];
        return w;
    }
    __device__ __forceinline__ int worstId(int k) const {
        int w = id[0];
#pragma unroll
        for (int s = 1; s < K; ++s)
            if (s == k - 1) w = id[s];
        return w;
    }
    __device__ __forceinline__ void insert(float nd, int nid, int k) {
        float cd = nd;
        int cid = nid;
#pragma unroll
        for (int s = 0; s < K; ++s) {
            if (s < k) {
                const bool sw = closer(cd, cid, d[s], id[s]);
                const float td = d[s];
                const int tid = id[s];
                d[s] = sw ? cd : td;
                id[s] = sw ? cid : tid;
                cd = sw ? td : cd;
                cid = sw ? tid : cid;
            }
        }
        n = n < k ? n + 1 : k;
    }
};

template <typename L, int K, bool kStats>
__global__ __launch_bounds__(kBlock) void knnBinaryKernel(
    const Node2* __restrict__ nodes, const typename L::Item* __restrict__ prims,
    const int* __restrict__ primIds, const V3* __restrict__ points,
    const uint32_t* __restrict__ order, int count, int k, int outStride,
    ClosestResult* __restrict__ out, unsigned long long* __restrict__ stats) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    int nodeTests = 0, leafTests = 0, overflow = 0;
    if (i < count) {
        const int q = order ? static_cast<int>(order[i]) : i;
        const V3 p = points[q];
        KnnList<K> list;
        list.init();
        int stack[kStack];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const int id = stack[--sp];
            const Node2 nd = nodes[id];
            ++nodeTests;
            if (nd.left < 0) {
                const int slot = -nd.left - 1;
                ++leafTests;
                const float d2 = L::dist2(p, prims[slot]);
                const int prim = primIds[slot];
                if (list.full(k) && !closer(d2, prim, list.worstD(k), list.worstId(k))) continue;
                list.insert(d2, prim, k);
                continue;
            }
            const float dl = boxDist2(p, nodeBox(nodes[nd.left]));
            const float dr = boxDist2(p, nodeBox(nodes[nd.right]));
            const bool isFull = list.full(k);
            const float bound = list.worstD(k);
            const bool okL = !isFull || dl <= bound;
            const bool okR = !isFull || dr <= bound;
            if (sp + 2 > kStack) {
                overflow = 1;
                break;
            }
            if (okL && okR) {
                const bool leftFirst = dl <= dr;
                stack[sp++] = leftFirst ? nd.right : nd.left;
                stack[sp++] = leftFirst ? nd.left : nd.right;
            } else if (okL) {
                stack[sp++] = nd.left;
            } else if (okR) {
                stack[sp++] = nd.right;
            }
        }
        ClosestResult* dst = out + static_cast<size_t>(q) * outStride;
#pragma unroll
        for (int s = 0; s < K; ++s) {
            if (s < k) {
                ClosestResult r{0u, -1};
                if (s < list.n) r = {floatBits(list.d[s]), list.id[s]};
                dst[s] = r;
            }
        }
    }
    flushStats<kStats>(stats, nodeTests, leafTests, overflow);
}

template <typename L, int K, bool kStats>
__global__ __launch_bounds__(kBlock) void knnWideKernel(
    const Wide4* __restrict__ wide, const typename L::Item* __restrict__ prims,
    const int* __restrict__ primIds, const V3* __restrict__ points,
    const uint32_t* __restrict__ order, int count, int k, int outStride,
    ClosestResult* __restrict__ out, unsigned long long* __restrict__ stats) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    int nodeTests = 0, leafTests = 0, overflow = 0;
    if (i < count) {
        const int q = order ? static_cast<int>(order[i]) : i;
        const V3 p = points[q];
        KnnList<K> list;
        list.init();
        int stack[kStackWide];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const int id = stack[--sp];
            if (id < 0) {
                const int slot = -id - 1;
                ++leafTests;
                const float d2 = L::dist2(p, prims[slot]);
                const int prim = primIds[slot];
                if (list.full(k) && !closer(d2, prim, list.worstD(k), list.worstId(k))) continue;
                list.insert(d2, prim, k);
                continue;
            }
            const Wide4& wn = wide[id];
            int child[4];
            float key[4];
            int hits = 0;
            const bool isFull = list.full(k);
            const float bound = list.worstD(k);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const WideEntry e = wn.c[s];
                if (e.child == kWideEmpty) continue;
                const float d = boxDist2(p, entryBox(e));
                if (isFull && d > bound) continue;
                ++nodeTests;
                int j = hits++;
                while (j > 0 && key[j - 1] > d) {
                    key[j] = key[j - 1];
                    child[j] = child[j - 1];
                    --j;
                }
                key[j] = d;
                child[j] = e.child;
            }
            if (sp + hits > kStackWide) {
                overflow = 1;
                break;
            }
            for (int j = hits - 1; j >= 0; --j) stack[sp++] = child[j];
        }
        ClosestResult* dst = out + static_cast<size_t>(q) * outStride;
#pragma unroll
        for (int s = 0; s < K; ++s) {
            if (s < k) {
                ClosestResult r{0u, -1};
                if (s < list.n) r = {floatBits(list.d[s]), list.id[s]};
                dst[s] = r;
            }
        }
    }
    flushStats<kStats>(stats, nodeTests, leafTests, overflow);
}

// ---- radius: count and fill passes ---------------------------------------------------
// kFill = false counts matches per query; kFill = true writes the 64-bit keys
// (dist2 bits << 32 | index) at the query's CSR offset.  Both passes visit the same
// nodes in the same order, so the fill never writes past its segment.
template <typename L, bool kWide, bool kFill, bool kStats>
__global__ __launch_bounds__(kBlock) void radiusKernel(
    const Node2* __restrict__ nodes, const Wide4* __restrict__ wide,
    const typename L::Item* __restrict__ prims, const int* __restrict__ primIds,
    const V3* __restrict__ points, const uint32_t* __restrict__ order, int count, float r2,
    int* __restrict__ counts, const int* __restrict__ offsets, uint64_t* __restrict__ keys,
    unsigned long long* __restrict__ stats) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    int nodeTests = 0, leafTests = 0, overflow = 0;
    if (i < count) {
        const int q = order ? static_cast<int>(order[i]) : i;
        const V3 p = points[q];
        int found = 0;
        int write = kFill ? offsets[q] : 0;
        const int end = kFill ? offsets[q + 1] : 0;
        int stack[kWide ? kStackWide : kStack];
        constexpr int cap = kWide ? kStackWide : kStack;
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const int id = stack[--sp];
            if (!kWide) {
                const Node2 nd = nodes[id];
                if (boxDist2(p, nodeBox(nd)) > r2) continue;
                ++nodeTests;
                if (nd.left < 0) {
                    const int slot = -nd.left - 1;
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
                    continue;
                }
                if (sp + 2 > cap) {
                    overflow = 1;
                    break;
                }
                stack[sp++] = nd.right;
                stack[sp++] = nd.left;
            } else {
                if (id < 0) {
                    const int slot = -id - 1;
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
                    continue;
                }
                const Wide4& wn = wide[id];
                int push[4];
                int hits = 0;
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const WideEntry e = wn.c[s];
                    if (e.child == kWideEmpty) continue;
                    if (boxDist2(p, entryBox(e)) > r2) continue;
                    ++nodeTests;
                    push[hits++] = e.child;
                }
                if (sp + hits > cap) {
                    overflow = 1;
                    break;
                }
                for (int j = hits - 1; j >= 0; --j) stack[sp++] = push[j];
            }
        }
        if (!kFill) counts[q] = overflow ? 0 : found;
    }
    flushStats<kStats>(stats, nodeTests, leafTests, overflow);
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

__global__ __launch_bounds__(kBlock) void segSortBlockKernel(const int* __restrict__ offsets,
                                                             const int* __restrict__ segList,
                                                             uint64_t* __restrict__ keys) {
    extern __shared__ uint64_t dyn[];
    const int seg = segList[blockIdx.x];
    const int begin = offsets[seg];
    const int m = offsets[seg + 1] - begin;
    int cap = 2;
    while (cap < m) cap <<= 1;
    for (int i = threadIdx.x; i < cap; i += kBlock) dyn[i] = i < m ? keys[begin + i] : ~0ull;
    __syncthreads();
    bitonicLds(dyn, cap, threadIdx.x, kBlock, true);
    for (int i = threadIdx.x; i < m; i += kBlock) keys[begin + i] = dyn[i];
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
                                 const uint32_t* __restrict__ bb, int fixedBox,
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
        const uint32_t code = spread10b(quant10(p[0], lo[0], hi[0])) |
                              (spread10b(quant10(p[1], lo[1], hi[1])) << 1) |
                              (spread10b(quant10(p[2], lo[2], hi[2])) << 2);
        keys[i] = code;
        atomicAdd(&h[0][code & 255u], 1u);
        atomicAdd(&h[1][(code >> 8) & 255u], 1u);
        atomicAdd(&h[2][(code >> 16) & 255u], 1u);
        atomicAdd(&h[3][code >> 24], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 4 * 256; i += kBlock) {
        const uint32_t v = (&h[0][0])[i];
        if (v) atomicAdd(&hist[i], v);
    }
}

__global__ void zeroU64Kernel(unsigned long long* __restrict__ p, int n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

}  // namespace

// ---------------------------------------------------------------------------------------
// Host launchers.

void zeroStats(unsigned long long* stats, hipStream_t s) {
    zeroU64Kernel<<<1, kBlock, 0, s>>>(stats, 4);
}

void mortonOrder(const float* base, int stride, int count, int fixedBox, SortWorkspace& ws,
                 uint32_t* keys, uint32_t* keysSorted, uint32_t* orderOut, uint32_t* bb,
                 hipStream_t s) {
    ws.reserve(static_cast<size_t>(count));
    sortClear(ws, static_cast<size_t>(count), 4, s);
    int grid = blocksFor(count, kBlock * 4);
    if (grid < 1) grid = 1;
    if (grid > 1024) grid = 1024;
    if (!fixedBox) {
        pointBoundsInitKernel<<<1, 64, 0, s>>>(bb);
        pointBoundsKernel<<<grid, kBlock, 0, s>>>(base, stride, count, bb);
    }
    mortonKeysKernel<<<grid, kBlock, 0, s>>>(base, stride, count, bb, fixedBox, keys, ws.hist.p);
    hipCheck(hipGetLastError(), "morton keys");
    sortPairs(ws, keys, nullptr, keysSorted, orderOut, static_cast<size_t>(count), 4, s);
}

void launchRays(const TraverseArgs& a, const Ray* rays, int count, uint64_t* out, hipStream_t s) {
    const int grid = blocksFor(count, kBlock);
    if (grid == 0) return;
    const Tri* tris = static_cast<const Tri*>(a.items);
    if (a.wide) {
        if (a.stats) {
            rayWideKernel<true><<<grid, kBlock, 0, s>>>(a.wideNodes, tris, rays, a.order, count, out, a.stats);
        } else {
            rayWideKernel<false><<<grid, kBlock, 0, s>>>(a.wideNodes, tris, rays, a.order, count, out, nullptr);
        }
    } else {
        if (a.stats) {
            rayBinaryKernel<true><<<grid, kBlock, 0, s>>>(a.nodes, tris, rays, a.order, count, out, a.stats);
        } else {
            rayBinaryKernel<false><<<grid, kBlock, 0, s>>>(a.nodes, tris, rays, a.order, count, out, nullptr);
        }
    }
    hipCheck(hipGetLastError(), "raycast");
}

template <typename L>
static void launchClosestT(const TraverseArgs& a, const V3* pts, int count, ClosestResult* out,
                           hipStream_t s) {
    const int grid = blocksFor(count, kBlock);
    const auto* prims = static_cast<const typename L::Item*>(a.items);
    if (a.wide) {
        if (a.stats)
            closestWideKernel<L, true><<<grid, kBlock, 0, s>>>(a.wideNodes, prims, a.primIds, pts, a.order, count, out, a.stats);
        else
            closestWideKernel<L, false><<<grid, kBlock, 0, s>>>(a.wideNodes, prims, a.primIds, pts, a.order, count, out, nullptr);
    } else {
        if (a.stats)
            closestBinaryKernel<L, true><<<grid, kBlock, 0, s>>>(a.nodes, prims, a.primIds, pts, a.order, count, out, a.stats);
        else
            closestBinaryKernel<L, false><<<grid, kBlock, 0, s>>>(a.nodes, prims, a.primIds, pts, a.order, count, out, nullptr);
    }
}

void launchClosest(const TraverseArgs& a, const V3* pts, int count, ClosestResult* out,
                   hipStream_t s) {
    if (count <= 0) return;
    switch (a.kind) {
        case 0: launchClosestT<TriLeaf>(a, pts, count, out, s); break;
        case 1: launchClosestT<PointLeaf>(a, pts, count, out, s); break;
        default: launchClosestT<GaussLeaf>(a, pts, count, out, s); break;
    }
    hipCheck(hipGetLastError(), "closest");
}

template <typename L, int K>
static void launchKnnTK(const TraverseArgs& a, const V3* pts, int count, int k, int outStride,
                        ClosestResult* out, hipStream_t s) {
    const int grid = blocksFor(count, kBlock);
    const auto* prims = static_cast<const typename L::Item*>(a.items);
    if (a.wide) {
        if (a.stats)
            knnWideKernel<L, K, true><<<grid, kBlock, 0, s>>>(a.wideNodes, prims, a.primIds, pts, a.order, count, k, outStride, out, a.stats);
        else
            knnWideKernel<L, K, false><<<grid, kBlock, 0, s>>>(a.wideNodes, prims, a.primIds, pts, a.order, count, k, outStride, out, nullptr);
    } else {
        if (a.stats)
            knnBinaryKernel<L, K, true><<<grid, kBlock, 0, s>>>(a.nodes, prims, a.primIds, pts, a.order, count, k, outStride, out, a.stats);
        else
            knnBinaryKernel<L, K, false><<<grid, kBlock, 0, s>>>(a.nodes, prims, a.primIds, pts, a.order, count, k, outStride, out, nullptr);
    }
}

template <typename L>
static void launchKnnT(const TraverseArgs& a, const V3* pts, int count, int k, int outStride,
                       ClosestResult* out, hipStream_t s) {
    if (k <= 4) {
        launchKnnTK<L, 4>(a, pts, count, k, outStride, out, s);
    } else if (k <= 8) {
        launchKnnTK<L, 8>(a, pts, count, k, outStride, out, s);
    } else if (k <= 16) {
        launchKnnTK<L, 16>(a, pts, count, k, outStride, out, s);
    } else {
        launchKnnTK<L, 32>(a, pts, count, k, outStride, out, s);
    }
}

void launchKnn(const TraverseArgs& a, const V3* pts, int count, int k, int outStride,
               ClosestResult* out, hipStream_t s) {
    if (count <= 0) return;
    switch (a.kind) {
        case 0: launchKnnT<TriLeaf>(a, pts, count, k, outStride, out, s); break;
        case 1: launchKnnT<PointLeaf>(a, pts, count, k, outStride, out, s); break;
        default: launchKnnT<GaussLeaf>(a, pts, count, k, outStride, out, s); break;
    }
    hipCheck(hipGetLastError(), "knn");
}

template <typename L, bool kFill>
static void launchRadiusT(const TraverseArgs& a, const V3* pts, int count, float r2, int* counts,
                          const int* offsets, uint64_t* keys, hipStream_t s) {
    const int grid = blocksFor(count, kBlock);
    const auto* prims = static_cast<const typename L::Item*>(a.items);
    unsigned long long* st = kFill ? nullptr : a.stats;
    if (a.wide) {
        if (st)
            radiusKernel<L, true, kFill, true><<<grid, kBlock, 0, s>>>(a.nodes, a.wideNodes, prims, a.primIds, pts, a.order, count, r2, counts, offsets, keys, st);
        else
            radiusKernel<L, true, kFill, false><<<grid, kBlock, 0, s>>>(a.nodes, a.wideNodes, prims, a.primIds, pts, a.order, count, r2, counts, offsets, keys, nullptr);
    } else {
        if (st)
            radiusKernel<L, false, kFill, true><<<grid, kBlock, 0, s>>>(a.nodes, a.wideNodes, prims, a.primIds, pts, a.order, count, r2, counts, offsets, keys, st);
        else
            radiusKernel<L, false, kFill, false><<<grid, kBlock, 0, s>>>(a.nodes, a.wideNodes, prims, a.primIds, pts, a.order, count, r2, counts, offsets, keys, nullptr);
    }
}

void launchRadiusCount(const TraverseArgs& a, const V3* pts, int count, float r2, int* counts,
                       hipStream_t s) {
    if (count <= 0) return;
    switch (a.kind) {
        case 0: launchRadiusT<TriLeaf, false>(a, pts, count, r2, counts, nullptr, nullptr, s); break;
        case 1: launchRadiusT<PointLeaf, false>(a, pts, count, r2, counts, nullptr, nullptr, s); break;
        default: launchRadiusT<GaussLeaf, false>(a, pts, count, r2, counts, nullptr, nullptr, s); break;
    }
    hipCheck(hipGetLastError(), "radius count");
}

void launchRadiusFill(const TraverseArgs& a, const V3* pts, int count, float r2,
                      const int* offsets, uint64_t* keys, hipStream_t s) {
    if (count <= 0) return;
    switch (a.kind) {
        case 0: launchRadiusT<TriLeaf, true>(a, pts, count, r2, nullptr, offsets, keys, s); break;
        case 1: launchRadiusT<PointLeaf, true>(a, pts, count, r2, nullptr, offsets, keys, s); break;
        default: launchRadiusT<GaussLeaf, true>(a, pts, count, r2, nullptr, offsets, keys, s); break;
    }
    hipCheck(hipGetLastError(), "radius fill");
}

void segmentSort(const int* offsets, uint64_t* keys, int count, int maxSeg, int* scratchInts,
                 hipStream_t s) {
    if (count <= 0) return;
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
    int h[2] = {0, 0};
    hipCheck(hipMemcpyAsync(h, info, sizeof(h), hipMemcpyDeviceToHost, s), "seg info");
    hipCheck(hipStreamSynchronize(s), "seg info sync");
    if (h[0] > 0) {
        segSortBlockKernel<<<h[0], kBlock, kBlockCap * sizeof(uint64_t), s>>>(offsets, blockList, keys);
        hipCheck(hipGetLastError(), "segment sort (block)");
    }
    if (h[1] > 0) {
        throw ArgError("radius: a query matched more than 8192 primitives; reduce the radius");
    }
}

}  // namespace qr
