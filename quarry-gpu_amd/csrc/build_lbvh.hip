// LBVH build for gfx950: centroid bounds -> 30-bit Morton codes (fused with the
// radix histogram) -> onesweep sort -> one pass that emits the hierarchy, gathers
// the leaves and merges boxes bottom-up (lbvhEmitKernel).
//
// Topology and node bytes follow the reference's LBVH exactly
// (src/gpu/lbvh.cu:277-315 Morton, :366-414 Karras, :345-354 leaf boxes,
// :472-514 parent box merge), so treeHash equals the reference's for the same
// input.  What differs is the schedule: the reference sorts 64-bit keys with 8
// digit passes, runs a Karras kernel (binary searches per node), computes depths
// by walking parent chains, depth-sorts the nodes and launches one merge kernel
// per level; here the sort runs 4 passes and the Karras-numbered tree is built
// agglomeratively in a single bottom-up pass where the second child to arrive at
// a split merges it; the batched build runs the same window and hand-up passes per
// scene.  A bottom-up climb over a given parent[] (leafUpLocalKernel) serves boxes
// over a topology that is already known (SAH, refits).
#include "qr_bvh.h"

namespace qr {

namespace {

constexpr int kBlock = 256;

// ---- coherent (agent-scope) accessors for data produced inside one kernel ------
// Boxes written by one workgroup are read by another that may sit on a different
// XCD with its own L2.  Every value that crosses workgroups (boxes, heights, the
// arrival counters) is an agent-scope atomic access, so each location is coherent;
// the explicit vmcnt wait before the arrival atomic makes this thread's box stores
// complete before the arrival is counted, and its "memory" clobber keeps the
// compiler from moving them after it.  A release RMW plus an acquire fence on the
// winning path would state the same order in the language's terms, but at agent
// scope it also writes back the XCD's L2 on every arrival: measured on the 2M
// terrain, the leaves + boxes pass of the SAH build went from 0.20 to 4.36 ms.
__device__ __forceinline__ void stAgent(float* p, float v) {
    __hip_atomic_store(reinterpret_cast<uint32_t*>(p), floatBits(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float ldAgent(const float* p) {
    return bitsFloat(__hip_atomic_load(reinterpret_cast<const uint32_t*>(p), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ void stAgentI(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ldAgentI(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void storesDone() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

__device__ __forceinline__ void writeBoxAgent(Node2* nd, const Box& b) {
    stAgent(&nd->lo[0], b.lo.x);
    stAgent(&nd->lo[1], b.lo.y);
    stAgent(&nd->lo[2], b.lo.z);
    stAgent(&nd->hi[0], b.hi.x);
    stAgent(&nd->hi[1], b.hi.y);
    stAgent(&nd->hi[2], b.hi.z);
}
__device__ __forceinline__ Box readBoxAgent(const Node2* nd) {
    return {{ldAgent(&nd->lo[0]), ldAgent(&nd->lo[1]), ldAgent(&nd->lo[2])},
            {ldAgent(&nd->hi[0]), ldAgent(&nd->hi[1]), ldAgent(&nd->hi[2])}};
}

// ---- centroid bounds -----------------------------------------------------------
// Each bounds workgroup w writes its partial bounds to part[w * kPartStride + 0..5]
// (ordered bits: min xyz, max xyz; src/gpu/lbvh.cu:233-275 takes the same min / max
// with atomics); the Morton pass reduces the partials, so nothing needs
// initialising and no same-address atomics serialize.  The bounds pass also zeroes
// what the rest of the build expects zero (sort histograms, look-back descriptors and
// tickets, arrival counters), so the build needs no preparation launch.
constexpr int kPartStride = 8;
constexpr int kMaxBoundsBlocks = 1024;
constexpr int kUnroll = 4;  // items in flight per thread in the streaming passes

struct BuildZero {
    uint32_t* hist;
    uint32_t* desc;
    long long descWords;
    uint32_t* tickets;
    int* arrive;
    int slots;
};
__device__ __forceinline__ void zeroForBuild(const BuildZero& z) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long j = i; j < z.descWords; j += stride) z.desc[j] = 0u;
    for (long long j = i; j < z.slots; j += stride) z.arrive[j] = 0;
    for (long long j = i; j < kHistCopies * kHistCopyWords; j += stride) z.hist[j] = 0u;
    if (i < 4) z.tickets[i] = 0u;
}

// Workgroup reduction of the per-thread centroid bounds into this workgroup's partial.
__device__ __forceinline__ void boundsFlush(float lo0, float lo1, float lo2, float hi0, float hi1, float hi2,
                                            uint32_t* __restrict__ cb) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lo0 = fminf(lo0, __shfl_xor(lo0, off, 64));
        lo1 = fminf(lo1, __shfl_xor(lo1, off, 64));
        lo2 = fminf(lo2, __shfl_xor(lo2, off, 64));
        hi0 = fmaxf(hi0, __shfl_xor(hi0, off, 64));
        hi1 = fmaxf(hi1, __shfl_xor(hi1, off, 64));
        hi2 = fmaxf(hi2, __shfl_xor(hi2, off, 64));
    }
    __shared__ float s[6][kBlock / 64];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s[0][w] = lo0;
        s[1][w] = lo1;
        s[2][w] = lo2;
        s[3][w] = hi0;
        s[4][w] = hi1;
        s[5][w] = hi2;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int c = threadIdx.x;
        float v = s[c][0];
        for (int k = 1; k < kBlock / 64; ++k) v = c < 3 ? fminf(v, s[c][k]) : fmaxf(v, s[c][k]);
        cb[blockIdx.x * kPartStride + c] = orderedBits(v);
    }
}

template <typename L>
__global__ __launch_bounds__(kBlock) void centroidBoundsKernel(
    const typename L::Item* __restrict__ items, int n, uint32_t* __restrict__ cb, BuildZero z) {
    zeroForBuild(z);
    float lo0 = kInf, lo1 = kInf, lo2 = kInf;
    float hi0 = -kInf, hi1 = -kInf, hi2 = -kInf;
    const int stride = kBlock * gridDim.x;
    int i = blockIdx.x * kBlock + threadIdx.x;
    for (; i + (kUnroll - 1) * stride < n; i += kUnroll * stride) {
        typename L::Item it[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) it[u] = items[i + u * stride];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const V3 c = L::centroid(it[u]);
            lo0 = fminf(lo0, c.x);
            lo1 = fminf(lo1, c.y);
            lo2 = fminf(lo2, c.z);
            hi0 = fmaxf(hi0, c.x);
            hi1 = fmaxf(hi1, c.y);
            hi2 = fmaxf(hi2, c.z);
        }
    }
    for (; i < n; i += stride) {
        const V3 c = L::centroid(items[i]);
        lo0 = fminf(lo0, c.x);
        lo1 = fminf(lo1, c.y);
        lo2 = fminf(lo2, c.z);
        hi0 = fmaxf(hi0, c.x);
        hi1 = fmaxf(hi1, c.y);
        hi2 = fmaxf(hi2, c.z);
    }
    boundsFlush(lo0, lo1, lo2, hi0, hi1, hi2, cb);
}

// ---- Morton codes + radix histogram ----------------------------------------------

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// quantize (src/gpu/lbvh.cu:289-294): (v - lo) / extent * 1023 + 0.5, clamped, truncated.
__device__ __forceinline__ uint32_t quantize1023(float v, float lo, float extent) {
    if (extent <= 0.0f) return 0u;
    const float q = (v - lo) / extent * 1023.0f + 0.5f;
    return static_cast<uint32_t>(static_cast<int>(fminf(1023.0f, fmaxf(0.0f, q))));
}

// The build's quantization frame: the centroid bounds folded over their copies.
struct MortonFrame {
    float lo0, lo1, lo2, ex0, ex1, ex2;
    __device__ uint32_t code(const V3& c) const {
        return spread10(quantize1023(c.x, lo0, ex0)) | (spread10(quantize1023(c.y, lo1, ex1)) << 1) |
               (spread10(quantize1023(c.z, lo2, ex2)) << 2);
    }
};
// The frame from the bounds partials of `parts` workgroups, reduced by the whole
// workgroup of kBlock threads (every thread returns it).
__device__ __forceinline__ MortonFrame mortonFrame(const uint32_t* __restrict__ cb, int parts) {
    __shared__ uint32_t sb[kBlock / 64][6];
    uint32_t b[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    // a partial is two 16-byte words (6 bounds words + 2 unused); every thread issues
    // all of its loads before it folds them (parts <= kMaxBoundsBlocks = 4 x 256)
    const uint4* part = reinterpret_cast<const uint4*>(cb);
    constexpr int kPer = kMaxBoundsBlocks / kBlock;
    uint4 lo[kPer], hi[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int w = u * kBlock + threadIdx.x;
        if (w < parts) {
            lo[u] = part[2 * w];
            hi[u] = part[2 * w + 1];
        } else {
            lo[u] = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u);
            hi[u] = make_uint4(0u, 0u, 0u, 0u);
        }
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        b[0] = lo[u].x < b[0] ? lo[u].x : b[0];
        b[1] = lo[u].y < b[1] ? lo[u].y : b[1];
        b[2] = lo[u].z < b[2] ? lo[u].z : b[2];
        b[3] = lo[u].w > b[3] ? lo[u].w : b[3];
        b[4] = hi[u].x > b[4] ? hi[u].x : b[4];
        b[5] = hi[u].y > b[5] ? hi[u].y : b[5];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        b[c] = reduce64<MinU>(b[c]);
        b[3 + c] = reduce64<MaxU>(b[3 + c]);
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int c = 0; c < 6; ++c) sb[threadIdx.x >> 6][c] = b[c];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        uint32_t x = sb[0][c];
        for (int w = 1; w < kBlock / 64; ++w) {
            const uint32_t v = sb[w][c];
            x = c < 3 ? (v < x ? v : x) : (v > x ? v : x);
        }
        b[c] = x;
    }
    const float lo0 = fromOrderedBits(b[0]), lo1 = fromOrderedBits(b[1]), lo2 = fromOrderedBits(b[2]);
    return {lo0, lo1, lo2, fromOrderedBits(b[3]) - lo0, fromOrderedBits(b[4]) - lo1, fromOrderedBits(b[5]) - lo2};
}

template <typename L>
__global__ __launch_bounds__(kBlock) void mortonHistKernel(const typename L::Item* __restrict__ items,
                                                           int n, const uint32_t* __restrict__ cb, int parts,
                                                           uint32_t* __restrict__ codes,
                                                           uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[4][256];
    for (int i = threadIdx.x; i < 4 * 256; i += kBlock) (&h[0][0])[i] = 0u;
    const MortonFrame f = mortonFrame(cb, parts);
    __syncthreads();
    auto emit = [&](int i, const typename L::Item& item) {
        const uint32_t code = f.code(L::centroid(item));
        codes[i] = code;
        atomicAdd(&h[0][code & 255u], 1u);
        atomicAdd(&h[1][(code >> 8) & 255u], 1u);
        atomicAdd(&h[2][(code >> 16) & 255u], 1u);
        atomicAdd(&h[3][code >> 24], 1u);
    };
    const int stride = kBlock * gridDim.x;
    int i = blockIdx.x * kBlock + threadIdx.x;
    for (; i + (kUnroll - 1) * stride < n; i += kUnroll * stride) {
        typename L::Item it[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) it[u] = items[i + u * stride];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) emit(i + u * stride, it[u]);
    }
    for (; i < n; i += stride) emit(i, items[i]);
    __syncthreads();
    flushHistogram(&h[0][0], 4 * 256, hist);
}

// ---- triangle bounds with coalesced loads ------------------------------------------------
// A triangle is 36 bytes, so a lane-per-item load makes every lane start its own
// 36-byte run; here a workgroup copies kStageTris consecutive triangles (a multiple of
// 16 bytes) into LDS with 16-byte loads on consecutive lanes, then each thread reads
// its triangles from LDS (a 9-word stride: no bank conflicts).  Used when the
// triangle array is 16-byte aligned.
constexpr int kStageTris = 4 * kBlock;
constexpr int kStageVec = kStageTris * 36 / 16 / kBlock;  // 16-byte loads per thread
static_assert(kStageTris * 36 % (16 * kBlock) == 0, "staging chunk");

__device__ __forceinline__ void stageTris(const Tri* __restrict__ tris, long long base, int cnt,
                                          uint4* __restrict__ lds) {
    if (cnt == kStageTris) {
        const uint4* src = reinterpret_cast<const uint4*>(tris + base);
        uint4 v[kStageVec];
#pragma unroll
        for (int k = 0; k < kStageVec; ++k) v[k] = src[k * kBlock + threadIdx.x];
#pragma unroll
        for (int k = 0; k < kStageVec; ++k) lds[k * kBlock + threadIdx.x] = v[k];
    } else {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(tris + base);
        uint32_t* dst = reinterpret_cast<uint32_t*>(lds);
        for (int k = threadIdx.x; k < cnt * 9; k += kBlock) dst[k] = src[k];
    }
}

__global__ __launch_bounds__(kBlock) void triBoundsStagedKernel(const Tri* __restrict__ tris, int n,
                                                                uint32_t* __restrict__ cb, BuildZero z) {
    zeroForBuild(z);
    __shared__ uint4 sStage[kStageTris * 36 / 16];
    float lo0 = kInf, lo1 = kInf, lo2 = kInf;
    float hi0 = -kInf, hi1 = -kInf, hi2 = -kInf;
    const int chunks = (n + kStageTris - 1) / kStageTris;
    for (int c = blockIdx.x; c < chunks; c += gridDim.x) {
        const long long base = static_cast<long long>(c) * kStageTris;
        const int cnt = n - base < kStageTris ? static_cast<int>(n - base) : kStageTris;
        stageTris(tris, base, cnt, sStage);
        __syncthreads();
        const Tri* st = reinterpret_cast<const Tri*>(sStage);
#pragma unroll
        for (int u = 0; u < kStageTris / kBlock; ++u) {
            const int j = u * kBlock + threadIdx.x;
            if (j < cnt) {
                const V3 q = triCentroid(st[j]);
                lo0 = fminf(lo0, q.x);
                lo1 = fminf(lo1, q.y);
                lo2 = fminf(lo2, q.z);
                hi0 = fmaxf(hi0, q.x);
                hi1 = fmaxf(hi1, q.y);
                hi2 = fmaxf(hi2, q.z);
            }
        }
        __syncthreads();
    }
    boundsFlush(lo0, lo1, lo2, hi0, hi1, hi2, cb);
}

// ---- Karras 2012 hierarchy ---------------------------------------------------------
// Common-prefix length of the 64-bit keys (code << 32 | index) at sorted
// positions a and b; -1 outside [0, n) (src/gpu/lbvh.cu:373-377).
__device__ __forceinline__ int prefixLen(const uint32_t* __restrict__ code,
                                         const uint32_t* __restrict__ idx, int n, int a, int b) {
    if (b < 0 || b >= n) return -1;
    const uint32_t ca = code[a], cb = code[b];
    if (ca != cb) return __clz(ca ^ cb);
    const uint32_t x = idx[a] ^ idx[b];
    return x == 0u ? 64 : 32 + __clz(x);
}

// ---- bottom-up boxes over a given topology --------------------------------------------
// Bottom-up boxes with the block-local part in LDS.  Workgroup b owns leaf slots
// [256 b, 256 b + 255] = [base, last].  An internal node p whose leaf range
// [min(p, far[p]), max(p, far[p])] lies inside that window has both subtrees in this
// workgroup: its arrival counter, box and height live in LDS (the stores to the node
// array are ordinary write-back stores).  Node ids are range endpoints (Karras
// numbering, which the SAH builder follows too), so the topology of every node that
// can be local -- parent[], far[], children of ids base..last, parents of the
// window's leaves -- is read into LDS up front with coalesced loads and the local
// climb makes no dependent global loads.  The first node whose range leaves the
// window is published with agent-scope stores and from there on the climb takes
// device-wide arrival counters: the first child to arrive stops, the second merges
// both boxes (left operand = left child, as the reference's refit does), returns the
// counter to zero and continues, with the height riding along.  Every node is merged
// from the same operands in the same order as by a climb from each leaf.
// One workgroup's window of leaf slots [base, base + kBlock) of a tree of n leaves
// (the whole body of leafUpLocalKernel).
template <typename L, bool kGather>
__device__ __forceinline__ void leafUpWindow(
    const typename L::Item* __restrict__ src, const uint32_t* __restrict__ order, int n,
    typename L::Item* __restrict__ sorted, int* __restrict__ primIds, Node2* __restrict__ nodes,
    const int* __restrict__ parent, int* __restrict__ arrive, int* __restrict__ height,
    const int* __restrict__ farEnd, int base) {
    __shared__ float sLeaf[kBlock][6];
    __shared__ float sNode[kBlock][6];
    __shared__ int sCount[kBlock];
    __shared__ int sHeight[kBlock];
    __shared__ int sLeafParent[kBlock];
    __shared__ int sParent[kBlock];
    __shared__ int sFar[kBlock];
    __shared__ int sLeft[kBlock];
    __shared__ int sRight[kBlock];
    const int last = (base + kBlock < n ? base + kBlock : n) - 1;
    const int k = base + threadIdx.x;
    const bool live = k < n;
    sCount[threadIdx.x] = 0;
    if (live) sLeafParent[threadIdx.x] = parent[n - 1 + k];
    if (k < n - 1) {
        sParent[threadIdx.x] = parent[k];
        sFar[threadIdx.x] = farEnd[k];
        sLeft[threadIdx.x] = nodes[k].left;
        sRight[threadIdx.x] = nodes[k].right;
    } else {
        sFar[threadIdx.x] = -1;  // not an internal node: never local
    }

    Box box{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
    if (live) {
        typename L::Item item;
        if (kGather) {
            const int sceneIdx = static_cast<int>(order[k]);
            item = src[sceneIdx];
            sorted[k] = item;
            primIds[k] = sceneIdx;
        } else {
            item = sorted[k];
        }
        box = L::bounds(item);
        Node2 leaf;
        leaf.lo[0] = box.lo.x;
        leaf.lo[1] = box.lo.y;
        leaf.lo[2] = box.lo.z;
        leaf.hi[0] = box.hi.x;
        leaf.hi[1] = box.hi.y;
        leaf.hi[2] = box.hi.z;
        leaf.left = -(k + 1);
        leaf.right = 1;
        nodes[n - 1 + k] = leaf;
        float* mine = sLeaf[threadIdx.x];
        mine[0] = box.lo.x;
        mine[1] = box.lo.y;
        mine[2] = box.lo.z;
        mine[3] = box.hi.x;
        mine[4] = box.hi.y;
        mine[5] = box.hi.z;
    }
    __syncthreads();
    if (!live) return;

    int node = n - 1 + k;
    int p = sLeafParent[threadIdx.x];
    int h = 0;
    // local part: node is a leaf of the window or an internal node with a local range
    while (p >= base && p <= last) {
        const int f = sFar[p - base];
        if (f < base || f > last) break;
        const int prior = __hip_atomic_fetch_add(&sCount[p - base], 1, __ATOMIC_ACQ_REL,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP);
        if (prior == 0) return;
        const int l = sLeft[p - base];
        const int r = sRight[p - base];
        const int sib = (l == node) ? r : l;
        const float* sb = sib >= n - 1 ? sLeaf[sib - (n - 1) - base] : sNode[sib - base];
        const Box other{{sb[0], sb[1], sb[2]}, {sb[3], sb[4], sb[5]}};
        const int sh = sib >= n - 1 ? 0 : sHeight[sib - base];
        box = (l == node) ? mergeChildren(box, other) : mergeChildren(other, box);
        h = (h > sh ? h : sh) + 1;
        float* dst = sNode[p - base];
        dst[0] = box.lo.x;
        dst[1] = box.lo.y;
        dst[2] = box.lo.z;
        dst[3] = box.hi.x;
        dst[4] = box.hi.y;
        dst[5] = box.hi.z;
        sHeight[p - base] = h;
        Node2* nd = &nodes[p];
        nd->lo[0] = box.lo.x;
        nd->lo[1] = box.lo.y;
        nd->lo[2] = box.lo.z;
        nd->hi[0] = box.hi.x;
        nd->hi[1] = box.hi.y;
        nd->hi[2] = box.hi.z;
        height[p] = h;
        node = p;
        p = sParent[p - base];
    }
    if (p < 0) return;  // the whole tree lies in this window
    // p's range leaves this workgroup's window: publish this subtree's root for the
    // other child's workgroup, then take the global arrival counters
    writeBoxAgent(&nodes[node], box);
    if (node < n - 1) stAgentI(&height[node], h);
    while (true) {
        storesDone();
        const int prior = __hip_atomic_fetch_add(&arrive[p], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (prior == 0) break;
        const int l = nodes[p].left;
        const int r = nodes[p].right;
        const int sib = (l == node) ? r : l;
        const Box sb = readBoxAgent(&nodes[sib]);
        const int sh = sib >= n - 1 ? 0 : ldAgentI(&height[sib]);
        box = (l == node) ? mergeChildren(box, sb) : mergeChildren(sb, box);
        h = (h > sh ? h : sh) + 1;
        writeBoxAgent(&nodes[p], box);
        stAgentI(&height[p], h);
        stAgentI(&arrive[p], 0);
        node = p;
        p = parent[p];
        if (p < 0) break;
    }
}

template <typename L, bool kGather>
__global__ __launch_bounds__(kBlock) void leafUpLocalKernel(
    const typename L::Item* __restrict__ src, const uint32_t* __restrict__ order, int n,
    typename L::Item* __restrict__ sorted, int* __restrict__ primIds, Node2* __restrict__ nodes,
    const int* __restrict__ parent, int* __restrict__ arrive, int* __restrict__ height,
    const int* __restrict__ farEnd, const int* __restrict__ gate) {
    // a builder whose topology failed (its error word set) leaves it unclimbed; the
    // host reports the error after the stream completes
    if (gate != nullptr && *gate != 0) return;
    leafUpWindow<L, kGather>(src, order, n, sorted, primIds, nodes, parent, arrive, height, farEnd,
                             static_cast<int>(blockIdx.x) * kBlock);
}

// ---- fused hierarchy emission + bottom-up boxes ------------------------------------------
// delta(i) = common-prefix length of sorted keys i and i+1, -1 outside [0, n-2].
__device__ __forceinline__ int adjPrefix(const uint32_t* __restrict__ code,
                                         const uint32_t* __restrict__ idx, int n, int i) {
    if (i < 0 || i >= n - 1) return -1;
    return prefixLen(code, idx, n, i, i + 1);
}

// Where node [l, r] attaches: the binary radix tree over distinct keys joins a node
// with the neighbour it shares the longer prefix with (delta(r) vs delta(l-1); they
// are never equal).  As a left child its parent splits at r, as a right child at l-1.
// Karras numbering names an internal node by the end of its range that faces its
// parent's split: a left child is node r, a right child node l, the root node 0
// (Karras 2012's numbering).
struct Attach {
    int split;  // parent's split position (keys split..split+1 straddle it)
    bool left;  // this node is the parent's left child
    int id;     // this node's Karras id (internal nodes)
};
__device__ __forceinline__ Attach attachOf(int l, int r, int dl, int dr) {
    const bool left = dr > dl;
    return {left ? r : l - 1, left, left ? r : l};
}

// A finished node on its way up: leaf range, Karras id, height, the prefixes
// delta(l-1) and delta(r) that decide where it attaches, and its box.
struct NodeState {
    int l, r, id, h, dl, dr;
    Box box;
};

// A node handed up to the next level (48 B).
struct FrontierEntry {
    int l, r, id, h, dl, dr;
    float box[6];
};
static_assert(sizeof(FrontierEntry) == 48, "FrontierEntry layout");

__device__ __forceinline__ FrontierEntry toEntry(const NodeState& s) {
    return {s.l, s.r, s.id, s.h, s.dl, s.dr, {s.box.lo.x, s.box.lo.y, s.box.lo.z, s.box.hi.x, s.box.hi.y, s.box.hi.z}};
}
__device__ __forceinline__ NodeState fromEntry(const FrontierEntry& e) {
    return {e.l, e.r, e.id, e.h, e.dl, e.dr, {{e.box[0], e.box[1], e.box[2]}, {e.box[3], e.box[4], e.box[5]}}};
}

// Rank of this thread's pending node among the workgroup's pending nodes by
// position (a leaf offset or an entry index, distinct per pending node), so the
// frontier is staged in leaf order; sWave[kThreads / 64] receives the count.
template <int kThreads = kBlock>
__device__ __forceinline__ int frontierRank(bool pend, int pos, int* sMark, int* sWave) {
    const int t = threadIdx.x;
    sMark[t] = 0;
    __syncthreads();
    if (pend) sMark[pos] = 1;
    __syncthreads();
    const uint64_t bal = __ballot(sMark[t]);
    const int lane = t & 63;
    const int before = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) sWave[t >> 6] = __popcll(bal);
    __syncthreads();
    int prefix = 0;
    for (int w = 0; w < (t >> 6); ++w) prefix += sWave[w];
    sMark[t] = prefix + before;
    __syncthreads();
    if (t == 0) {
        int total = 0;
        for (int w = 0; w < kThreads / 64; ++w) total += sWave[w];
        sWave[kThreads / 64] = total;
    }
    const int rank = pend ? sMark[pos] : 0;
    __syncthreads();
    return rank;
}

// Ids of the two children merged at split g, given this node [l, r] and the far end
// of its sibling; returns the parent's range through lo / hi and the sibling's id.
__device__ __forceinline__ void mergeIds(int n, int l, int r, int my, int g, bool left, int sibFar,
                                         int& lo, int& hi, int& lid, int& rid, int& sib) {
    if (left) {
        lo = l;
        hi = sibFar;
        lid = my;
        rid = sib = (sibFar == g + 1) ? n + g : g + 1;
    } else {
        lo = sibFar;
        hi = r;
        lid = sib = (sibFar == g) ? n - 1 + g : g;
        rid = my;
    }
}

// The agent-scope arrival protocol, from node s up to where its thread arrives
// first: arrive[split] takes the far end + 1 by an exchange (0 = empty; the second
// child resets it, so every slot is zero between kernels), boxes and heights cross
// workgroups with agent-scope stores.
__device__ void globalClimb(NodeState s, const uint32_t* __restrict__ code,
                            const uint32_t* __restrict__ order, int n, Node2* __restrict__ nodes,
                            int* __restrict__ parent, int* __restrict__ arrive, int* __restrict__ height,
                            int* __restrict__ farEnd) {
    writeBoxAgent(&nodes[s.id], s.box);
    if (s.id < n - 1) stAgentI(&height[s.id], s.h);
    while (!(s.l == 0 && s.r == n - 1)) {
        const Attach at = attachOf(s.l, s.r, s.dl, s.dr);
        const int g = at.split;
        storesDone();
        const int prior = __hip_atomic_exchange(&arrive[g], (at.left ? s.l : s.r) + 1, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
        if (prior == 0) return;
        stAgentI(&arrive[g], 0);
        int lo, hi, lid, rid, sib;
        mergeIds(n, s.l, s.r, s.id, g, at.left, prior - 1, lo, hi, lid, rid, sib);
        const Box sb = readBoxAgent(&nodes[sib]);
        const int sh = sib >= n - 1 ? 0 : ldAgentI(&height[sib]);
        const Box box = at.left ? mergeChildren(s.box, sb) : mergeChildren(sb, s.box);
        const int h = (s.h > sh ? s.h : sh) + 1;
        const int ndl = adjPrefix(code, order, n, lo - 1);
        const int ndr = adjPrefix(code, order, n, hi);
        const int pid = (lo == 0 && hi == n - 1) ? 0 : attachOf(lo, hi, ndl, ndr).id;
        writeBoxAgent(&nodes[pid], box);
        nodes[pid].left = lid;
        nodes[pid].right = rid;
        stAgentI(&height[pid], h);
        parent[lid] = pid;
        parent[rid] = pid;
        farEnd[pid] = pid == lo ? hi : lo;
        s = {lo, hi, pid, h, ndl, ndr, box};
    }
}

// Frontier hand-up between kernels.  The nodes a workgroup cannot finish inside
// its leaf window -- its frontier, the maximal subtrees in the window, in leaf
// order -- are staged in global memory with their ranges, prefixes, heights and
// boxes.  The next kernel merges the frontiers of kMergeUnits neighbouring units per
// workgroup in LDS, with the boundaries between consecutive entries as split slots,
// and stages the group's own frontier; the last level makes the root.  Kernel
// boundaries order the staging, so no workgroup waits and no device-scope atomic
// is needed.  A unit whose frontier has more than the capacity (a tree with long
// spines across windows) sends its nodes through globalClimb instead and stages a
// count of -1, which sends the whole group, and every group above it, the same
// way -- so the two protocols never meet at one split.
constexpr int kMergeUnits = 32;
// The batched build merges 8 units per group (256 threads, 25 KB of LDS: six groups per
// CU, where the thousands of small scenes' groups are latency bound) at the cost of a
// level more for its largest scenes.
constexpr int kBatchMergeUnits = 8;
constexpr int kFrontierCap = 32;
constexpr int kMergeBlock = kMergeUnits * kFrontierCap;  // one thread per entry
constexpr int kEmitMaxLevels = 10;

struct EmitLevels {
    int levels;                    // merge levels above the windows (0: one window holds all)
    int units[kEmitMaxLevels];     // units per level (level 0: the windows)
    int stageOff[kEmitMaxLevels];  // first FrontierEntry of level L's staging (kFrontierCap per unit)
    int countOff[kEmitMaxLevels];  // first frontier count of level L's units
};

int gFrontierCap = kFrontierCap;  // qr_set_frontier_cap

// One thread per leaf slot k: gather the item into leaf order, write leaf node
// n-1+k, then build the tree bottom-up without a separate hierarchy pass
// (agglomerative order, as in Apetrei 2014): a finished node [l, r] knows its
// parent's split from two adjacent-key prefixes, and the split slot's arrival hands
// the second child the far end of its sibling, from which both child ids, the
// parent range and the parent's Karras id follow.  The result is byte-identical to
// a Karras pass (Karras 2012) + leafUpLocalKernel (same ids, parent[], farEnd[], boxes merged with
// the left child's box as the left operand, heights).
//
// Workgroup b owns leaf slots [256 b, 256 b + 255] = [base, last].  Split slots
// base..last-1 have both neighbouring leaves in the window: arrivals from this
// workgroup meet there in LDS (prefixes of the window come from an LDS table).  A
// node whose parent splits outside the window, or the first child at an LDS slot
// whose sibling was not finished here (known after the workgroup barrier), is on the
// window's frontier and is handed up.
// One window (win, of kBlock leaf slots) of a tree of n leaves; its frontier goes to
// stageUnit / *countUnit; single: the window holds the whole tree.
template <typename L>
__device__ __forceinline__ void emitWindow(
    const typename L::Item* __restrict__ src, const uint32_t* __restrict__ code,
    const uint32_t* __restrict__ order, int n, typename L::Item* __restrict__ sorted,
    int* __restrict__ primIds, Node2* __restrict__ nodes, int* __restrict__ parent,
    int* __restrict__ arrive, int* __restrict__ height, int* __restrict__ farEnd,
    FrontierEntry* __restrict__ stageUnit, int* __restrict__ countUnit, int frontierCap, int win, bool single) {
    __shared__ float sLeaf[kBlock][6];
    __shared__ float sNode[kBlock][6];
    __shared__ int sSlot[kBlock];
    __shared__ int sHeight[kBlock];
    __shared__ int sDelta[kBlock + 1];  // delta(base - 1 + j)
    __shared__ int sMark[kBlock];
    __shared__ int sWave[kBlock / 64 + 1];
    const int t = threadIdx.x;
    const int base = win * kBlock;
    const int last = (base + kBlock < n ? base + kBlock : n) - 1;
    const int k = base + t;
    const bool live = k < n;
    sSlot[t] = 0;
    sDelta[t] = adjPrefix(code, order, n, base - 1 + t);
    if (t == 0) {
        sDelta[kBlock] = adjPrefix(code, order, n, base + kBlock - 1);
        if (win == 0) parent[0] = -1;
    }

    NodeState s{k, k, n - 1 + k, 0, -1, -1, {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}}};
    if (live) {
        const int sceneIdx = static_cast<int>(order[k]);
        const typename L::Item item = src[sceneIdx];
        sorted[k] = item;
        primIds[k] = sceneIdx;
        s.box = L::bounds(item);
        Node2 leaf;
        leaf.lo[0] = s.box.lo.x;
        leaf.lo[1] = s.box.lo.y;
        leaf.lo[2] = s.box.lo.z;
        leaf.hi[0] = s.box.hi.x;
        leaf.hi[1] = s.box.hi.y;
        leaf.hi[2] = s.box.hi.z;
        leaf.left = -(k + 1);
        leaf.right = 1;
        nodes[n - 1 + k] = leaf;
        float* mine = sLeaf[t];
        mine[0] = s.box.lo.x;
        mine[1] = s.box.lo.y;
        mine[2] = s.box.lo.z;
        mine[3] = s.box.hi.x;
        mine[4] = s.box.hi.y;
        mine[5] = s.box.hi.z;
    }
    __syncthreads();

    // ---- the window ----------------------------------------------------------------
    bool pend = false;
    int dangling = -1;
    Attach at{0, false, 0};
    if (live) {
        s.dl = sDelta[t];
        s.dr = sDelta[t + 1];
    }
    bool done = !live;
    if (live) at = attachOf(s.l, s.r, s.dl, s.dr);
    while (!done) {
        const int g = at.split;
        if (g < base || g >= last) {
            pend = true;
            break;
        }
        const int prior = __hip_atomic_exchange(&sSlot[g - base], (at.left ? s.l : s.r) + 1, __ATOMIC_ACQ_REL,
                                                __HIP_MEMORY_SCOPE_WORKGROUP);
        if (prior == 0) {
            dangling = g;
            break;
        }
        int lo, hi, lid, rid, sib;
        mergeIds(n, s.l, s.r, s.id, g, at.left, prior - 1, lo, hi, lid, rid, sib);
        const float* sb = sib >= n - 1 ? sLeaf[sib - (n - 1) - base] : sNode[sib - base];
        const Box other{{sb[0], sb[1], sb[2]}, {sb[3], sb[4], sb[5]}};
        const int sh = sib >= n - 1 ? 0 : sHeight[sib - base];
        s.box = at.left ? mergeChildren(s.box, other) : mergeChildren(other, s.box);
        s.h = (s.h > sh ? s.h : sh) + 1;
        s.dl = sDelta[lo - base];
        s.dr = sDelta[hi - base + 1];
        int pid = 0;
        if (lo == 0 && hi == n - 1) {
            done = true;
        } else {
            at = attachOf(lo, hi, s.dl, s.dr);
            pid = at.id;
        }
        float* dst = sNode[pid - base];
        dst[0] = s.box.lo.x;
        dst[1] = s.box.lo.y;
        dst[2] = s.box.lo.z;
        dst[3] = s.box.hi.x;
        dst[4] = s.box.hi.y;
        dst[5] = s.box.hi.z;
        sHeight[pid - base] = s.h;
        Node2 nd;
        nd.lo[0] = s.box.lo.x;
        nd.lo[1] = s.box.lo.y;
        nd.lo[2] = s.box.lo.z;
        nd.hi[0] = s.box.hi.x;
        nd.hi[1] = s.box.hi.y;
        nd.hi[2] = s.box.hi.z;
        nd.left = lid;
        nd.right = rid;
        nodes[pid] = nd;
        if (pid == 0) height[0] = s.h;  // the tree depth; other heights stay in LDS / entries
        parent[lid] = pid;
        parent[rid] = pid;
        farEnd[pid] = pid == lo ? hi : lo;
        s.l = lo;
        s.r = hi;
        s.id = pid;
    }
    __syncthreads();
    // a first arrival whose sibling never came in this workgroup is on the frontier
    if (dangling >= 0 && sSlot[dangling - base] == (at.left ? s.l : s.r) + 1) pend = true;
    int pos = pend ? s.l - base : 0;  // position of the entry inside its unit

    // ---- hand-up: stage the window's frontier in leaf order ----------------------------
    if (single) {  // one window holds the whole tree: the root is made
        if (pend) globalClimb(s, code, order, n, nodes, parent, arrive, height, farEnd);
        return;
    }
    const int rank = frontierRank(pend, pos, sMark, sWave);
    const int total = sWave[kBlock / 64];
    if (total > frontierCap) {
        if (pend) globalClimb(s, code, order, n, nodes, parent, arrive, height, farEnd);
        if (t == 0) *countUnit = -1;
    } else {
        if (pend) stageUnit[rank] = toEntry(s);
        if (t == 0) *countUnit = total;
    }
}

template <typename L>
__global__ __launch_bounds__(kBlock) void lbvhEmitKernel(
    const typename L::Item* __restrict__ src, const uint32_t* __restrict__ code,
    const uint32_t* __restrict__ order, int n, typename L::Item* __restrict__ sorted,
    int* __restrict__ primIds, Node2* __restrict__ nodes, int* __restrict__ parent,
    int* __restrict__ arrive, int* __restrict__ height, int* __restrict__ farEnd,
    FrontierEntry* __restrict__ stage, int* __restrict__ counts, int frontierCap) {
    emitWindow<L>(src, code, order, n, sorted, primIds, nodes, parent, arrive, height, farEnd,
                  stage + static_cast<size_t>(blockIdx.x) * kFrontierCap, counts + blockIdx.x, frontierCap,
                  static_cast<int>(blockIdx.x), gridDim.x == 1);
}

// Level L >= 1 of the hand-up: a group (one workgroup) merges the frontiers of
// consecutive units of level L-1 (32 for one tree, 8 in the batched build) and stages
// its own frontier as one unit of level L.
//
// The group's entries partition its leaf range in leaf order, so the nodes above
// them are fixed by the prefixes at the boundaries between consecutive entries
// (slot j between entries j and j+1, plus the group's two outer boundaries): the node
// splitting at slot j spans the entries between the nearest boundaries on either
// side with a shorter prefix (prefixes are distinct along any range, so "nearest
// shorter" is well defined), and it hangs below whichever of those two boundaries
// has the longer prefix -- the attach rule of the bottom-up pass.  A node is formed
// here when both nearest shorter boundaries are among the group's boundaries (the
// outer two included); a formed node or entry whose parent boundary is an outer one
// or an unformed slot is on the group's frontier.  One thread per slot finds both
// boundaries by scanning the prefix bytes, 8 per LDS read within its own word and 64
// per read beyond it through a table of word minima; the boxes then go up the formed forest by
// arrival counters, each merge taking the left child's box as the left operand,
// so every node is byte-identical to the one-node-at-a-time climb.

// High bit of each byte of w whose value is below x (bytes and x in [0, 127], x >= 1).
__device__ __forceinline__ uint64_t bytesBelow(uint64_t w, uint32_t x) {
    constexpr uint64_t kHigh = 0x8080808080808080ull;
    return ~((w | kHigh) - x * 0x0101010101010101ull) & kHigh;
}
// Bytes 0..k of a word (bits 0 .. 8k + 7).
__device__ __forceinline__ uint64_t bytesUpTo(int k) {
    return k == 7 ? ~0ull : ((1ull << (8 * k + 8)) - 1ull);
}
__device__ __forceinline__ int highByte(uint64_t m) {
    return (63 - __clzll(static_cast<long long>(m))) >> 3;
}
__device__ __forceinline__ int lowByte(uint64_t m) {
    return (__ffsll(static_cast<long long>(m)) - 1) >> 3;
}
// Largest i <= p with byte v[i] < x, or -1.  vm holds the minimum byte of each word
// of v, so a scan passes 64 prefixes per read after the first word.
__device__ __forceinline__ int lastBelow(const uint64_t* v, const uint64_t* vm, int p, uint32_t x) {
    const int w = p >> 3;
    const uint64_t m = bytesBelow(v[w], x) & bytesUpTo(p & 7);
    if (m != 0) return w * 8 + highByte(m);
    const int q = w - 1;
    if (q < 0) return -1;
    int mw = q >> 3;
    uint64_t mm = bytesBelow(vm[mw], x) & bytesUpTo(q & 7);
    while (mm == 0) {
        if (--mw < 0) return -1;
        mm = bytesBelow(vm[mw], x);
    }
    const int fw = mw * 8 + highByte(mm);
    return fw * 8 + highByte(bytesBelow(v[fw], x));
}
// Smallest i in [p, last] with byte v[i] < x, or -1 (bytes past last are padding
// that is never below x).
__device__ __forceinline__ int firstBelow(const uint64_t* v, const uint64_t* vm, int p, int last, uint32_t x) {
    const int w = p >> 3;
    const uint64_t m = bytesBelow(v[w], x) & (~0ull << ((p & 7) * 8));
    int i = -1;
    if (m != 0) {
        i = w * 8 + lowByte(m);
    } else {
        const int q = w + 1, wl = last >> 3;
        if (q > wl) return -1;
        int mw = q >> 3;
        const int mwl = wl >> 3;
        uint64_t mm = bytesBelow(vm[mw], x) & (~0ull << ((q & 7) * 8));
        while (mm == 0) {
            if (++mw > mwl) return -1;
            mm = bytesBelow(vm[mw], x);
        }
        const int fw = mw * 8 + lowByte(mm);
        if (fw > wl) return -1;
        i = fw * 8 + lowByte(bytesBelow(v[fw], x));
    }
    return i <= last ? i : -1;
}

// One group: the frontiers of `members` consecutive units (their staging from inStage,
// kFrontierCap entries each, their counts from inCounts) merged; the group's own
// frontier goes to outStageUnit / *outCountUnit unless top (the group makes the root).
template <int kUnits>
__device__ __forceinline__ void mergeGroup(
    const FrontierEntry* __restrict__ inStage, const int* __restrict__ inCounts, int members,
    FrontierEntry* __restrict__ outStageUnit, int* __restrict__ outCountUnit, bool top,
    const uint32_t* __restrict__ code, const uint32_t* __restrict__ order, int n,
    Node2* __restrict__ nodes, int* __restrict__ parent, int* __restrict__ arrive,
    int* __restrict__ height, int* __restrict__ farEnd, int frontierCap) {
    constexpr int kThreads = kUnits * kFrontierCap;  // one thread per entry; LDS node j: entry j,
                                                     // slot j at kThreads + j
    constexpr int kPreWords = (kThreads + 8) / 8 + 1;
    constexpr int kPreMinWords = (kPreWords + 7) / 8 + 1;
    __shared__ float sBox[2 * kThreads][6];
    __shared__ int sH[2 * kThreads];
    __shared__ int sId[2 * kThreads];
    __shared__ int sLo[kThreads], sHi[kThreads];  // leaf range of entry e
    __shared__ uint64_t sPre[kPreWords];  // byte i: prefix + 1 at boundary i - 1 (boundary -1 = left edge)
    __shared__ int4 sSlot[kThreads];      // formed slot: parent slot (-1: none here), child 0, child 1, formed
    __shared__ int sArrive[kThreads];
    __shared__ int sMark[kThreads];
    __shared__ int sWave[kThreads / 64 + 1];
    __shared__ uint64_t sPreMin[kPreMinWords];  // byte w: the smallest byte of sPre[w]
    __shared__ int sOff[kUnits + 1];
    __shared__ unsigned char sMember[kThreads];  // unit of entry e
    __shared__ int sBroken;
    const int t = threadIdx.x;
    if (t < 64) {  // wave 0: unit counts, their offsets, each entry's unit, the broken flag
        const int raw = t < members ? inCounts[t] : 0;
        const int own = raw < 0 ? 0 : (raw > kFrontierCap ? kFrontierCap : raw);
        int c = own;
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(c, d);
            if (t >= d) c += y;
        }
        const bool anyBroken = __any(raw < 0);
        if (t < members) sOff[t + 1] = c;
        if (t == 0) {
            sOff[0] = 0;
            sBroken = anyBroken ? 1 : 0;
        }
        for (int i = c - own; i < c; ++i) sMember[i] = static_cast<unsigned char>(t);
    }
    __syncthreads();
    const int count = sOff[members];
    const bool broken = sBroken != 0;
    NodeState s{};
    const bool entry = t < count;
    if (entry) {
        const int m = sMember[t];
        s = fromEntry(inStage[m * kFrontierCap + (t - sOff[m])]);
    }
    if (broken) {  // a unit below went global: so does the whole group
        if (entry) globalClimb(s, code, order, n, nodes, parent, arrive, height, farEnd);
        if (t == 0 && !top) *outCountUnit = -1;
        return;
    }

    // ---- entries and boundary prefixes into LDS ------------------------------------------
    auto* pre = reinterpret_cast<unsigned char*>(sPre);
    if (t < kPreWords) sPre[t] = 0x7F7F7F7F7F7F7F7Full;  // padding: never below any prefix
    if (t < kPreMinWords) sPreMin[t] = 0x7F7F7F7F7F7F7F7Full;
    sArrive[t] = 0;
    __syncthreads();
    if (entry) {
        float* b = sBox[t];
        b[0] = s.box.lo.x;
        b[1] = s.box.lo.y;
        b[2] = s.box.lo.z;
        b[3] = s.box.hi.x;
        b[4] = s.box.hi.y;
        b[5] = s.box.hi.z;
        sH[t] = s.h;
        sId[t] = s.id;
        sLo[t] = s.l;
        sHi[t] = s.r;
        pre[t + 1] = static_cast<unsigned char>(s.dr + 1);  // boundary t (slot t, or the right edge)
        if (t == 0) pre[0] = static_cast<unsigned char>(s.dl + 1);
    }
    __syncthreads();
    if (t <= (count >> 3)) {
        const uint64_t w = sPre[t];
        uint32_t mn = 0x7F;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t byte = static_cast<uint32_t>(w >> (8 * k)) & 0xFFu;
            mn = byte < mn ? byte : mn;
        }
        reinterpret_cast<unsigned char*>(sPreMin)[t] = static_cast<unsigned char>(mn);
    }
    __syncthreads();

    // ---- the slots: nearest shorter boundaries, parent, id ---------------------------------
    const int slots = count - 1;
    const bool isSlot = t < slots;
    int sa = -1, sb = -1;  // byte indices of the nearest shorter boundaries (boundary = index - 1)
    bool formed = false, slotRoot = false;
    int slotParent = -1, slotSide = 0;  // boundary index of the parent split (-1: the left edge)
    if (isSlot) {
        const uint32_t x = pre[t + 1];
        sa = lastBelow(sPre, sPreMin, t, x);
        sb = firstBelow(sPre, sPreMin, t + 2, count, x);
        formed = sa >= 0 && sb >= 0;
        if (formed) {
            const int dl = static_cast<int>(pre[sa]) - 1, dr = static_cast<int>(pre[sb]) - 1;
            const int lo = sLo[sa], hi = sHi[sb - 1];
            const bool left = dr > dl;
            slotRoot = lo == 0 && hi == n - 1;
            sId[kThreads + t] = slotRoot ? 0 : (left ? hi : lo);
            slotParent = left ? sb - 1 : sa - 1;
            slotSide = left ? 0 : 1;
        }
    }
    sSlot[t] = make_int4(-1, -1, -1, formed ? 1 : 0);
    __syncthreads();
    // a node's parent counts only if it is an inner slot formed here
    auto innerFormed = [&](int p) { return p >= 0 && p < slots && sSlot[p].w != 0; };
    int entryParent = -1, entrySide = 0;
    if (entry) {
        const bool left = pre[t + 1] > pre[t];  // dr > dl
        entryParent = left ? t : t - 1;
        entrySide = left ? 0 : 1;
        if (s.l == 0 && s.r == n - 1) entryParent = -1;  // the root (never staged)
    }
    const bool entryPend = entry && !innerFormed(entryParent) && !(s.l == 0 && s.r == n - 1);
    const bool slotUp = formed && !slotRoot && innerFormed(slotParent);
    const bool slotPend = formed && !slotRoot && !slotUp;
    if (entry && !entryPend && entryParent >= 0) {
        if (entrySide == 0) sSlot[entryParent].y = t;
        else sSlot[entryParent].z = t;
    }
    if (slotUp) {
        if (slotSide == 0) sSlot[slotParent].y = kThreads + t;
        else sSlot[slotParent].z = kThreads + t;
    }
    __syncthreads();
    if (formed) sSlot[t].x = slotUp ? slotParent : -1;
    __syncthreads();

    // ---- boxes and heights up the formed forest ---------------------------------------------
    if (entry && !entryPend && entryParent >= 0) {
        int me = t, p = entryParent, h = s.h;
        Box box = s.box;
        while (p >= 0) {
            const int4 tp = sSlot[p];  // fixed since the last barrier: read beside the arrival
            const int prior = __hip_atomic_fetch_add(&sArrive[p], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (prior == 0) break;
            const bool mineLeft = tp.y == me;
            const int sib = mineLeft ? tp.z : tp.y;
            const float* c = sBox[sib];
            const Box other{{c[0], c[1], c[2]}, {c[3], c[4], c[5]}};
            const int hs = sH[sib];
            box = mineLeft ? mergeChildren(box, other) : mergeChildren(other, box);
            h = (h > hs ? h : hs) + 1;
            float* d = sBox[kThreads + p];
            d[0] = box.lo.x;
            d[1] = box.lo.y;
            d[2] = box.lo.z;
            d[3] = box.hi.x;
            d[4] = box.hi.y;
            d[5] = box.hi.z;
            sH[kThreads + p] = h;
            me = kThreads + p;
            p = tp.x;
        }
    }
    __syncthreads();

    // ---- the formed nodes to memory -----------------------------------------------------------
    const int4 tp = sSlot[t];
    if (formed && tp.y >= 0 && tp.z >= 0) {  // (both children are always linked)
        const int pid = sId[kThreads + t];
        const int lid = sId[tp.y], rid = sId[tp.z];
        const float* b = sBox[kThreads + t];
        Node2 nd;
        nd.lo[0] = b[0];
        nd.lo[1] = b[1];
        nd.lo[2] = b[2];
        nd.hi[0] = b[3];
        nd.hi[1] = b[4];
        nd.hi[2] = b[5];
        nd.left = lid;
        nd.right = rid;
        nodes[pid] = nd;
        parent[lid] = pid;
        parent[rid] = pid;
        const int lo = sLo[sa], hi = sHi[sb - 1];
        farEnd[pid] = pid == lo ? hi : lo;
        if (pid == 0) height[0] = sH[kThreads + t];  // the tree depth
    }

    // ---- hand-up: stage the group's frontier in leaf order ------------------------------------
    if (top) {  // the whole tree: every node was formed (kept for safety)
        if (entryPend) globalClimb(s, code, order, n, nodes, parent, arrive, height, farEnd);
        if (slotPend) {
            const float* b = sBox[kThreads + t];
            const NodeState q{sLo[sa], sHi[sb - 1], sId[kThreads + t], sH[kThreads + t],
                              static_cast<int>(pre[sa]) - 1, static_cast<int>(pre[sb]) - 1,
                              {{b[0], b[1], b[2]}, {b[3], b[4], b[5]}}};
            globalClimb(q, code, order, n, nodes, parent, arrive, height, farEnd);
        }
        return;
    }
    // positions: an entry at its index, a slot node at its first entry (distinct)
    sMark[t] = 0;
    __syncthreads();
    if (entryPend) sMark[t] = 1;
    if (slotPend) sMark[sa] = 1;
    __syncthreads();
    const uint64_t bal = __ballot(sMark[t]);
    const int lane = t & 63;
    const int before = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) sWave[t >> 6] = __popcll(bal);
    __syncthreads();
    int prefix = 0;
    for (int w = 0; w < (t >> 6); ++w) prefix += sWave[w];
    sMark[t] = prefix + before;
    int total = 0;
    for (int w = 0; w < kThreads / 64; ++w) total += sWave[w];
    __syncthreads();
    if (total > frontierCap) {
        if (entryPend) globalClimb(s, code, order, n, nodes, parent, arrive, height, farEnd);
        if (slotPend) {
            const float* b = sBox[kThreads + t];
            const NodeState q{sLo[sa], sHi[sb - 1], sId[kThreads + t], sH[kThreads + t],
                              static_cast<int>(pre[sa]) - 1, static_cast<int>(pre[sb]) - 1,
                              {{b[0], b[1], b[2]}, {b[3], b[4], b[5]}}};
            globalClimb(q, code, order, n, nodes, parent, arrive, height, farEnd);
        }
        if (t == 0) *outCountUnit = -1;
    } else {
        if (entryPend) outStageUnit[sMark[t]] = toEntry(s);
        if (slotPend) {
            const float* b = sBox[kThreads + t];
            outStageUnit[sMark[sa]] =
                FrontierEntry{sLo[sa], sHi[sb - 1], sId[kThreads + t], sH[kThreads + t],
                              static_cast<int>(pre[sa]) - 1, static_cast<int>(pre[sb]) - 1,
                              {b[0], b[1], b[2], b[3], b[4], b[5]}};
        }
        if (t == 0) *outCountUnit = total;
    }
}

// Level L >= 1 of the hand-up of one tree: workgroup g merges units
// [g kMergeUnits, g kMergeUnits + kMergeUnits) of level L-1.
__global__ __launch_bounds__(kMergeBlock) void lbvhMergeKernel(
    const FrontierEntry* __restrict__ inStage, const int* __restrict__ inCounts, int inUnits,
    FrontierEntry* __restrict__ outStage, int* __restrict__ outCounts, bool top,
    const uint32_t* __restrict__ code, const uint32_t* __restrict__ order, int n,
    Node2* __restrict__ nodes, int* __restrict__ parent, int* __restrict__ arrive,
    int* __restrict__ height, int* __restrict__ farEnd, int frontierCap) {
    const int g = static_cast<int>(blockIdx.x);
    const int first = g * kMergeUnits;
    const int members = inUnits - first < kMergeUnits ? inUnits - first : kMergeUnits;
    mergeGroup<kMergeUnits>(inStage + static_cast<size_t>(first) * kFrontierCap, inCounts + first, members,
               top ? nullptr : outStage + static_cast<size_t>(g) * kFrontierCap, top ? nullptr : outCounts + g, top,
               code, order, n, nodes, parent, arrive, height, farEnd, frontierCap);
}

// Hand-up levels for n leaves: windows of kBlock leaves, then groups of kMergeUnits
// units until one unit is left.  Returns the sizes of the staging (entries) and of
// the counts.
static EmitLevels emitLevels(int n, int* stageEntries, int* countInts) {
    EmitLevels lv{};
    int u = blocksFor(n, kBlock);
    int stage = 0, cnt = 0, level = 0;
    lv.units[0] = u;
    while (u > 1) {
        if (level + 1 >= kEmitMaxLevels) throw ArgError("build: too many primitives for the hand-up levels");
        lv.stageOff[level] = stage;
        stage += u * kFrontierCap;
        lv.countOff[level] = cnt;
        cnt += u;
        u = (u + kMergeUnits - 1) / kMergeUnits;
        ++level;
        lv.units[level] = u;
    }
    lv.levels = level;
    *stageEntries = stage > 0 ? stage : 1;
    *countInts = cnt > 0 ? cnt : 1;
    return lv;
}

// ---- batched (segmented) LBVH: many independent scenes per launch -----------------------
// Scene s owns items [primOff[s], primOff[s+1]) (n_s >= 2), nodes from 2 primOff[s] - s
// (2 n_s - 1 of them) and internal-node slots (arrival counters, heights) from
// primOff[s] - s.  Each tree is built by the plain build's own device code on local
// ids through offset pointers, so it equals the plain build of that scene alone,
// byte for byte.  The per-scene Morton order is a least-significant-digit sort:
// first by code, then (stably) by scene id.

__global__ void batchSceneIdKernel(const int* __restrict__ primOff, int* __restrict__ sceneId) {
    const int sc = blockIdx.x;
    for (int i = primOff[sc] + threadIdx.x; i < primOff[sc + 1]; i += kBlock) sceneId[i] = sc;
}

__global__ void batchBoundsInitKernel(uint32_t* __restrict__ cb, int scenes) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= scenes * 6) return;
    cb[i] = (i % 6) < 3 ? orderedBits(kInf) : orderedBits(-kInf);
}

// Per-scene centroid bounds.  Each wave takes kBoundsSegs consecutive 64-item
// segments (coalesced loads); a segment inside one scene is reduced across the wave
// (DPP, on ordered bits: exact and order-free) into the wave's running bounds, which
// go out with 6 atomics when the scene changes or the wave ends; a segment that
// straddles scenes falls back to per-lane atomics.
constexpr int kBoundsSegs = 16;

template <typename L>
__global__ __launch_bounds__(kBlock) void batchBoundsKernel(const typename L::Item* __restrict__ items, int n,
                                                            const int* __restrict__ sceneId,
                                                            uint32_t* __restrict__ cb) {
    const int lane = threadIdx.x & 63;
    const long long first = ((static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x) >> 6) * 64 * kBoundsSegs;
    int cur = -1;  // the scene of the running bounds (the same in every lane)
    uint32_t acc[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    auto flush = [&]() {
        if (cur >= 0 && lane == 0) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                atomicMin(&cb[cur * 6 + k], acc[k]);
                atomicMax(&cb[cur * 6 + 3 + k], acc[3 + k]);
            }
        }
    };
    for (int j = 0; j < kBoundsSegs; ++j) {
        const long long seg = first + static_cast<long long>(j) * 64;
        if (seg >= n) break;
        const long long i = seg + lane;
        const bool live = i < n;
        const int sc = live ? sceneId[i] : -1;
        const int s0 = __builtin_amdgcn_readlane(sc, 0);  // lane 0 is live
        uint32_t v[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
        if (live) {
            const V3 c = L::centroid(items[i]);
            v[0] = v[3] = orderedBits(c.x);
            v[1] = v[4] = orderedBits(c.y);
            v[2] = v[5] = orderedBits(c.z);
        }
        if (__all(!live || sc == s0)) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                v[k] = reduce64<MinU>(v[k]);
                v[3 + k] = reduce64<MaxU>(v[3 + k]);
            }
            if (s0 != cur) {
                flush();
                cur = s0;
#pragma unroll
                for (int k = 0; k < 6; ++k) acc[k] = v[k];
            } else {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    acc[k] = MinU::op(acc[k], v[k]);
                    acc[3 + k] = MaxU::op(acc[3 + k], v[3 + k]);
                }
            }
        } else if (live) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                atomicMin(&cb[sc * 6 + k], v[k]);
                atomicMax(&cb[sc * 6 + 3 + k], v[3 + k]);
            }
        }
    }
    flush();
}

template <typename L>
__global__ __launch_bounds__(kBlock) void batchMortonKernel(const typename L::Item* __restrict__ items, int n,
                                                            const int* __restrict__ sceneId,
                                                            const uint32_t* __restrict__ cb,
                                                            uint32_t* __restrict__ codes,
                                                            uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[4][256];
    for (int i = threadIdx.x; i < 4 * 256; i += kBlock) (&h[0][0])[i] = 0u;
    __syncthreads();
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const int sc = sceneId[i];
        const uint32_t* b = cb + sc * 6;
        const float lo0 = fromOrderedBits(b[0]), lo1 = fromOrderedBits(b[1]), lo2 = fromOrderedBits(b[2]);
        const float ex0 = fromOrderedBits(b[3]) - lo0, ex1 = fromOrderedBits(b[4]) - lo1,
                    ex2 = fromOrderedBits(b[5]) - lo2;
        const V3 c = L::centroid(items[i]);
        const uint32_t code = spread10(quantize1023(c.x, lo0, ex0)) |
                              (spread10(quantize1023(c.y, lo1, ex1)) << 1) |
                              (spread10(quantize1023(c.z, lo2, ex2)) << 2);
        codes[i] = code;
        atomicAdd(&h[0][code & 255u], 1u);
        atomicAdd(&h[1][(code >> 8) & 255u], 1u);
        atomicAdd(&h[2][(code >> 16) & 255u], 1u);
        atomicAdd(&h[3][code >> 24], 1u);
    }
    __syncthreads();
    flushHistogram(&h[0][0], 4 * 256, hist);
}

__global__ void batchGatherKernel(const uint32_t* __restrict__ idx, const int* __restrict__ src, int n,
                                  uint32_t* __restrict__ dst) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) dst[i] = static_cast<uint32_t>(src[idx[i]]);
}

// Scene-local codes and indices of the sorted order; also returns every scene's
// arrival counters to zero (the hand-up's device-wide path expects them so).
__global__ void batchLocalKernel(const uint32_t* __restrict__ order2, const uint32_t* __restrict__ s2,
                                 const uint32_t* __restrict__ codes, const int* __restrict__ primOff, int n,
                                 uint32_t* __restrict__ codes2, uint32_t* __restrict__ idxLocal,
                                 int* __restrict__ arrive) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = order2[i];
    const int sc = static_cast<int>(s2[i]);
    const int ps = primOff[sc];
    codes2[i] = codes[o];
    idxLocal[i] = o - static_cast<uint32_t>(ps);
    if (i - ps < primOff[sc + 1] - ps - 1) arrive[i - sc] = 0;  // scene slot i - ps of arrive + (ps - sc)
}

// The last scene sc with off[sc] <= w (off: S + 1 ascending offsets; a scene with an
// empty range is passed over, so the scene found holds w).
__device__ __forceinline__ int sceneOf(const int* __restrict__ off, int scenes, int w) {
    int lo = 0, hi = scenes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= w) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Hierarchy, leaf gather and boxes of every scene: workgroup w runs the plain build's
// window pass (emitWindow) on window w - winOff[sc] of scene sc, on that scene's
// arrays; a scene of one window makes its root there, the others hand their frontiers
// up through batchMergeKernel.
template <typename L>
__global__ __launch_bounds__(kBlock) void batchEmitKernel(const typename L::Item* __restrict__ items,
                                                          const uint32_t* __restrict__ codes2,
                                                          const uint32_t* __restrict__ idxLocal,
                                                          const int* __restrict__ primOff,
                                                          const int* __restrict__ winOff, int scenes,
                                                          typename L::Item* __restrict__ sorted,
                                                          int* __restrict__ primIds, Node2* __restrict__ nodes,
                                                          int* __restrict__ parent, int* __restrict__ arrive,
                                                          int* __restrict__ height, int* __restrict__ farEnd,
                                                          FrontierEntry* __restrict__ stage, int* __restrict__ counts,
                                                          int frontierCap) {
    const int w = static_cast<int>(blockIdx.x);
    const int sc = sceneOf(winOff, scenes, w);
    const int ps = primOff[sc], ns = primOff[sc + 1] - ps;
    emitWindow<L>(items + ps, codes2 + ps, idxLocal + ps, ns, sorted + ps, primIds + ps, nodes + (2 * ps - sc),
                  parent + (2 * ps - sc), arrive + (ps - sc), height + (ps - sc), farEnd + (ps - sc),
                  stage + static_cast<size_t>(w) * kFrontierCap, counts + w, frontierCap, w - winOff[sc],
                  winOff[sc + 1] - winOff[sc] == 1);
}

// One hand-up level of every scene that still has more than one unit: workgroup g is
// group g - grpOff[sc] of scene sc (grpOff: this level's groups, scene after scene);
// its members are the scene's units from unitOff[sc] on, in the level below's
// numbering; a scene's only group makes its root.
template <int kUnits>
__global__ __launch_bounds__(kUnits * kFrontierCap) void batchMergeKernel(
    const FrontierEntry* __restrict__ inStage, const int* __restrict__ inCounts, const int* __restrict__ unitOff,
    const int* __restrict__ grpOff, int scenes, FrontierEntry* __restrict__ outStage, int* __restrict__ outCounts,
    const uint32_t* __restrict__ codes2, const uint32_t* __restrict__ idxLocal, const int* __restrict__ primOff,
    Node2* __restrict__ nodes, int* __restrict__ parent, int* __restrict__ arrive, int* __restrict__ height,
    int* __restrict__ farEnd, int frontierCap) {
    const int g = static_cast<int>(blockIdx.x);
    const int sc = sceneOf(grpOff, scenes, g);
    const int first = (g - grpOff[sc]) * kUnits;
    const int unitsIn = unitOff[sc + 1] - unitOff[sc];
    const int members = unitsIn - first < kUnits ? unitsIn - first : kUnits;
    const bool top = grpOff[sc + 1] - grpOff[sc] == 1;
    const int ps = primOff[sc], ns = primOff[sc + 1] - ps;
    mergeGroup<kUnits>(inStage + static_cast<size_t>(unitOff[sc] + first) * kFrontierCap, inCounts + unitOff[sc] + first,
               members, top ? nullptr : outStage + static_cast<size_t>(g) * kFrontierCap,
               top ? nullptr : outCounts + g, top, codes2 + ps, idxLocal + ps, ns, nodes + (2 * ps - sc),
               parent + (2 * ps - sc), arrive + (ps - sc), height + (ps - sc), farEnd + (ps - sc), frontierCap);
}

__global__ void batchDepthKernel(const int* __restrict__ height, const int* __restrict__ primOff, int scenes,
                                 int* __restrict__ depth) {
    const int sc = blockIdx.x * kBlock + threadIdx.x;
    if (sc < scenes) depth[sc] = height[primOff[sc] - sc];
}

}  // namespace

template <typename L>
static void buildLbvhT(BvhState& b, const typename L::Item* dSrc, int n, hipStream_t s,
                       BuildTimingsC* t) {
    using Item = typename L::Item;
    b.n = n;
    b.nodeCount = 2 * n - 1;
    b.nodes.reserve(static_cast<size_t>(b.nodeCount));
    b.items.reserve(static_cast<size_t>(n) * sizeof(Item));
    b.primIds.reserve(static_cast<size_t>(n));
    b.primSlot.reserve(static_cast<size_t>(n));  // filled on first use (refits, adjoints)
    b.parent.reserve(static_cast<size_t>(b.nodeCount));
    b.arrive.reserve(static_cast<size_t>(n - 1));
    b.height.reserve(static_cast<size_t>(n - 1));
    b.farEnd.reserve(static_cast<size_t>(n - 1));
    b.sortWs.reserve(static_cast<size_t>(n));
    b.counters.reserve(static_cast<size_t>(kMaxBoundsBlocks) * kPartStride);
    int stageEntries = 0, countInts = 0;
    const EmitLevels lv = emitLevels(n, &stageEntries, &countInts);
    b.emitStage.reserve(static_cast<size_t>(stageEntries) * (sizeof(FrontierEntry) / sizeof(int)));
    b.emitCounts.reserve(static_cast<size_t>(countInts));

    b.keys.reserve(static_cast<size_t>(n));
    b.keysSorted.reserve(static_cast<size_t>(n));
    b.order.reserve(static_cast<size_t>(n));
    uint32_t* codes = b.keys.p;
    uint32_t* codesSorted = b.keysSorted.p;
    uint32_t* order = b.order.p;
    uint32_t* cb = reinterpret_cast<uint32_t*>(b.counters.p);

    EventSet<6> ev;
    ev.record(0, s);

    const int tiles = sortTiles(static_cast<size_t>(n));
    const long long descWords = 4LL * tiles * 256;
    const BuildZero zero{b.sortWs.hist.p, b.sortWs.desc.p, descWords, b.sortWs.tickets.p, b.arrive.p, n - 1};
    // two workgroups per CU, 16 items per thread: enough loads in flight to stream,
    // few enough workgroups that the histogram atomics stay cheap
    int grid = blocksFor(n, kBlock * 4 * kUnroll);
    if (grid > 512) grid = 512;
    int parts = grid;
    if (L::kKind == 0 && (reinterpret_cast<uintptr_t>(dSrc) & 15) == 0) {
        // staged triangle bounds: four workgroups per CU (36 KB of LDS each); 2M
        // triangles 14.4 -> 13.3 us
        parts = blocksFor(n, kStageTris);
        if (parts > kMaxBoundsBlocks) parts = kMaxBoundsBlocks;
        const Tri* tris = reinterpret_cast<const Tri*>(dSrc);
        triBoundsStagedKernel<<<parts, kBlock, 0, s>>>(tris, n, cb, zero);
    } else {
        centroidBoundsKernel<L><<<grid, kBlock, 0, s>>>(dSrc, n, cb, zero);
    }
    // (a staged variant of this pass measured slower: 21.4 against 18.2 us on 2M
    // triangles; the codes are written per item, so the lane-per-item loads suffice)
    mortonHistKernel<L><<<grid, kBlock, 0, s>>>(dSrc, n, cb, parts, codes, b.sortWs.hist.p);
    hipCheck(hipGetLastError(), "morton");
    // phase events only for a caller that asks for the timings with phases on: an event
    // between two kernels costs ~12 us of idle device time here (2M triangles: 0.318
    // against 0.342 ms from outside with the two)
    const bool phases = t && gBuildPhaseTiming >= 1;
    if (phases) ev.record(1, s);

    sortPairs(b.sortWs, codes, nullptr, codesSorted, order, static_cast<size_t>(n), 4, s);
    if (phases) ev.record(2, s);

    // hierarchy, leaf gather and boxes: the windows, then one merge kernel per level
    // (no separate Karras kernel)
    auto* stage = reinterpret_cast<FrontierEntry*>(b.emitStage.p);
    lbvhEmitKernel<L><<<lv.units[0], kBlock, 0, s>>>(
        dSrc, codesSorted, order, n, reinterpret_cast<Item*>(b.items.p), b.primIds.p, b.nodes.p,
        b.parent.p, b.arrive.p, b.height.p, b.farEnd.p, stage + lv.stageOff[0], b.emitCounts.p + lv.countOff[0],
        gFrontierCap);
    hipCheck(hipGetLastError(), "lbvh emit");
    for (int level = 1; level <= lv.levels; ++level) {
        const bool top = level == lv.levels;
        lbvhMergeKernel<<<lv.units[level], kMergeBlock, 0, s>>>(
            stage + lv.stageOff[level - 1], b.emitCounts.p + lv.countOff[level - 1], lv.units[level - 1],
            top ? nullptr : stage + lv.stageOff[level], top ? nullptr : b.emitCounts.p + lv.countOff[level], top,
            codesSorted, order, n, b.nodes.p, b.parent.p, b.arrive.p, b.height.p, b.farEnd.p, gFrontierCap);
        hipCheck(hipGetLastError(), "lbvh merge");
    }
    ev.record(4, s);

    int depth = 0;
    hipCheck(hipMemcpyAsync(&depth, b.height.p, sizeof(int), hipMemcpyDeviceToHost, s), "depth");
    ev.record(5, s);
    ev.sync(5);
    b.depth = depth;

    if (t) {
        t->mortonMs = phases ? ev.ms(0, 1) : 0.0;
        t->sortMs = phases ? ev.ms(1, 2) : 0.0;
        t->treeMs = 0.0;  // the hierarchy is formed with the boxes, in the same kernels
        t->scatterMs = 0.0;
        t->refitMs = phases ? ev.ms(2, 4) : 0.0;
        t->sahBinMs = t->sahSplitMs = t->sahPartMs = 0.0;
        t->totalMs = ev.ms(0, 4);
    }
    b.primSlotValid = false;
    b.depthValid = false;
    b.wideValid = false;
    b.wideCount = 0;
}

void setFrontierCap(int cap) {
    if (cap < 0 || cap > kFrontierCap) throw ArgError("frontier capacity must be in [0, 32]");
    gFrontierCap = cap;
}

void buildLbvhDevice(BvhState& b, int kind, const void* dItems, int n, hipStream_t s,
                     BuildTimingsC* t) {
    if (n < 2) throw ArgError("build: at least 2 primitives required");
    b.kind = kind;
    b.builder = kBuildLbvh;
    switch (kind) {
        case kPrimTri:
            buildLbvhT<TriLeaf>(b, static_cast<const Tri*>(dItems), n, s, t);
            break;
        case kPrimPoint:
            buildLbvhT<PointLeaf>(b, static_cast<const V3*>(dItems), n, s, t);
            break;
        case kPrimGauss:
            buildLbvhT<GaussLeaf>(b, static_cast<const Gaussian*>(dItems), n, s, t);
            break;
        default:
            throw ArgError("build: unknown primitive kind");
    }
}

// Leaf gather, leaf nodes and bottom-up boxes for a topology built elsewhere (the
// SAH builder): internal nodes' left/right, parent[] (root -1) and zeroed arrival
// counters must be in place; order[k] is the scene index of leaf slot k.
void finishTopology(BvhState& b, const void* dSrc, const uint32_t* order, hipStream_t s, const int* gate) {
    const int n = b.n;
    const int grid = blocksFor(n, kBlock);
    switch (b.kind) {
        case kPrimTri:
            leafUpLocalKernel<TriLeaf, true><<<grid, kBlock, 0, s>>>(
                static_cast<const Tri*>(dSrc), order, n, reinterpret_cast<Tri*>(b.items.p), b.primIds.p,
                b.nodes.p, b.parent.p, b.arrive.p, b.height.p, b.farEnd.p, gate);
            break;
        case kPrimPoint:
            leafUpLocalKernel<PointLeaf, true><<<grid, kBlock, 0, s>>>(
                static_cast<const V3*>(dSrc), order, n, reinterpret_cast<V3*>(b.items.p), b.primIds.p,
                b.nodes.p, b.parent.p, b.arrive.p, b.height.p, b.farEnd.p, gate);
            break;
        default:
            leafUpLocalKernel<GaussLeaf, true><<<grid, kBlock, 0, s>>>(
                static_cast<const Gaussian*>(dSrc), order, n, reinterpret_cast<Gaussian*>(b.items.p),
                b.primIds.p, b.nodes.p, b.parent.p, b.arrive.p, b.height.p, b.farEnd.p, gate);
            break;
    }
    hipCheck(hipGetLastError(), "finish topology");
}

// Refit of every leaf from the leaf-ordered items already in b.items: the same
// bottom-up pass without the gather (arrival counters are zero between kernels).
void propagateAll(BvhState& b, hipStream_t s) {
    const int n = b.n;
    const int grid = blocksFor(n, kBlock);
    switch (b.kind) {
        case kPrimTri:
            leafUpLocalKernel<TriLeaf, false><<<grid, kBlock, 0, s>>>(
                nullptr, nullptr, n, reinterpret_cast<Tri*>(b.items.p), nullptr, b.nodes.p,
                b.parent.p, b.arrive.p, b.height.p, b.farEnd.p, nullptr);
            break;
        case kPrimPoint:
            leafUpLocalKernel<PointLeaf, false><<<grid, kBlock, 0, s>>>(
                nullptr, nullptr, n, reinterpret_cast<V3*>(b.items.p), nullptr, b.nodes.p,
                b.parent.p, b.arrive.p, b.height.p, b.farEnd.p, nullptr);
            break;
        default:
            leafUpLocalKernel<GaussLeaf, false><<<grid, kBlock, 0, s>>>(
                nullptr, nullptr, n, reinterpret_cast<Gaussian*>(b.items.p), nullptr, b.nodes.p,
                b.parent.p, b.arrive.p, b.height.p, b.farEnd.p, nullptr);
            break;
    }
    hipCheck(hipGetLastError(), "propagate");
}

// ---- depth by pointer jumping --------------------------------------------------------
namespace {
// Each node walks towards the root by jumping through ancestors' (anc, dist)
// pairs; pairs are published as one 64-bit word so a reader always sees a
// consistent jump.  Every step strictly shortens the remaining path, so the
// loop ends within depth iterations (typically ~log2(depth)).
__global__ void depthJumpKernel(const int* __restrict__ parent, int nodeCount,
                                unsigned long long* __restrict__ jump, int* __restrict__ depth) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nodeCount) return;
    int anc = parent[i];
    int dist = anc < 0 ? 0 : 1;
    while (anc > 0) {
        const unsigned long long pj = __hip_atomic_load(&jump[anc], __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
        const int a2 = static_cast<int>(pj & 0xffffffffu);
        const int d2 = static_cast<int>(pj >> 32);
        dist += d2;
        anc = a2;
        const unsigned long long mine = (static_cast<unsigned long long>(static_cast<uint32_t>(dist)) << 32) |
                                        static_cast<uint32_t>(anc);
        __hip_atomic_store(&jump[i], mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    depth[i] = dist;
}

__global__ void depthJumpInitKernel(const int* __restrict__ parent, int nodeCount,
                                    unsigned long long* __restrict__ jump) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nodeCount) return;
    const int p = parent[i];
    // Root: (anc 0, dist 0) so anything that reaches it stops there.
    const int anc = p < 0 ? 0 : p;
    const int dist = p < 0 ? 0 : 1;
    jump[i] = (static_cast<unsigned long long>(static_cast<uint32_t>(dist)) << 32) |
              static_cast<uint32_t>(anc);
}
}  // namespace

void computeDepth(BvhState& b, hipStream_t s) {
    if (b.depthValid) return;
    const int nc = b.nodeCount;
    b.depthArr.reserve(static_cast<size_t>(nc));
    b.jump.reserve(static_cast<size_t>(nc));
    depthJumpInitKernel<<<blocksFor(nc, kBlock), kBlock, 0, s>>>(b.parent.p, nc, b.jump.p);
    depthJumpKernel<<<blocksFor(nc, kBlock), kBlock, 0, s>>>(b.parent.p, nc, b.jump.p, b.depthArr.p);
    hipCheck(hipGetLastError(), "depth");
    b.depthValid = true;
}


// ---- batched build ---------------------------------------------------------------------------

template <typename L>
static void batchOrderT(const typename L::Item* items, int n, const int* sceneId, int scenes, SortWorkspace& ws,
                        uint32_t* cb, uint32_t* codes, uint32_t* tmpA, uint32_t* tmpB, uint32_t* s2,
                        uint32_t* order2, hipStream_t s) {
    batchBoundsInitKernel<<<blocksFor(static_cast<long long>(scenes) * 6, kBlock), kBlock, 0, s>>>(cb, scenes);
    batchBoundsKernel<L><<<blocksFor(n, kBlock * kBoundsSegs), kBlock, 0, s>>>(items, n, sceneId, cb);
    sortClear(ws, static_cast<size_t>(n), 4, s);
    batchMortonKernel<L><<<blocksFor(n, kBlock), kBlock, 0, s>>>(items, n, sceneId, cb, codes, ws.hist.p);
    hipCheck(hipGetLastError(), "batch morton");
    // by code: tmpA = sorted codes (unused), tmpB = order1
    sortPairs(ws, codes, nullptr, tmpA, tmpB, static_cast<size_t>(n), 4, s);
    // then stably by scene: keys s1 = scene of order1
    batchGatherKernel<<<blocksFor(n, kBlock), kBlock, 0, s>>>(tmpB, sceneId, n, tmpA);
    sortClear(ws, static_cast<size_t>(n), 2, s);
    sortHistogram(ws, tmpA, static_cast<size_t>(n), 2, s);
    sortPairs(ws, tmpA, tmpB, s2, order2, static_cast<size_t>(n), 2, s);
}

void batchMortonOrder(int kind, const void* items, int n, const int* sceneId, int scenes, SortWorkspace& ws,
                      uint32_t* cb, uint32_t* codes, uint32_t* tmpA, uint32_t* tmpB, uint32_t* sceneSorted,
                      uint32_t* order, hipStream_t s) {
    if (n <= 0) return;
    switch (kind) {
        case kPrimTri:
            batchOrderT<TriLeaf>(static_cast<const Tri*>(items), n, sceneId, scenes, ws, cb, codes, tmpA, tmpB,
                                 sceneSorted, order, s);
            break;
        case kPrimPoint:
            batchOrderT<PointLeaf>(static_cast<const V3*>(items), n, sceneId, scenes, ws, cb, codes, tmpA, tmpB,
                                   sceneSorted, order, s);
            break;
        default:
            batchOrderT<GaussLeaf>(static_cast<const Gaussian*>(items), n, sceneId, scenes, ws, cb, codes, tmpA,
                                   tmpB, sceneSorted, order, s);
            break;
    }
}

void batchSceneIds(const int* primOff, int scenes, int* sceneId, hipStream_t s) {
    if (scenes <= 0) return;
    batchSceneIdKernel<<<scenes, kBlock, 0, s>>>(primOff, sceneId);
    hipCheck(hipGetLastError(), "batch scene ids");
}

template <typename L>
static void buildBatchT(BatchState& b, const typename L::Item* dSrc, hipStream_t s, BuildTimingsC* t) {
    using Item = typename L::Item;
    const int n = b.total, S = b.scenes;
    b.nodes.alloc(static_cast<size_t>(2 * n - S));
    b.items.alloc(static_cast<size_t>(n) * sizeof(Item));
    b.primIds.alloc(static_cast<size_t>(n));
    b.parent.alloc(static_cast<size_t>(2 * n - S));
    b.arrive.alloc(static_cast<size_t>(n - S));
    b.height.alloc(static_cast<size_t>(n - S));
    b.depth.alloc(static_cast<size_t>(S));
    b.sceneId.alloc(static_cast<size_t>(n));
    b.ws.reserve(static_cast<size_t>(n));
    DevBuf<uint32_t> cb(static_cast<size_t>(S) * 6), codes(n), tmpA(n), tmpB(n), s2(n), order2(n), codes2(n),
        idxLocal(n);
    DevBuf<int> farEnd(static_cast<size_t>(n - S)), winOff(static_cast<size_t>(S) + 1);
    // each scene's windows of kBlock leaf slots, numbered scene after scene
    std::vector<int> winOffH(static_cast<size_t>(S) + 1, 0);
    for (int sc = 0; sc < S; ++sc)
        winOffH[sc + 1] = winOffH[sc] + blocksFor(b.primOffH[sc + 1] - b.primOffH[sc], kBlock);
    hipCheck(hipMemcpyAsync(winOff.p, winOffH.data(), winOffH.size() * sizeof(int), hipMemcpyHostToDevice, s),
             "batch windows");
    // hand-up levels per scene: level 0 = the windows; at level L a scene with u > 1
    // units of level L-1 has ceil(u / kBatchMergeUnits) groups (one: it makes the root).
    // Offsets of every level, scene after scene, in one array; unit numbering per level.
    std::vector<int> lvOffH;             // (S + 1) ints per level
    std::vector<size_t> lvOff;           // where level L's offsets start in lvOffH
    std::vector<int> lvTotal, lvUnitBase;  // units (groups) per level and their first staging unit
    lvOff.push_back(0);
    lvOffH.insert(lvOffH.end(), winOffH.begin(), winOffH.end());
    lvTotal.push_back(winOffH[S]);
    lvUnitBase.push_back(0);
    for (int level = 1;; ++level) {
        const int* prev = lvOffH.data() + lvOff[level - 1];
        std::vector<int> off(static_cast<size_t>(S) + 1, 0);
        for (int sc = 0; sc < S; ++sc) {
            const int u = prev[sc + 1] - prev[sc];
            off[sc + 1] = off[sc] + (u > 1 ? (u + kBatchMergeUnits - 1) / kBatchMergeUnits : 0);
        }
        if (off[S] == 0) break;
        if (level >= kEmitMaxLevels) throw ArgError("batch build: too many primitives for the hand-up levels");
        // (a scene whose only group makes its root counts one unit at this level and
        // so has no group at the next)
        lvOff.push_back(lvOffH.size());
        lvOffH.insert(lvOffH.end(), off.begin(), off.end());
        lvTotal.push_back(off[S]);
        lvUnitBase.push_back(lvUnitBase.back() + lvTotal[lvTotal.size() - 2]);
    }
    DevBuf<int> lvOffDev(lvOffH.size());
    hipCheck(hipMemcpyAsync(lvOffDev.p, lvOffH.data(), lvOffH.size() * sizeof(int), hipMemcpyHostToDevice, s),
             "batch hand-up levels");
    const int stageUnits = lvUnitBase.back() + lvTotal.back();
    DevBuf<int> bstage(static_cast<size_t>(stageUnits) * kFrontierCap * (sizeof(FrontierEntry) / sizeof(int)));
    DevBuf<int> bcounts(static_cast<size_t>(stageUnits));

    EventSet<5> ev;
    ev.record(0, s);
    batchSceneIds(b.primOff.p, S, b.sceneId.p, s);
    batchOrderT<L>(dSrc, n, b.sceneId.p, S, b.ws, cb.p, codes.p, tmpA.p, tmpB.p, s2.p, order2.p, s);
    const bool phases = t && gBuildPhaseTiming >= 1;  // (events between kernels idle the device)
    if (phases) ev.record(1, s);
    batchLocalKernel<<<blocksFor(n, kBlock), kBlock, 0, s>>>(order2.p, s2.p, codes.p, b.primOff.p, n, codes2.p,
                                                             idxLocal.p, b.arrive.p);
    hipCheck(hipGetLastError(), "batch local");
    if (phases) ev.record(2, s);
    // hierarchy, leaves and boxes: every scene's windows, then the hand-up levels
    auto* stage = reinterpret_cast<FrontierEntry*>(bstage.p);
    batchEmitKernel<L><<<winOffH[S], kBlock, 0, s>>>(dSrc, codes2.p, idxLocal.p, b.primOff.p, winOff.p, S,
                                                     reinterpret_cast<Item*>(b.items.p), b.primIds.p, b.nodes.p,
                                                     b.parent.p, b.arrive.p, b.height.p, farEnd.p, stage, bcounts.p,
                                                     gFrontierCap);
    hipCheck(hipGetLastError(), "batch emit");
    for (size_t level = 1; level < lvOff.size(); ++level) {
        const int groups = lvTotal[level];
        if (groups == 0) break;
        batchMergeKernel<kBatchMergeUnits><<<groups, kBatchMergeUnits * kFrontierCap, 0, s>>>(
            stage + static_cast<size_t>(lvUnitBase[level - 1]) * kFrontierCap, bcounts.p + lvUnitBase[level - 1],
            lvOffDev.p + lvOff[level - 1], lvOffDev.p + lvOff[level], S,
            stage + static_cast<size_t>(lvUnitBase[level]) * kFrontierCap, bcounts.p + lvUnitBase[level], codes2.p,
            idxLocal.p, b.primOff.p, b.nodes.p, b.parent.p, b.arrive.p, b.height.p, farEnd.p, gFrontierCap);
        hipCheck(hipGetLastError(), "batch merge");
    }
    ev.record(3, s);
    batchDepthKernel<<<blocksFor(S, kBlock), kBlock, 0, s>>>(b.height.p, b.primOff.p, S, b.depth.p);
    std::vector<int> d(static_cast<size_t>(S));
    hipCheck(hipMemcpyAsync(d.data(), b.depth.p, sizeof(int) * d.size(), hipMemcpyDeviceToHost, s), "batch depth");
    ev.record(4, s);
    ev.sync(4);
    b.maxDepth = 0;
    for (int v : d) b.maxDepth = v > b.maxDepth ? v : b.maxDepth;
    if (t) {
        t->mortonMs = phases ? ev.ms(0, 1) : 0.0;
        t->sortMs = 0.0;
        t->treeMs = phases ? ev.ms(1, 2) : 0.0;
        t->scatterMs = 0.0;
        t->refitMs = phases ? ev.ms(2, 3) : 0.0;
        t->sahBinMs = t->sahSplitMs = t->sahPartMs = 0.0;
        t->totalMs = ev.ms(0, 3);
    }
}

void buildBatchDevice(BatchState& b, int kind, const void* dItems, hipStream_t s, BuildTimingsC* t) {
    b.kind = kind;
    switch (kind) {
        case kPrimTri: buildBatchT<TriLeaf>(b, static_cast<const Tri*>(dItems), s, t); break;
        case kPrimPoint: buildBatchT<PointLeaf>(b, static_cast<const V3*>(dItems), s, t); break;
        case kPrimGauss: buildBatchT<GaussLeaf>(b, static_cast<const Gaussian*>(dItems), s, t); break;
        default: throw ArgError("batch build: unknown primitive kind");
    }
}

}  // namespace qr
