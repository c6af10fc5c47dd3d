// Tree maintenance: 4-wide collapse, refit (full and dirty-only), structural
// validation and the node-array hash.
#include "qr_bvh.h"

namespace qr {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ void stAgent(float* p, float v) {
    __hip_atomic_store(reinterpret_cast<uint32_t*>(p), floatBits(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float ldAgent(const float* p) {
    return bitsFloat(__hip_atomic_load(reinterpret_cast<const uint32_t*>(p), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT));
}
// Completes this thread's (agent-scope atomic) box stores before the arrival that
// follows; the "memory" clobber also keeps the compiler from sinking them past it.
// (A release RMW states this order in the language's terms but writes back the
// XCD's L2 on each arrival -- 20 times slower on the bottom-up pass.)
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

// ---- 4-wide collapse ----------------------------------------------------------------
// Even-depth internal nodes become 4-wide nodes holding their children, or, for
// internal children, their grandchildren (src/gpu/lbvh.cu:731-807).  Leaves sit
// at odd depth (direct children) or even depth (grandchildren), so every leaf is
// referenced exactly once.

__global__ void wideFlagKernel(const int* __restrict__ depth, int internalCount,
                               int* __restrict__ flags) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < internalCount) flags[i] = (depth[i] & 1) == 0 ? 1 : 0;
}

__device__ __forceinline__ void setEntry(WideEntry& e, const Node2& box, int child) {
    e.lo[0] = box.lo[0];
    e.lo[1] = box.lo[1];
    e.lo[2] = box.lo[2];
    e.hi[0] = box.hi[0];
    e.hi[1] = box.hi[1];
    e.hi[2] = box.hi[2];
    e.child = child;
    e.pad = 0;
}

__device__ __forceinline__ void clearEntry(WideEntry& e) {
    e.lo[0] = e.lo[1] = e.lo[2] = kInf;
    e.hi[0] = e.hi[1] = e.hi[2] = -kInf;
    e.child = kWideEmpty;
    e.pad = 0;
}

// Four lanes per internal node: lane s writes entry s (32 B) of the node's 4-wide
// node, so a wide node is one contiguous 128 B write.  Even-depth internal node i
// holds its children, with each interior child replaced by its two children.
// A node id >= n - 1 is a leaf, so the entry layout needs no node reads.
__global__ __launch_bounds__(kBlock) void wideEmitKernel(const Node2* __restrict__ nodes, int n,
                                                         const int* __restrict__ depth,
                                                         const int* __restrict__ wideId,
                                                         Wide4* __restrict__ wide) {
    const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    const int i = static_cast<int>(t >> 2);
    const int slot = static_cast<int>(t & 3);
    if (i >= n - 1 || (depth[i] & 1) != 0) return;
    const Node2 nd = nodes[i];
    const int a0 = nd.left >= n - 1 ? 1 : 2;
    const int a1 = nd.right >= n - 1 ? 1 : 2;
    WideEntry e;
    clearEntry(e);
    int src = -1;
    if (slot < a0) {
        src = a0 == 1 ? nd.left : (slot == 0 ? nodes[nd.left].left : nodes[nd.left].right);
    } else if (slot - a0 < a1) {
        const int j = slot - a0;
        src = a1 == 1 ? nd.right : (j == 0 ? nodes[nd.right].left : nodes[nd.right].right);
    }
    if (src >= 0) {
        const Node2 sn = nodes[src];
        setEntry(e, sn, src >= n - 1 ? sn.left : wideId[src]);
    }
    wide[wideId[i]].c[slot] = e;
}

// ---- refit --------------------------------------------------------------------------

// Full refit: gather scene-order items into leaf order, then the shared
// bottom-up pass.  primSlot (scene index -> slot) is the inverse of primIds.
__global__ void invertPermKernel(const int* __restrict__ primIds, int n, int* __restrict__ primSlot) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k < n) primSlot[primIds[k]] = k;
}

template <typename L>
__global__ void gatherKernel(const typename L::Item* __restrict__ src, const int* __restrict__ primIds,
                             int n, typename L::Item* __restrict__ dst) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k < n) dst[k] = src[primIds[k]];
}

// Dirty refit, pass 1: write the new items and leaf boxes, then mark the paths to
// the root.  need[p] ends as the number of p's children that lie on a dirty path
// (1 or 2): the first visitor of p continues upward, the second stops there.
template <typename L>
__global__ void dirtyMarkKernel(const typename L::Item* __restrict__ staged,
                                const int* __restrict__ sceneIds, int count,
                                const int* __restrict__ primSlot, int n,
                                typename L::Item* __restrict__ items, Node2* __restrict__ nodes,
                                const int* __restrict__ parent, int* __restrict__ need,
                                int* __restrict__ leafOut) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    const int slot = primSlot[sceneIds[i]];
    const typename L::Item it = staged[i];
    items[slot] = it;
    const Box b = L::bounds(it);
    Node2* leaf = &nodes[n - 1 + slot];
    leaf->lo[0] = b.lo.x;
    leaf->lo[1] = b.lo.y;
    leaf->lo[2] = b.lo.z;
    leaf->hi[0] = b.hi.x;
    leaf->hi[1] = b.hi.y;
    leaf->hi[2] = b.hi.z;
    leafOut[i] = n - 1 + slot;
    int p = parent[n - 1 + slot];
    while (p >= 0) {
        if (atomicAdd(&need[p], 1) != 0) break;
        p = parent[p];
    }
}

// Dense dirty refit: only the new items are written; every box is then rebuilt by
// the bottom-up pass over the whole tree (same merge order, so the same bytes).
template <typename L>
__global__ void scatterDirtyKernel(const typename L::Item* __restrict__ staged,
                                   const int* __restrict__ sceneIds, int count,
                                   const int* __restrict__ primSlot,
                                   typename L::Item* __restrict__ items) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < count) items[primSlot[sceneIds[i]]] = staged[i];
}

// Dirty refit, pass 2: climb from every dirty leaf; at a node the last of its
// dirty children to arrive merges it and resets the node's counters, so the
// counters are all zero again when the pass ends.
__global__ void dirtyClimbKernel(const int* __restrict__ leaves, int count, Node2* __restrict__ nodes,
                                 const int* __restrict__ parent, int* __restrict__ need,
                                 int* __restrict__ arrive) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    int node = leaves[i];
    while (true) {
        const int p = parent[node];
        if (p < 0) break;
        // need[p] is read before this arrival is counted: the last arrival resets it,
        // so a read after the increment could see the reset value.  The vmcnt wait
        // also orders this thread's box stores ahead of the increment.
        const int want = __hip_atomic_load(&need[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        storesDone();
        const int prior = __hip_atomic_fetch_add(&arrive[p], 1, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
        if (prior + 1 < want) break;
        const int l = nodes[p].left;
        const int r = nodes[p].right;
        const Box lb = readBoxAgent(&nodes[l]);
        const Box rb = readBoxAgent(&nodes[r]);
        writeBoxAgent(&nodes[p], mergeChildren(lb, rb));
        __hip_atomic_store(&arrive[p], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&need[p], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        node = p;
    }
}

__global__ void zeroIntKernel(int* __restrict__ p, int n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---- validation (src/gpu/lbvh.cu:613-729) --------------------------------------------
// errors: 0 bad children or parent box, 1 bad leaf box, 2 bad reference count,
// 3 bad leaf id, 4 bad slot -> scene map or payload, 5 bad depth.

template <typename L>
__global__ void checkNodesKernel(const Node2* __restrict__ nodes,
                                 const typename L::Item* __restrict__ sorted, int n, int nodeCount,
                                 int* __restrict__ visited, int* __restrict__ errors) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nodeCount) return;
    const Node2 nd = nodes[i];
    if (i < n - 1) {
        if (nd.left < 0 || nd.left >= nodeCount || nd.right < 0 || nd.right >= nodeCount) {
            atomicAdd(&errors[0], 1);
            return;
        }
        atomicAdd(&visited[nd.left], 1);
        atomicAdd(&visited[nd.right], 1);
        const Box mine = nodeBox(nd);
        Box uni = emptyBox();
        growBox(uni, nodeBox(nodes[nd.left]));
        growBox(uni, nodeBox(nodes[nd.right]));
        if (mine.lo.x != uni.lo.x || mine.lo.y != uni.lo.y || mine.lo.z != uni.lo.z ||
            mine.hi.x != uni.hi.x || mine.hi.y != uni.hi.y || mine.hi.z != uni.hi.z) {
            atomicAdd(&errors[0], 1);
        }
    } else {
        const int slot = -nd.left - 1;
        if (slot < 0 || slot >= n) {
            atomicAdd(&errors[0], 1);
            return;
        }
        if (i != n - 1 + slot || nd.right != 1) atomicAdd(&errors[3], 1);
        const Box mine = nodeBox(nd);
        const Box prim = L::bounds(sorted[slot]);
        if (mine.lo.x != prim.lo.x || mine.lo.y != prim.lo.y || mine.lo.z != prim.lo.z ||
            mine.hi.x != prim.hi.x || mine.hi.y != prim.hi.y || mine.hi.z != prim.hi.z) {
            atomicAdd(&errors[1], 1);
        }
    }
}

__global__ void checkRefsKernel(const int* __restrict__ visited, int nodeCount,
                                int* __restrict__ errors) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nodeCount) return;
    if (visited[i] != (i == 0 ? 0 : 1)) atomicAdd(&errors[2], 1);
}

__global__ void checkPermKernel(const int* __restrict__ primIds, int n, int* __restrict__ hist,
                                int* __restrict__ errors) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int id = primIds[i];
    if (id < 0 || id >= n) {
        atomicAdd(&errors[4], 1);
        return;
    }
    atomicAdd(&hist[id], 1);
}

__global__ void checkPermCountKernel(const int* __restrict__ hist, int n, int* __restrict__ errors) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n && hist[i] != 1) atomicAdd(&errors[4], 1);
}

template <typename L>
__global__ void checkPayloadKernel(const typename L::Item* __restrict__ sorted,
                                   const typename L::Item* __restrict__ items,
                                   const int* __restrict__ primIds, int n, int* __restrict__ errors) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const int id = primIds[s];
    if (id < 0 || id >= n) return;  // counted by checkPermKernel
    const uint32_t* a = reinterpret_cast<const uint32_t*>(&sorted[s]);
    const uint32_t* b = reinterpret_cast<const uint32_t*>(&items[id]);
    uint32_t diff = 0;
    for (size_t k = 0; k < sizeof(typename L::Item) / 4; ++k) diff |= a[k] ^ b[k];
    if (diff) atomicAdd(&errors[4], 1);
}

__global__ void checkDepthKernel(const Node2* __restrict__ nodes, const int* __restrict__ depth,
                                 int n, int nodeCount, int* __restrict__ errors) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nodeCount) return;
    if (i == 0 && depth[0] != 0) atomicAdd(&errors[5], 1);
    if (i < n - 1) {
        const int l = nodes[i].left;
        const int r = nodes[i].right;
        if (l >= 0 && l < nodeCount && depth[l] != depth[i] + 1) atomicAdd(&errors[5], 1);
        if (r >= 0 && r < nodeCount && depth[r] != depth[i] + 1) atomicAdd(&errors[5], 1);
    }
    if (depth[i] >= 63) atomicAdd(&errors[5], 1);
}

}  // namespace

void reserveWide(BvhState& b) {
    if (b.n < 2) throw ArgError("buildWideNodes: tree is empty");
    const int internal = b.n - 1;
    b.depthArr.reserve(static_cast<size_t>(b.nodeCount));
    b.jump.reserve(static_cast<size_t>(b.nodeCount));
    b.wideId.reserve(static_cast<size_t>(internal));
    b.wideFlags.reserve(static_cast<size_t>(internal));
    b.wideScan.reserve(static_cast<size_t>(scanScratchInts(internal)) + 1);
    // at most one wide node per internal node: no count round trip before the emit
    b.wide.reserve(static_cast<size_t>(internal));
}

void buildWide(BvhState& b, hipStream_t s) {
    reserveWide(b);
    computeDepth(b, s);
    const int internal = b.n - 1;
    wideFlagKernel<<<blocksFor(internal, kBlock), kBlock, 0, s>>>(b.depthArr.p, internal, b.wideFlags.p);
    int* total = b.wideScan.p + scanScratchInts(internal);
    scanExclusive(b.wideFlags.p, b.wideId.p, internal, total, b.wideScan.p, s);
    wideEmitKernel<<<blocksFor(4LL * internal, kBlock), kBlock, 0, s>>>(b.nodes.p, b.n, b.depthArr.p,
                                                                       b.wideId.p, b.wide.p);
    hipCheck(hipGetLastError(), "wide emit");
    int count = 0;
    hipCheck(hipMemcpyAsync(&count, total, sizeof(int), hipMemcpyDeviceToHost, s), "wide count");
    hipCheck(hipStreamSynchronize(s), "wide sync");
    if (count <= 0) throw DeviceError("buildWideNodes: no wide nodes");
    b.wideCount = count;
    b.wideValid = true;
}

// ---- 4-wide collapse of a batch ------------------------------------------------------
// Scene sc's binary nodes sit at 2 primOff[sc] - sc (local ids, root 0).  Depths come
// from pointer jumping over global node indices (a root jumps to itself), the wide
// nodes of all scenes are numbered by one scan over the flags of every node, and each
// scene's wide nodes keep local child indices (global index minus the scene's root
// index), so a scene's walk runs on wide + wideRoot[sc] as on a plain tree.
namespace {

// Scene of a global node index: the last sc with 2 primOff[sc] - sc <= g.
__device__ __forceinline__ int nodeScene(const int* __restrict__ primOff, int scenes, int g) {
    int lo = 0, hi = scenes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (2 * primOff[mid] - mid <= g) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    return lo;
}

__global__ void batchJumpInitKernel(const int* __restrict__ parent, const int* __restrict__ primOff, int scenes,
                                    int nodeCount, unsigned long long* __restrict__ jump) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= nodeCount) return;
    const int sc = nodeScene(primOff, scenes, g);
    const int off = 2 * primOff[sc] - sc;
    const int p = parent[g];
    const int anc = p < 0 ? g : off + p;
    const int dist = p < 0 ? 0 : 1;
    jump[g] = (static_cast<unsigned long long>(static_cast<uint32_t>(dist)) << 32) | static_cast<uint32_t>(anc);
}

// Every step moves to a strict ancestor, so the loop ends at the scene's root.
__global__ void batchJumpKernel(int nodeCount, unsigned long long* __restrict__ jump, int* __restrict__ depth) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= nodeCount) return;
    const unsigned long long j0 = __hip_atomic_load(&jump[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int anc = static_cast<int>(j0 & 0xffffffffu);
    int dist = static_cast<int>(j0 >> 32);
    while (true) {
        const unsigned long long pj = __hip_atomic_load(&jump[anc], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int a2 = static_cast<int>(pj & 0xffffffffu);
        if (a2 == anc) break;
        dist += static_cast<int>(pj >> 32);
        anc = a2;
        const unsigned long long mine = (static_cast<unsigned long long>(static_cast<uint32_t>(dist)) << 32) |
                                        static_cast<uint32_t>(anc);
        __hip_atomic_store(&jump[g], mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    depth[g] = dist;
}

// flags[g] = 1 for an internal node (local id below the scene's n - 1) at even depth.
__global__ void batchWideFlagKernel(const int* __restrict__ depth, const int* __restrict__ primOff, int scenes,
                                    int nodeCount, int* __restrict__ flags) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= nodeCount) return;
    const int sc = nodeScene(primOff, scenes, g);
    const int local = g - (2 * primOff[sc] - sc);
    const int ns = primOff[sc + 1] - primOff[sc];
    flags[g] = (local < ns - 1 && (depth[g] & 1) == 0) ? 1 : 0;
}

__global__ void batchWideRootKernel(const int* __restrict__ primOff, int scenes, const int* __restrict__ wideId,
                                    const int* __restrict__ total, int* __restrict__ wideRoot) {
    const int sc = blockIdx.x * kBlock + threadIdx.x;
    if (sc < scenes) wideRoot[sc] = wideId[2 * primOff[sc] - sc];
    if (sc == scenes) wideRoot[sc] = *total;
}

// Four lanes per node, as wideEmitKernel; interior entries hold scene-local wide ids.
__global__ __launch_bounds__(kBlock) void batchWideEmitKernel(const Node2* __restrict__ nodes,
                                                              const int* __restrict__ primOff, int scenes,
                                                              int nodeCount, const int* __restrict__ flags,
                                                              const int* __restrict__ wideId,
                                                              const int* __restrict__ wideRoot,
                                                              Wide4* __restrict__ wide) {
    const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    const int g = static_cast<int>(t >> 2);
    const int slot = static_cast<int>(t & 3);
    if (g >= nodeCount || flags[g] == 0) return;
    const int sc = nodeScene(primOff, scenes, g);
    const int off = 2 * primOff[sc] - sc;
    const int ns = primOff[sc + 1] - primOff[sc];
    const Node2* tn = nodes + off;
    const Node2 nd = nodes[g];
    const int a0 = nd.left >= ns - 1 ? 1 : 2;
    const int a1 = nd.right >= ns - 1 ? 1 : 2;
    WideEntry e;
    clearEntry(e);
    int src = -1;
    if (slot < a0) {
        src = a0 == 1 ? nd.left : (slot == 0 ? tn[nd.left].left : tn[nd.left].right);
    } else if (slot - a0 < a1) {
        const int j = slot - a0;
        src = a1 == 1 ? nd.right : (j == 0 ? tn[nd.right].left : tn[nd.right].right);
    }
    if (src >= 0) {
        const Node2 sn = tn[src];
        setEntry(e, sn, src >= ns - 1 ? sn.left : wideId[off + src] - wideRoot[sc]);
    }
    wide[wideId[g]].c[slot] = e;
}

}  // namespace

void buildBatchWide(BatchState& b, hipStream_t s) {
    if (b.wideValid) return;
    const int nc = 2 * b.total - b.scenes;
    DevBuf<unsigned long long> jump(static_cast<size_t>(nc));
    DevBuf<int> depth(static_cast<size_t>(nc)), flags(static_cast<size_t>(nc)),
        scan(static_cast<size_t>(scanScratchInts(nc)) + 1);
    b.wideId.alloc(static_cast<size_t>(nc));
    b.wideRoot.alloc(static_cast<size_t>(b.scenes) + 1);
    const int grid = blocksFor(nc, kBlock);
    batchJumpInitKernel<<<grid, kBlock, 0, s>>>(b.parent.p, b.primOff.p, b.scenes, nc, jump.p);
    batchJumpKernel<<<grid, kBlock, 0, s>>>(nc, jump.p, depth.p);
    batchWideFlagKernel<<<grid, kBlock, 0, s>>>(depth.p, b.primOff.p, b.scenes, nc, flags.p);
    int* total = scan.p + scanScratchInts(nc);
    scanExclusive(flags.p, b.wideId.p, nc, total, scan.p, s);
    batchWideRootKernel<<<blocksFor(b.scenes + 1, kBlock), kBlock, 0, s>>>(b.primOff.p, b.scenes, b.wideId.p,
                                                                          total, b.wideRoot.p);
    int count = 0;
    hipCheck(hipMemcpyAsync(&count, total, sizeof(int), hipMemcpyDeviceToHost, s), "batch wide count");
    hipCheck(hipStreamSynchronize(s), "batch wide count sync");
    if (count < b.scenes) throw DeviceError("batch wide: fewer wide nodes than scenes");
    b.wide.alloc(static_cast<size_t>(count));
    batchWideEmitKernel<<<blocksFor(4LL * nc, kBlock), kBlock, 0, s>>>(b.nodes.p, b.primOff.p, b.scenes, nc,
                                                                       flags.p, b.wideId.p, b.wideRoot.p, b.wide.p);
    hipCheck(hipGetLastError(), "batch wide");
    hipCheck(hipStreamSynchronize(s), "batch wide sync");
    b.wideValid = true;
}

// Re-emits every wide node from the current binary boxes (after a refit).
void reemitWide(BvhState& b, hipStream_t s) {
    if (!b.wideValid) return;
    const int internal = b.n - 1;
    wideEmitKernel<<<blocksFor(4LL * internal, kBlock), kBlock, 0, s>>>(b.nodes.p, b.n, b.depthArr.p,
                                                                       b.wideId.p, b.wide.p);
    hipCheck(hipGetLastError(), "wide re-emit");
}

namespace {
// New-scene item j of a topology edit: appended and inserted items come from the
// upload (inserted first, then appended); a survivor is the old item, found
// through the old tree's scene -> slot map.  The r-th survivor (0-based) is old
// index r + d with d = #{i : deleted[i] - i <= r} (deleted is ascending and unique,
// so deleted[i] - i does not decrease).
template <typename L>
__global__ void composeKernel(const typename L::Item* __restrict__ oldItems, const int* __restrict__ oldSlot,
                              const int* __restrict__ deleted, int nDel, const int* __restrict__ inserted,
                              int nIns, int tailBeg, int nNew, const typename L::Item* __restrict__ staged,
                              typename L::Item* __restrict__ out) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= nNew) return;
    if (j >= tailBeg) {
        out[j] = staged[nIns + (j - tailBeg)];
        return;
    }
    int lo = 0, hi = nIns;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (inserted[mid] < j) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    if (lo < nIns && inserted[lo] == j) {
        out[j] = staged[lo];
        return;
    }
    const int r = j - lo;
    int a = 0, c = nDel;
    while (a < c) {
        const int mid = (a + c) >> 1;
        if (deleted[mid] - mid <= r) {
            a = mid + 1;
        } else {
            c = mid;
        }
    }
    out[j] = oldItems[oldSlot[r + a]];
}

template <typename L>
__global__ void scatterMovedKernel(const int* __restrict__ moved, int nMoved,
                                   const typename L::Item* __restrict__ staged, typename L::Item* __restrict__ out) {
    const int m = blockIdx.x * kBlock + threadIdx.x;
    if (m < nMoved) out[moved[m]] = staged[m];
}

template <typename L>
void composeT(BvhState& b, const int* dDel, int nDel, const int* dIns, int nIns, int tailBeg, int nNew,
              const void* dStaged, const int* dMoved, int nMoved, const void* dStagedMoved, void* dOut,
              hipStream_t s) {
    using Item = typename L::Item;
    composeKernel<L><<<blocksFor(nNew, kBlock), kBlock, 0, s>>>(
        reinterpret_cast<const Item*>(b.items.p), b.primSlot.p, dDel, nDel, dIns, nIns, tailBeg, nNew,
        static_cast<const Item*>(dStaged), static_cast<Item*>(dOut));
    if (nMoved > 0)
        scatterMovedKernel<L><<<blocksFor(nMoved, kBlock), kBlock, 0, s>>>(
            dMoved, nMoved, static_cast<const Item*>(dStagedMoved), static_cast<Item*>(dOut));
    hipCheck(hipGetLastError(), "compose edit");
}
}  // namespace

void composeEdit(BvhState& b, const int* dDel, int nDel, const int* dIns, int nIns, int tailBeg, int nNew,
                 const void* dStaged, const int* dMoved, int nMoved, const void* dStagedMoved, void* dOut,
                 hipStream_t s) {
    ensurePrimSlot(b, s);
    switch (b.kind) {
        case kPrimTri:
            composeT<TriLeaf>(b, dDel, nDel, dIns, nIns, tailBeg, nNew, dStaged, dMoved, nMoved, dStagedMoved, dOut, s);
            break;
        case kPrimPoint:
            composeT<PointLeaf>(b, dDel, nDel, dIns, nIns, tailBeg, nNew, dStaged, dMoved, nMoved, dStagedMoved, dOut,
                                s);
            break;
        default:
            composeT<GaussLeaf>(b, dDel, nDel, dIns, nIns, tailBeg, nNew, dStaged, dMoved, nMoved, dStagedMoved, dOut,
                                s);
            break;
    }
}

void ensurePrimSlot(BvhState& b, hipStream_t s) {
    if (b.primSlotValid) return;
    b.primSlot.reserve(static_cast<size_t>(b.n));
    invertPermKernel<<<blocksFor(b.n, kBlock), kBlock, 0, s>>>(b.primIds.p, b.n, b.primSlot.p);
    hipCheck(hipGetLastError(), "invert perm");
    b.primSlotValid = true;
}

template <typename L>
static void refitFullT(BvhState& b, const void* dSceneItems, hipStream_t s) {
    gatherKernel<L><<<blocksFor(b.n, kBlock), kBlock, 0, s>>>(
        static_cast<const typename L::Item*>(dSceneItems), b.primIds.p, b.n,
        reinterpret_cast<typename L::Item*>(b.items.p));
    hipCheck(hipGetLastError(), "refit gather");
    propagateAll(b, s);
}

void refitFull(BvhState& b, const void* dSceneItems, hipStream_t s) {
    switch (b.kind) {
        case kPrimTri: refitFullT<TriLeaf>(b, dSceneItems, s); break;
        case kPrimPoint: refitFullT<PointLeaf>(b, dSceneItems, s); break;
        default: refitFullT<GaussLeaf>(b, dSceneItems, s); break;
    }
}

// The sparse path (mark the dirty paths, then merge only the marked nodes) is a
// chain of dependent device-scope atomics about as long as the tree is deep, whatever
// the dirty count; the bottom-up pass over all leaves streams the tree with its
// local part in LDS.  Past 1 dirty leaf in 256 the full pass is the faster one.
static bool denseRefit(int count, int n) { return static_cast<long long>(count) * 256 > n; }

template <typename L>
static void refitDirtyT(BvhState& b, const void* dStaged, const int* dSceneIds, int count,
                        int* dLeaves, hipStream_t s, hipEvent_t mid) {
    if (denseRefit(count, b.n)) {
        scatterDirtyKernel<L><<<blocksFor(count, kBlock), kBlock, 0, s>>>(
            static_cast<const typename L::Item*>(dStaged), dSceneIds, count, b.primSlot.p,
            reinterpret_cast<typename L::Item*>(b.items.p));
        hipCheck(hipGetLastError(), "dirty scatter");
        if (mid) hipCheck(hipEventRecord(mid, s), "event");
        propagateAll(b, s);
        return;
    }
    dirtyMarkKernel<L><<<blocksFor(count, kBlock), kBlock, 0, s>>>(
        static_cast<const typename L::Item*>(dStaged), dSceneIds, count, b.primSlot.p, b.n,
        reinterpret_cast<typename L::Item*>(b.items.p), b.nodes.p, b.parent.p, b.need.p, dLeaves);
    hipCheck(hipGetLastError(), "dirty mark");
    if (mid) hipCheck(hipEventRecord(mid, s), "event");
    dirtyClimbKernel<<<blocksFor(count, kBlock), kBlock, 0, s>>>(dLeaves, count, b.nodes.p,
                                                                b.parent.p, b.need.p, b.arrive.p);
    hipCheck(hipGetLastError(), "dirty climb");
}

void refitDirty(BvhState& b, const void* dStaged, const int* dSceneIds, int count, int* dLeaves,
                hipStream_t s, hipEvent_t mid) {
    if (count <= 0) return;
    ensurePrimSlot(b, s);
    if (!denseRefit(count, b.n) && (!b.need.p || b.need.n < static_cast<size_t>(b.n - 1))) {
        b.need.alloc(static_cast<size_t>(b.n - 1));
        zeroIntKernel<<<blocksFor(b.n - 1, kBlock), kBlock, 0, s>>>(b.need.p, b.n - 1);
    }
    switch (b.kind) {
        case kPrimTri: refitDirtyT<TriLeaf>(b, dStaged, dSceneIds, count, dLeaves, s, mid); break;
        case kPrimPoint: refitDirtyT<PointLeaf>(b, dStaged, dSceneIds, count, dLeaves, s, mid); break;
        default: refitDirtyT<GaussLeaf>(b, dStaged, dSceneIds, count, dLeaves, s, mid); break;
    }
}

template <typename L>
static void validateT(const BvhState& b, const void* dSceneItems, int* errors, hipStream_t s) {
    const int n = b.n;
    const int nc = b.nodeCount;
    DevBuf<int> visited(static_cast<size_t>(nc));
    DevBuf<int> hist(static_cast<size_t>(n));
    zeroIntKernel<<<blocksFor(nc, kBlock), kBlock, 0, s>>>(visited.p, nc);
    zeroIntKernel<<<blocksFor(n, kBlock), kBlock, 0, s>>>(hist.p, n);
    const auto* sorted = reinterpret_cast<const typename L::Item*>(b.items.p);
    checkNodesKernel<L><<<blocksFor(nc, kBlock), kBlock, 0, s>>>(b.nodes.p, sorted, n, nc, visited.p, errors);
    checkRefsKernel<<<blocksFor(nc, kBlock), kBlock, 0, s>>>(visited.p, nc, errors);
    checkPermKernel<<<blocksFor(n, kBlock), kBlock, 0, s>>>(b.primIds.p, n, hist.p, errors);
    checkPermCountKernel<<<blocksFor(n, kBlock), kBlock, 0, s>>>(hist.p, n, errors);
    checkPayloadKernel<L><<<blocksFor(n, kBlock), kBlock, 0, s>>>(
        sorted, static_cast<const typename L::Item*>(dSceneItems), b.primIds.p, n, errors);
    checkDepthKernel<<<blocksFor(nc, kBlock), kBlock, 0, s>>>(b.nodes.p, b.depthArr.p, n, nc, errors);
    hipCheck(hipGetLastError(), "validate");
    hipCheck(hipStreamSynchronize(s), "validate sync");
}

void validateTree(BvhState& b, const void* dSceneItems, int* hostErrors, hipStream_t s) {
    computeDepth(b, s);
    DevBuf<int> errors(8);
    zeroIntKernel<<<1, kBlock, 0, s>>>(errors.p, 8);
    switch (b.kind) {
        case kPrimTri: validateT<TriLeaf>(b, dSceneItems, errors.p, s); break;
        case kPrimPoint: validateT<PointLeaf>(b, dSceneItems, errors.p, s); break;
        default: validateT<GaussLeaf>(b, dSceneItems, errors.p, s); break;
    }
    hipCheck(hipMemcpy(hostErrors, errors.p, 6 * sizeof(int), hipMemcpyDeviceToHost), "errors D2H");
}

}  // namespace qr
