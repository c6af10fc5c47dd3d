// The state behind an opaque qr_bvh handle and the build entry points.
#pragma once

#include <vector>

#include "qr_internal.h"

namespace qr {

enum PrimKind : int { kPrimTri = 0, kPrimPoint = 1, kPrimGauss = 2 };
enum Builder : int { kBuildLbvh = 0, kBuildSah = 1 };

inline size_t itemSize(int kind) {
    return kind == kPrimTri ? sizeof(Tri) : (kind == kPrimPoint ? sizeof(V3) : sizeof(Gaussian));
}

struct BuildTimingsC {
    double mortonMs, sortMs, treeMs, scatterMs, refitMs, sahBinMs, sahSplitMs, sahPartMs, totalMs;
};

// Work of the calling thread's last binned-SAH build: levels of the split loop,
// primitive slots binned (levels that score SAH planes) and partitioned (every
// level), and slots left to the serial tails (tasks of at most 64).
struct SahBuildStats {
    int levels = 0;
    long long binnedSlots = 0, partitionedSlots = 0, serialSlots = 0;
};
extern thread_local SahBuildStats tSahStats;
// SAH builds record per-level phase events (bin / split / partition timings) when set.
extern int gBuildPhaseTiming;

struct BvhState {
    int kind = kPrimTri;
    int builder = kBuildLbvh;
    int n = 0;
    int nodeCount = 0;
    int depth = 0;

    DevBuf<Node2> nodes;            // 2n-1
    DevBuf<unsigned char> items;    // leaf order, n * itemSize(kind)
    DevBuf<int> primIds;            // slot -> scene index
    DevBuf<int> parent;             // per node, root -1
    DevBuf<int> arrive;             // per internal node: bottom-up arrival counters
    DevBuf<int> height;             // per internal node: subtree height (leaves are 0)
    DevBuf<int> farEnd;             // per internal node i: the other end j of its leaf range [i, j] or [j, i]

    // scene index -> leaf slot; rebuilt lazily after a build or topology update
    DevBuf<int> primSlot;
    bool primSlotValid = false;

    // per node depth (root 0); computed on demand for wide collapse and validate
    DevBuf<int> depthArr;
    bool depthValid = false;
    DevBuf<unsigned long long> jump;  // pointer-jumping scratch of the depth pass

    // 4-wide collapse
    DevBuf<Wide4> wide;
    DevBuf<int> wideId;             // per internal node: wide index or -1
    DevBuf<int> wideFlags, wideScan;  // collapse scratch (flags, scan partials + total)
    int wideCount = 0;
    bool wideValid = false;

    // refit scratch (grows, reused across calls)
    DevBuf<int> need;               // per internal node: dirty children (dirty refit)
    DevBuf<unsigned char> stage;    // uploaded items (full or dirty)
    DevBuf<int> dirtyIds;
    DevBuf<int> dirtyLeaves;
    DevBuf<int> counters;           // small device counters
    DevBuf<int> emitStage;          // LBVH build: frontier entries handed up between levels
    DevBuf<int> emitCounts;         //   frontier sizes per unit (-1: the unit went device-wide)

    // scene-order primitives kept for validate / refit-by-device-pointer paths
    SortWorkspace sortWs;

    // build work buffers, kept with the tree so that a rebuild or topology update of
    // the same tree allocates nothing once they are large enough: Morton codes, sorted
    // codes and leaf order (LBVH), and one arena for the SAH builder's records and
    // task tables
    DevBuf<uint32_t> keys, keysSorted, order;
    DevBuf<unsigned char> sahArena;
    DevBuf<unsigned char> editBuf;  // topology edit: staged items and index lists
};

// Many independent trees built at once (the batched mode).  Scene s owns items
// [primOff[s], primOff[s+1]), nodes from 2 primOff[s] - s, internal slots from
// primOff[s] - s; node ids, leaf slots and prim ids are local to the scene.
struct BatchState {
    int kind = kPrimTri;
    int scenes = 0;
    int total = 0;
    int maxDepth = 0;
    std::vector<int> primOffH;
    DevBuf<int> primOff;             // scenes + 1
    DevBuf<Node2> nodes;             // 2 total - scenes
    DevBuf<unsigned char> items;     // leaf order per scene
    DevBuf<int> primIds;             // per slot: scene-local index
    DevBuf<int> parent, arrive, height, depth, sceneId;
    SortWorkspace ws;
    // 4-wide nodes of every scene (buildBatchWide): scene sc's wide nodes start at
    // wideRoot[sc] (its root) and hold scene-local wide ids; wideRoot[scenes] = total
    bool wideValid = false;
    DevBuf<Wide4> wide;
    DevBuf<int> wideId, wideRoot;
};

// 4-wide collapse of every scene of a batch (once; synchronous).
void buildBatchWide(BatchState& b, hipStream_t s);
// Segmented LBVH of all scenes; b.primOff / primOffH / scenes / total must be set.
void buildBatchDevice(BatchState& b, int kind, const void* dItems, hipStream_t s, BuildTimingsC* t);
// sceneId[i] = scene of item i, for scene-major items with offsets primOff.
void batchSceneIds(const int* primOff, int scenes, int* sceneId, hipStream_t s);
// Stable order of items (any primitive kind; query points use kPrimPoint) by
// (scene, Morton code of the centroid within its scene's centroid bounds, index).
// cb: 6 x scenes words; codes, tmpA, tmpB: n words of scratch; outputs
// sceneSorted[i] = scene of the i-th item in that order and order[i] = its index.
void batchMortonOrder(int kind, const void* items, int n, const int* sceneId, int scenes, SortWorkspace& ws,
                      uint32_t* cb, uint32_t* codes, uint32_t* tmpA, uint32_t* tmpB, uint32_t* sceneSorted,
                      uint32_t* order, hipStream_t s);

// Builds into an empty state from scene-order items already on the device.
// Frontier nodes a unit of the LBVH build may hand up (0..16, see qr_set_frontier_cap).
void setFrontierCap(int cap);
void buildLbvhDevice(BvhState& b, int kind, const void* dItems, int n, hipStream_t s,
                     BuildTimingsC* t);
void buildSahDevice(BvhState& b, int kind, const void* dItems, int n, hipStream_t s,
                    BuildTimingsC* t);

// Per-node depth by pointer jumping over the parent array (root 0).
void computeDepth(BvhState& b, hipStream_t s);

// Leaf init + bottom-up box propagation for all leaves, given items already in
// leaf order (refit path shares it).
void propagateAll(BvhState& b, hipStream_t s);

// Leaf gather + leaf nodes + bottom-up boxes and heights for a topology already in
// b.nodes (internal left/right), b.parent (root -1) and zeroed b.arrive; order[k]
// is the scene index of leaf slot k.  The root's height (b.height[0]) is the depth.
// gate: a device word that, when non-zero at launch, leaves the topology unclimbed
// (a builder's error flag), or nullptr.
void finishTopology(BvhState& b, const void* dSrc, const uint32_t* order, hipStream_t s,
                    const int* gate = nullptr);

// Scene-order items of the new scene of a topology edit, on the device, from the
// old tree's items (survivors) and the uploaded items: dStaged = inserted items
// then appended ones, dStagedMoved[m] = the new item at new index dMoved[m].
void composeEdit(BvhState& b, const int* dDel, int nDel, const int* dIns, int nIns, int tailBeg, int nNew,
                 const void* dStaged, const int* dMoved, int nMoved, const void* dStagedMoved, void* dOut,
                 hipStream_t s);

// Allocates the buffers of the 4-wide collapse (depth table, flags, scan, wide nodes)
// so that buildWide itself allocates nothing more.
void reserveWide(BvhState& b);
void buildWide(BvhState& b, hipStream_t s);
void reemitWide(BvhState& b, hipStream_t s);
void ensurePrimSlot(BvhState& b, hipStream_t s);
// Full refit from scene-order items on the device.
void refitFull(BvhState& b, const void* dSceneItems, hipStream_t s);
// Dirty refit: dStaged[i] is the new item of scene index dSceneIds[i] (unique ids).
// dLeaves is scratch of count ints.  mid (optional) is recorded between the mark
// and merge passes.
void refitDirty(BvhState& b, const void* dStaged, const int* dSceneIds, int count, int* dLeaves,
                hipStream_t s, hipEvent_t mid);
// errors[6] as the reference's validate (bad children/box, leaf box, refcount,
// leaf id, prim map, depth).
void validateTree(BvhState& b, const void* dSceneItems, int* hostErrors, hipStream_t s);

}  // namespace qr
