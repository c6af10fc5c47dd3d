// Launchers for the traversal kernels (rays, closest, kNN, radius) and the
// Morton reordering of queries.
#pragma once

#include "qr_internal.h"

namespace qr {

struct TraverseArgs {
    int kind = 0;                       // 0 tri, 1 point, 2 Gaussian
    bool wide = false;
    const Node2* nodes = nullptr;
    int nodeCount = 0;
    const Wide4* wideNodes = nullptr;
    int wideCount = 0;
    const void* items = nullptr;        // leaf-ordered primitives
    const int* primIds = nullptr;       // leaf slot -> scene index
    const uint32_t* order = nullptr;    // optional visiting order (lane i -> query order[i])
    unsigned long long* stats = nullptr;  // optional [node tests, leaf tests, overflows]
    // false: the walk shape comes from an earlier tuning of the same tree and size
    // class or from the measured crossovers, never from timing trials on this call
    bool tune = true;
};

// Traversal counters: kStatCopies copies of [node tests, leaf tests, stack
// overflows, nodes opened], block b adding into copy b % kStatCopies (same-address
// atomics serialize; one copy for every wave of a 100k-wave launch took
// milliseconds), summed when read.
constexpr int kStatCopies = 64;
constexpr int kStatWords = 4 * kStatCopies;
void zeroStats(unsigned long long* stats, hipStream_t s);
void zeroInts(int* p, int n, hipStream_t s);

// Quantisation domain of a Morton reordering.
enum MortonBox : int {
    kBoxMeasure = 0,  // measure the bounds of the elements on the device (into bb)
    kBoxUnit = 1,     // [-1, 1]^3 (ray directions)
    kBoxGiven = 2,    // bb already holds the bounds (ordered bits: lo xyz, hi xyz)
};

// Computes a visiting order by 30-bit Morton code of the first three floats of
// each element (elements are `stride` floats apart) quantised over the domain
// given by boxMode.  bb: 6 device words.
void mortonOrder(const float* base, int stride, int count, int boxMode, SortWorkspace& ws,
                 uint32_t* keys, uint32_t* keysSorted, uint32_t* orderOut, uint32_t* bb,
                 hipStream_t s);

void launchRays(const TraverseArgs& a, const Ray* rays, int count, uint64_t* out, hipStream_t s);
// Same, also writing the hit triangle's scene index (-1 for a miss) to hitPrim.
void launchRaysSave(const TraverseArgs& a, const Ray* rays, int count, uint64_t* out, int* hitPrim,
                    hipStream_t s);
void launchClosest(const TraverseArgs& a, const V3* pts, int count, ClosestResult* out,
                   hipStream_t s);
// Result j of query q lands at out[q * outStride + j] (outStride >= k).
void launchKnn(const TraverseArgs& a, const V3* pts, int count, int k, int outStride,
               ClosestResult* out, hipStream_t s);
// maxCount (may be null): kMaxCopies ints, each raised atomically to the largest
// per-query count of the blocks b with b % kMaxCopies equal to its index (one word for
// every wave of a 100k-wave launch serialized the atomics); the largest count is
// the largest word.
constexpr int kMaxCopies = 64;
// Radius, one walk: counts hits per query and keeps the first `cap` of each query in
// its slot of `capture` (count * cap keys); returns false when the tree or mode takes
// the per-lane kernels, which only count (nothing is captured).
int radiusCaptureCap(int count);  // hits kept per query: 64, fewer for huge query sets
bool launchRadiusCapture(const TraverseArgs& a, const V3* pts, int count, float r2, int* counts,
                         uint64_t* capture, int cap, int* maxCount, hipStream_t s);
// True when the capture runs one wave per query, which leaves every captured segment
// of at most 64 hits sorted by (dist2, index).
bool radiusCapturePresorts(const TraverseArgs& a, int count);
// Writes the CSR segments of the queries the capture does not hold (all queries when
// nothing was captured: cap 0); counts are those of the capture / count pass.
void launchRadiusFill(const TraverseArgs& a, const V3* pts, int count, float r2, const int* counts,
                      int cap, const int* offsets, uint64_t* keys, hipStream_t s);
// Copies the captured segments into the CSR keys.
void radiusCompact(const int* counts, const int* offsets, const uint64_t* capture, int count, int cap,
                   uint64_t* keys, long long capacity, hipStream_t s);
// Writes the total (*total) and the largest of the kMaxCopies per-query maxima to
// hostDev[0] and hostDev[1] (device view of pinned host words).
void radiusSummary(const int* total, const int* maxCopies, int* hostDev, hipStream_t s);
// scratchInts must hold 4 + 2 * count ints.
void segmentSort(const int* offsets, uint64_t* keys, int count, int maxSeg, int* scratchInts,
                 hipStream_t s);

// ---- batched mode: persistent work-queue traversal over many trees ------------
struct BatchQueryArgs {
    int kind = 0;
    const Node2* nodes = nullptr;   // all trees (scene s from 2 primOff[s] - s)
    const Wide4* wide = nullptr;    // 4-wide nodes of all trees (scene s from wideRoot[s]), wide variants
    const int* wideRoot = nullptr;  // scenes + 1
    const void* items = nullptr;    // leaf-order items of all trees
    const int* primIds = nullptr;   // per slot: scene-local index
    const int* primOff = nullptr;   // scenes + 1
    const V3* points = nullptr;     // queries, scene-major
    const int* qScene = nullptr;    // scene of each query
    const uint32_t* order = nullptr;  // optional processing order
    int count = 0;
    int k = 1;                      // 1: closest; up to 32
    int outStride = 1;
    ClosestResult* out = nullptr;   // per query: outStride records, scene-local indices
    int* counter = nullptr;         // work-queue head (reset before each launch)
};
// Variants the autotuner chooses from: binary or 4-wide nodes; the stack in LDS
// (256- or 64-lane workgroups, spilling to scratch), in scratch, or its top entries in
// registers.
enum BatchVariant : int {
    kBatchLds256 = 0, kBatchLds64 = 1, kBatchScratch64 = 2, kBatchReg64 = 3,
    kBatchWideLds256 = 4, kBatchWideLds64 = 5, kBatchWideScratch64 = 6, kBatchWideReg64 = 7,
    kBatchVariants = 8
};
inline bool batchVariantWide(int v) { return v >= kBatchWideLds256; }
void launchBatchNearest(const BatchQueryArgs& a, int variant, hipStream_t s);
void launchBatchReset(int* counter, hipStream_t s);

// Overflow reporting: armOverflowWord() points the traversal kernels at a pinned host
// word of the current device (once per device; it allocates that word, so
// allocation-free entry points do not call it); checkOverflow(), after a call's
// kernels have completed, throws DeviceError if a walk overflowed its stack since
// the last check, and clears the word.
void armOverflowWord();
void checkOverflow();

// Traversal shape of the wide point queries: 0 automatic, 1 one lane per query,
// 2 four lanes per query (process-wide).
void setTraversalMode(int mode);

}  // namespace qr
