// Binned SAH builder for gfx950: top-down, level-synchronous, 32 bins per axis.
//
// Reference: src/gpu/lbvh.cu:827-1291 (binned SAH kernels), :1777-1952 (level
// driver), :2043-2122 (entry points).  The reference bins 16 ways with one global
// bin table per task; this build bins 32 ways and keeps the histograms in LDS:
//   * tasks above kChunk primitives are cut into kChunk-position chunks; each
//     workgroup builds its chunk's histogram (3 axes x 32 bins x {count, box,
//     centroid box}) with LDS atomics and flushes it once with global atomics;
//     one wave per task then scores the 3 x 31 planes, and a stable two-pass
//     partition (chunk counts -> per-task chunk offsets -> block scan) moves the
//     primitive records;
//   * tasks of kSerial+1 .. kChunk primitives are handled by one workgroup each:
//     histogram, plane choice and stable partition never leave the workgroup;
//   * ranges of kSerial or fewer primitives get balanced count/2 splits in their
//     current order, one thread per range (the reference's serial tail,
//     src/gpu/lbvh.cu:1239-1291);
//   * from level kMedianLevel on every split is the count/2 split, which bounds
//     the depth by kMedianLevel + log2(n) (< 64).
// Plane cost: area(L) * nL + area(R) * nR with area = dx*dy + dy*dz + dz*dx
// (src/gpu/lbvh.cu:1045-1125); a plane must leave at least max(1, count/8)
// primitives on each side; without such a plane the split is count/2.
//
// Node numbering is the LBVH's: a range [s, e] split after position g gives the
// left child id g (or leaf n-1+g when s == g) and the right child id g+1 (or leaf
// n-1+e when g+1 == e), the root is node 0.  This numbering is a bijection onto
// 0..n-2 for any sequence of splits, so leaf boxes, bottom-up merging, heights,
// the 4-wide collapse, refit and every traversal kernel are shared with the LBVH
// path unchanged.  The tree is deterministic: bin boxes are min/max (order-free)
// and the partition is stable.
#include <algorithm>
#include <vector>

#include "qr_bvh.h"

namespace qr {

namespace {

constexpr int kBins = 32;
constexpr int kChunk = 4096;
constexpr int kSerial = 64;
constexpr int kMedianLevel = 32;
constexpr int kBlock = 256;
constexpr int kPerThread = kChunk / kBlock;
constexpr int kBinVals = 13;  // count, box lo xyz, box hi xyz, centroid lo xyz, centroid hi xyz
constexpr int kHist = 3 * kBins * kBinVals;

struct PrimRec {
    float c[3];
    float lo[3];
    float hi[3];
    int id;
};
static_assert(sizeof(PrimRec) == 40, "PrimRec layout");

struct SahTask {
    int start, count, pad0, pad1;
    float cLo[4], cHi[4];
};

struct SahResult {
    int axis, plane, nL, pad;
    float lLo[4], lHi[4], rLo[4], rHi[4];
};

struct Chunk {
    int task, start, count, pad;
};

struct NodeTriple {
    int id, left, right, far;  // far: the other end of the node's leaf range
};

__device__ __forceinline__ int hidx(int a, int b, int v) { return (a * kBins + b) * kBinVals + v; }

struct AxisMap {
    float lo[3], scale[3];
    bool use[3];
};

__device__ __forceinline__ AxisMap axisMap(const SahTask& t) {
    AxisMap m;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float ext = t.cHi[a] - t.cLo[a];
        const float sc = static_cast<float>(kBins) / ext;
        m.lo[a] = t.cLo[a];
        m.scale[a] = sc;
        m.use[a] = ext > 0.0f && sc < 3.0e38f;
    }
    return m;
}

// Bin of a centroid coordinate: floor((c - lo) * 32 / extent), clamped to [0, 31].
__device__ __forceinline__ int binOf(float c, float lo, float scale) {
    float f = (c - lo) * scale;
    f = fminf(fmaxf(f, 0.0f), static_cast<float>(kBins - 1));
    return static_cast<int>(f);
}

__device__ __forceinline__ void histInit(uint32_t* h) {
    for (int i = threadIdx.x; i < kHist; i += blockDim.x) {
        const int v = i % kBinVals;
        h[i] = v == 0 ? 0u : ((v >= 1 && v <= 3) || (v >= 7 && v <= 9)) ? orderedBits(kInf) : orderedBits(-kInf);
    }
}

// Bins of a contiguous range of records.  Each thread takes a run of consecutive
// records and keeps, per axis, the values of its current bin in registers, flushing
// them to the LDS histogram with 13 atomics only when the bin changes: neighbouring
// records (a mesh in scene order, the halves of a partition) mostly share a bin,
// and one-record-per-lane atomics on one shared bin serialize across the wave.
// Count, min and max are exact and order-free, so the histogram is the same bits as
// one histAdd per record.
struct BinRun {
    int bin;
    uint32_t cnt;
    uint32_t v[12];  // box lo xyz, box hi xyz, centroid lo xyz, centroid hi xyz
};

__device__ __forceinline__ void binRunFlush(uint32_t* h, int a, const BinRun& r) {
    if (r.cnt == 0u) return;
    atomicAdd(&h[hidx(a, r.bin, 0)], r.cnt);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        atomicMin(&h[hidx(a, r.bin, 1 + k)], r.v[k]);
        atomicMax(&h[hidx(a, r.bin, 4 + k)], r.v[3 + k]);
        atomicMin(&h[hidx(a, r.bin, 7 + k)], r.v[6 + k]);
        atomicMax(&h[hidx(a, r.bin, 10 + k)], r.v[9 + k]);
    }
}

__device__ __forceinline__ void histAddRange(uint32_t* h, const PrimRec* __restrict__ rec, int count,
                                             const AxisMap& m) {
    const int per = (count + kBlock - 1) / kBlock;
    const int beg = threadIdx.x * per;
    const int end = beg + per < count ? beg + per : count;
    BinRun run[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        run[a].bin = -1;
        run[a].cnt = 0u;
    }
    for (int i = beg; i < end; ++i) {
        const PrimRec r = rec[i];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (!m.use[a]) continue;
            const int b = binOf(r.c[a], m.lo[a], m.scale[a]);
            BinRun& u = run[a];
            if (b != u.bin) {
                binRunFlush(h, a, u);
                u.bin = b;
                u.cnt = 0u;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    u.v[k] = u.v[6 + k] = 0xFFFFFFFFu;
                    u.v[3 + k] = u.v[9 + k] = 0u;
                }
            }
            ++u.cnt;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const uint32_t lo = orderedBits(r.lo[k]), hi = orderedBits(r.hi[k]), c = orderedBits(r.c[k]);
                u.v[k] = lo < u.v[k] ? lo : u.v[k];
                u.v[3 + k] = hi > u.v[3 + k] ? hi : u.v[3 + k];
                u.v[6 + k] = c < u.v[6 + k] ? c : u.v[6 + k];
                u.v[9 + k] = c > u.v[9 + k] ? c : u.v[9 + k];
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (m.use[a]) binRunFlush(h, a, run[a]);
}

__device__ __forceinline__ float areaOf(float lx, float ly, float lz, float hx, float hy, float hz) {
    const float dx = hx - lx, dy = hy - ly, dz = hz - lz;
    return dx * dy + dy * dz + dz * dx;
}

// Plane choice by one wave (lanes 1..31 each score plane `lane` on all usable
// axes); lane 0 writes the result.  h is the task's histogram (LDS or global).
__device__ void chooseSplit(const uint32_t* h, const SahTask& t, const AxisMap& m, SahResult* out) {
    const int lane = threadIdx.x & 63;
    const int count = t.count;
    const int minSide = count / 8 > 1 ? count / 8 : 1;
    float best = kInf;
    int bestKey = 0x7fffffff;  // axis * 32 + plane
    if (lane >= 1 && lane < kBins) {
        for (int a = 0; a < 3; ++a) {
            if (!m.use[a]) continue;
            int nL = 0, nR = 0;
            float l[6] = {kInf, kInf, kInf, -kInf, -kInf, -kInf};
            float r[6] = {kInf, kInf, kInf, -kInf, -kInf, -kInf};
            for (int b = 0; b < kBins; ++b) {
                const int c = static_cast<int>(h[hidx(a, b, 0)]);
                if (c == 0) continue;
                float* dst = b < lane ? l : r;
                if (b < lane) {
                    nL += c;
                } else {
                    nR += c;
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    dst[k] = fminf(dst[k], fromOrderedBits(h[hidx(a, b, 1 + k)]));
                    dst[3 + k] = fmaxf(dst[3 + k], fromOrderedBits(h[hidx(a, b, 4 + k)]));
                }
            }
            if (nL < minSide || nR < minSide) continue;
            const float cost = areaOf(l[0], l[1], l[2], l[3], l[4], l[5]) * static_cast<float>(nL) +
                               areaOf(r[0], r[1], r[2], r[3], r[4], r[5]) * static_cast<float>(nR);
            const int key = a * kBins + lane;
            if (cost < best || (cost == best && key < bestKey)) {
                best = cost;
                bestKey = key;
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float ob = __shfl_xor(best, off, 64);
        const int ok = __shfl_xor(bestKey, off, 64);
        if (ob < best || (ob == best && ok < bestKey)) {
            best = ob;
            bestKey = ok;
        }
    }
    if (lane != 0) return;
    SahResult res;
    if (bestKey == 0x7fffffff) {
        res.axis = -1;
        res.plane = 0;
        res.nL = count / 2;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            res.lLo[k] = res.rLo[k] = t.cLo[k];
            res.lHi[k] = res.rHi[k] = t.cHi[k];
        }
    } else {
        const int a = bestKey / kBins, plane = bestKey % kBins;
        res.axis = a;
        res.plane = plane;
        int nL = 0;
        float lc[6] = {kInf, kInf, kInf, -kInf, -kInf, -kInf};
        float rc[6] = {kInf, kInf, kInf, -kInf, -kInf, -kInf};
        for (int b = 0; b < kBins; ++b) {
            const int c = static_cast<int>(h[hidx(a, b, 0)]);
            if (c == 0) continue;
            float* dst = b < plane ? lc : rc;
            if (b < plane) nL += c;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                dst[k] = fminf(dst[k], fromOrderedBits(h[hidx(a, b, 7 + k)]));
                dst[3 + k] = fmaxf(dst[3 + k], fromOrderedBits(h[hidx(a, b, 10 + k)]));
            }
        }
        res.nL = nL;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            res.lLo[k] = lc[k];
            res.lHi[k] = lc[3 + k];
            res.rLo[k] = rc[k];
            res.rHi[k] = rc[3 + k];
        }
        res.lLo[3] = res.lHi[3] = res.rLo[3] = res.rHi[3] = 0.0f;
    }
    res.pad = 0;
    *out = res;
}

__device__ __forceinline__ bool goesLeft(const PrimRec& r, int pos, const SahTask& t, const SahResult& s,
                                         const AxisMap& m) {
    if (s.axis < 0) return pos - t.start < s.nL;
    const int a = s.axis;
    return binOf(r.c[a], m.lo[a], m.scale[a]) < s.plane;
}

// Block-wide exclusive scan of one int per thread (kBlock threads); returns the
// prefix and writes the total to *total.
__device__ __forceinline__ int blockExclusive(int v, int* lds, int* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) lds[w] = x;
    __syncthreads();
    int base = 0;
    for (int k = 0; k < w; ++k) base += lds[k];
    const int all = lds[0] + lds[1] + lds[2] + lds[3];
    __syncthreads();
    *total = all;
    return base + x - v;
}

// ---- kernels ------------------------------------------------------------------------------

template <typename L>
__global__ __launch_bounds__(kBlock) void sahPrepKernel(const typename L::Item* __restrict__ items, int n,
                                                        PrimRec* __restrict__ rec, uint32_t* __restrict__ cb,
                                                        int* __restrict__ parent) {
    float lo0 = kInf, lo1 = kInf, lo2 = kInf, hi0 = -kInf, hi1 = -kInf, hi2 = -kInf;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kBlock * gridDim.x) {
        const typename L::Item it = items[i];
        const V3 c = L::centroid(it);
        const Box b = L::bounds(it);
        PrimRec r;
        r.c[0] = c.x;
        r.c[1] = c.y;
        r.c[2] = c.z;
        r.lo[0] = b.lo.x;
        r.lo[1] = b.lo.y;
        r.lo[2] = b.lo.z;
        r.hi[0] = b.hi.x;
        r.hi[1] = b.hi.y;
        r.hi[2] = b.hi.z;
        r.id = i;
        rec[i] = r;
        lo0 = fminf(lo0, c.x);
        lo1 = fminf(lo1, c.y);
        lo2 = fminf(lo2, c.z);
        hi0 = fmaxf(hi0, c.x);
        hi1 = fmaxf(hi1, c.y);
        hi2 = fmaxf(hi2, c.z);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) parent[0] = -1;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lo0 = fminf(lo0, __shfl_xor(lo0, off, 64));
        lo1 = fminf(lo1, __shfl_xor(lo1, off, 64));
        lo2 = fminf(lo2, __shfl_xor(lo2, off, 64));
        hi0 = fmaxf(hi0, __shfl_xor(hi0, off, 64));
        hi1 = fmaxf(hi1, __shfl_xor(hi1, off, 64));
        hi2 = fmaxf(hi2, __shfl_xor(hi2, off, 64));
    }
    // the workgroup's waves meet in LDS: one set of six atomics per workgroup (all
    // waves on six words serialize at the L2)
    __shared__ float wb[kBlock / 64][6];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        wb[w][0] = lo0;
        wb[w][1] = lo1;
        wb[w][2] = lo2;
        wb[w][3] = hi0;
        wb[w][4] = hi1;
        wb[w][5] = hi2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlock / 64; ++k) {
            lo0 = fminf(lo0, wb[k][0]);
            lo1 = fminf(lo1, wb[k][1]);
            lo2 = fminf(lo2, wb[k][2]);
            hi0 = fmaxf(hi0, wb[k][3]);
            hi1 = fmaxf(hi1, wb[k][4]);
            hi2 = fmaxf(hi2, wb[k][5]);
        }
        atomicMin(&cb[0], orderedBits(lo0));
        atomicMin(&cb[1], orderedBits(lo1));
        atomicMin(&cb[2], orderedBits(lo2));
        atomicMax(&cb[3], orderedBits(hi0));
        atomicMax(&cb[4], orderedBits(hi1));
        atomicMax(&cb[5], orderedBits(hi2));
    }
}

__global__ void sahResetKernel(uint32_t* __restrict__ cb, int* __restrict__ arrive, int internal) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i == 8) cb[8] = 0u;  // partition error counter
    if (i < 3) cb[i] = orderedBits(kInf);
    if (i >= 3 && i < 6) cb[i] = orderedBits(-kInf);
    for (int j = i; j < internal; j += kBlock * gridDim.x) arrive[j] = 0;
}

__global__ void sahClearBinsKernel(uint32_t* __restrict__ bins, int tasks) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= tasks * kHist) return;
    const int v = (i % kHist) % kBinVals;
    bins[i] = v == 0 ? 0u : ((v >= 1 && v <= 3) || (v >= 7 && v <= 9)) ? orderedBits(kInf) : orderedBits(-kInf);
}

// One workgroup per chunk of a large task: LDS histogram, then one global flush.
__global__ __launch_bounds__(kBlock) void sahBinKernel(const PrimRec* __restrict__ rec,
                                                       const SahTask* __restrict__ tasks,
                                                       const Chunk* __restrict__ chunks,
                                                       uint32_t* __restrict__ bins) {
    __shared__ uint32_t h[kHist];
    const Chunk ch = chunks[blockIdx.x];
    const SahTask t = tasks[ch.task];
    const AxisMap m = axisMap(t);
    histInit(h);
    __syncthreads();
    histAddRange(h, rec + ch.start, ch.count, m);
    __syncthreads();
    uint32_t* g = bins + static_cast<size_t>(ch.task) * kHist;
    for (int i = threadIdx.x; i < kHist; i += kBlock) {
        const int v = i % kBinVals;
        if (h[i - v] == 0u) continue;  // empty bin
        const uint32_t x = h[i];
        if (v == 0) {
            atomicAdd(&g[i], x);
        } else if ((v >= 1 && v <= 3) || (v >= 7 && v <= 9)) {
            atomicMin(&g[i], x);
        } else {
            atomicMax(&g[i], x);
        }
    }
}

// One wave per large task.
__global__ __launch_bounds__(64) void sahSplitKernel(const uint32_t* __restrict__ bins,
                                                     const SahTask* __restrict__ tasks,
                                                     SahResult* __restrict__ results) {
    const SahTask t = tasks[blockIdx.x];
    chooseSplit(bins + static_cast<size_t>(blockIdx.x) * kHist, t, axisMap(t), &results[blockIdx.x]);
}

__global__ __launch_bounds__(kBlock) void sahPartCountKernel(const PrimRec* __restrict__ rec,
                                                             const SahTask* __restrict__ tasks,
                                                             const SahResult* __restrict__ results,
                                                             const Chunk* __restrict__ chunks,
                                                             int* __restrict__ chunkLeft) {
    __shared__ int lds[4];
    const Chunk ch = chunks[blockIdx.x];
    const SahTask t = tasks[ch.task];
    const SahResult s = results[ch.task];
    const AxisMap m = axisMap(t);
    int c = 0;
    for (int i = threadIdx.x; i < ch.count; i += kBlock) c += goesLeft(rec[ch.start + i], ch.start + i, t, s, m) ? 1 : 0;
    int total = 0;
    blockExclusive(c, lds, &total);
    if (threadIdx.x == 0) chunkLeft[blockIdx.x] = total;
}

// One thread per large task: exclusive offsets of its chunks' left counts
// (chunks of one task are consecutive, first[task] .. first[task+1]).
__global__ void sahChunkScanKernel(const int* __restrict__ first, int tasks, const int* __restrict__ chunkLeft,
                                   int* __restrict__ chunkOff) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= tasks) return;
    int acc = 0;
    for (int c = first[t]; c < first[t + 1]; ++c) {
        chunkOff[c] = acc;
        acc += chunkLeft[c];
    }
}

__global__ __launch_bounds__(kBlock) void sahPartScatterKernel(const PrimRec* __restrict__ rec,
                                                               PrimRec* __restrict__ out,
                                                               const SahTask* __restrict__ tasks,
                                                               const SahResult* __restrict__ results,
                                                               const Chunk* __restrict__ chunks,
                                                               const int* __restrict__ chunkOff,
                                                               int* __restrict__ err) {
    __shared__ int lds[4];
    const Chunk ch = chunks[blockIdx.x];
    const SahTask t = tasks[ch.task];
    const SahResult s = results[ch.task];
    const AxisMap m = axisMap(t);
    const int base = threadIdx.x * kPerThread;
    unsigned flags = 0u;
    int c = 0;
    for (int j = 0; j < kPerThread; ++j) {
        const int i = base + j;
        if (i < ch.count && goesLeft(rec[ch.start + i], ch.start + i, t, s, m)) {
            flags |= 1u << j;
            ++c;
        }
    }
    int total = 0;
    int lrank = blockExclusive(c, lds, &total);
    const int leftBase = t.start + chunkOff[blockIdx.x];
    const int chunkPos = ch.start - t.start;
    const int rightBase = t.start + s.nL + (chunkPos - chunkOff[blockIdx.x]);
    int rrank = base - lrank;
    const int leftEnd = t.start + s.nL, rightEnd = t.start + t.count;
    for (int j = 0; j < kPerThread; ++j) {
        const int i = base + j;
        if (i >= ch.count) break;
        const PrimRec r = rec[ch.start + i];
        const int dst = (flags & (1u << j)) ? leftBase + lrank++ : rightBase + rrank++;
        const bool ok = (flags & (1u << j)) ? dst < leftEnd : (dst >= leftEnd && dst < rightEnd);
        if (ok) {
            out[dst] = r;
        } else {
            atomicAdd(err, 1);  // counts disagree with the histogram: reported, never written
        }
    }
}

// One workgroup per task of kSerial+1 .. kChunk primitives: histogram, plane and
// stable partition inside the workgroup.
__global__ __launch_bounds__(kBlock) void sahLocalKernel(const PrimRec* __restrict__ rec,
                                                         PrimRec* __restrict__ out,
                                                         const SahTask* __restrict__ tasks,
                                                         SahResult* __restrict__ results, int firstTask,
                                                         int* __restrict__ err) {
    __shared__ uint32_t h[kHist];
    __shared__ SahResult sres;
    __shared__ int lds[4];
    const int ti = firstTask + blockIdx.x;
    const SahTask t = tasks[ti];
    const AxisMap m = axisMap(t);
    histInit(h);
    __syncthreads();
    histAddRange(h, rec + t.start, t.count, m);
    __syncthreads();
    if (threadIdx.x < 64) chooseSplit(h, t, m, &sres);
    __syncthreads();
    const SahResult s = sres;
    if (threadIdx.x == 0) results[ti] = s;
    const int base = threadIdx.x * kPerThread;
    unsigned flags = 0u;
    int c = 0;
    for (int j = 0; j < kPerThread; ++j) {
        const int i = base + j;
        if (i < t.count && goesLeft(rec[t.start + i], t.start + i, t, s, m)) {
            flags |= 1u << j;
            ++c;
        }
    }
    int total = 0;
    int lrank = blockExclusive(c, lds, &total);
    int rrank = base - lrank;
    for (int j = 0; j < kPerThread; ++j) {
        const int i = base + j;
        if (i >= t.count) break;
        const PrimRec r = rec[t.start + i];
        const bool left = (flags & (1u << j)) != 0;
        const int k = left ? lrank++ : rrank++;
        if (left ? k < s.nL : k < t.count - s.nL) {
            out[t.start + (left ? 0 : s.nL) + k] = r;
        } else {
            atomicAdd(err, 1);  // counts disagree with the histogram: reported, never written
        }
    }
}

// Internal nodes decided on the host (levels of large and local tasks).
__global__ void sahWriteNodesKernel(const NodeTriple* __restrict__ tri, int count, Node2* __restrict__ nodes,
                                    int* __restrict__ parent, int* __restrict__ farEnd) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    const NodeTriple t = tri[i];
    nodes[t.id].left = t.left;
    nodes[t.id].right = t.right;
    parent[t.left] = t.id;
    parent[t.right] = t.id;
    farEnd[t.id] = t.far;
}

struct SerialTask {
    int start, count, node, buf;
};

// One thread per final range of at most kSerial primitives: its leaf order (the
// scene indices, read from whichever record buffer holds the range) and, for two
// or more primitives, a balanced subtree of count/2 splits.
__global__ void sahSerialKernel(const PrimRec* __restrict__ recA, const PrimRec* __restrict__ recB,
                                const SerialTask* __restrict__ tasks, int count, int n,
                                uint32_t* __restrict__ order, Node2* __restrict__ nodes,
                                int* __restrict__ parent, int* __restrict__ farEnd) {
    const int ti = blockIdx.x * kBlock + threadIdx.x;
    if (ti >= count) return;
    const SerialTask t = tasks[ti];
    const PrimRec* src = t.buf ? recB : recA;
    for (int i = 0; i < t.count; ++i) order[t.start + i] = static_cast<uint32_t>(src[t.start + i].id);
    if (t.count < 2) return;
    int stS[8], stE[8], stId[8];
    int sp = 0;
    stS[0] = t.start;
    stE[0] = t.start + t.count - 1;
    stId[0] = t.node;
    sp = 1;
    while (sp > 0) {
        --sp;
        const int s = stS[sp], e = stE[sp], id = stId[sp];
        const int g = s + (e - s + 1) / 2 - 1;
        const int l = (g == s) ? n - 1 + s : g;
        const int r = (g + 1 == e) ? n - 1 + e : g + 1;
        nodes[id].left = l;
        nodes[id].right = r;
        parent[l] = id;
        parent[r] = id;
        farEnd[id] = id == s ? e : s;
        if (g > s) {
            stS[sp] = s;
            stE[sp] = g;
            stId[sp] = g;
            ++sp;
        }
        if (e > g + 1) {
            stS[sp] = g + 1;
            stE[sp] = e;
            stId[sp] = g + 1;
            ++sp;
        }
    }
}

struct HostTask {
    int start, count, node;
    float cLo[3], cHi[3];
};

// Pinned host staging for the level loop's tables (tasks, chunks, chunk index,
// results, error flag).  A copy from pageable memory goes through a driver bounce
// buffer and the loop makes five per level; from pinned memory they are plain DMA.
// One block per host thread, grown on demand and kept for the process (no copy is
// in flight when it grows: every level ends with a stream synchronize).
char* pinnedStage(size_t bytes) {
    static thread_local char* p = nullptr;
    static thread_local size_t cap = 0;
    if (bytes > cap) {
        if (p) hipCheck(hipHostFree(p), "sah stage free");
        p = nullptr;
        cap = 0;
        hipCheck(hipHostMalloc(reinterpret_cast<void**>(&p), bytes, hipHostMallocDefault), "sah stage");
        cap = bytes;
    }
    return p;
}

}  // namespace

thread_local SahBuildStats tSahStats;

template <typename L>
static void buildSahT(BvhState& b, const typename L::Item* dSrc, int n, hipStream_t s, BuildTimingsC* t) {
    using Item = typename L::Item;
    b.n = n;
    b.nodeCount = 2 * n - 1;
    b.nodes.alloc(static_cast<size_t>(b.nodeCount));
    b.items.alloc(static_cast<size_t>(n) * sizeof(Item));
    b.primIds.alloc(static_cast<size_t>(n));
    b.primSlot.reserve(static_cast<size_t>(n));  // filled on first use (refits, adjoints)
    b.parent.alloc(static_cast<size_t>(b.nodeCount));
    b.arrive.alloc(static_cast<size_t>(n - 1));
    b.height.alloc(static_cast<size_t>(n - 1));
    b.farEnd.alloc(static_cast<size_t>(n - 1));
    b.sortWs.reserve(static_cast<size_t>(n));
    b.counters.reserve(16);
    uint32_t* cb = reinterpret_cast<uint32_t*>(b.counters.p);
    int* errDev = b.counters.p + 8;

    DevBuf<PrimRec> rec[2];
    rec[0].alloc(static_cast<size_t>(n));
    rec[1].alloc(static_cast<size_t>(n));
    DevBuf<uint32_t> order(static_cast<size_t>(n));

    const int maxTasks = n / (kSerial + 1) + 2;
    const int maxBig = n / (kChunk + 1) + 2;
    const int maxChunks = n / kChunk + maxBig + 2;
    DevBuf<SahTask> dTasks(static_cast<size_t>(maxTasks));
    DevBuf<SahResult> dRes(static_cast<size_t>(maxTasks));
    DevBuf<Chunk> dChunks(static_cast<size_t>(maxChunks));
    DevBuf<int> dFirst(static_cast<size_t>(maxBig + 1));
    DevBuf<int> dChunkLeft(static_cast<size_t>(maxChunks));
    DevBuf<int> dChunkOff(static_cast<size_t>(maxChunks));
    DevBuf<uint32_t> dBins(static_cast<size_t>(maxBig) * kHist);
    size_t stageBytes = 0;
    auto carve = [&](size_t bytes) {
        const size_t at = stageBytes;
        stageBytes += (bytes + 255) & ~static_cast<size_t>(255);
        return at;
    };
    const size_t oTasks = carve(sizeof(SahTask) * static_cast<size_t>(maxTasks));
    const size_t oRes = carve(sizeof(SahResult) * static_cast<size_t>(maxTasks));
    const size_t oChunks = carve(sizeof(Chunk) * static_cast<size_t>(maxChunks));
    const size_t oFirst = carve(sizeof(int) * static_cast<size_t>(maxBig + 1));
    const size_t oErr = carve(sizeof(int) * 8);
    const size_t oBounds = carve(sizeof(uint32_t) * 6);
    char* stage = pinnedStage(stageBytes);
    SahTask* hTasks = reinterpret_cast<SahTask*>(stage + oTasks);
    SahResult* hRes = reinterpret_cast<SahResult*>(stage + oRes);
    Chunk* hChunks = reinterpret_cast<Chunk*>(stage + oChunks);
    int* hFirst = reinterpret_cast<int*>(stage + oFirst);
    int* hErr = reinterpret_cast<int*>(stage + oErr);
    uint32_t* cbh = reinterpret_cast<uint32_t*>(stage + oBounds);

    EventSet<4> ev;
    ev.record(0, s);
    sahResetKernel<<<blocksFor(n - 1 > 256 ? n - 1 : 256, kBlock), kBlock, 0, s>>>(cb, b.arrive.p, n - 1);
    int grid = blocksFor(n, kBlock * 4);
    if (grid > 512) grid = 512;  // two workgroups per CU stream the items; few bounds atomics
    sahPrepKernel<L><<<grid, kBlock, 0, s>>>(dSrc, n, rec[0].p, cb, b.parent.p);
    hipCheck(hipGetLastError(), "sah prep");
    hipCheck(hipMemcpyAsync(cbh, cb, sizeof(uint32_t) * 6, hipMemcpyDeviceToHost, s), "sah bounds");
    ev.record(1, s);
    ev.sync(1);

    std::vector<NodeTriple> triples;
    triples.reserve(static_cast<size_t>(n / kSerial + 16) * 2);
    std::vector<SerialTask> serial;
    std::vector<HostTask> level, next;
    auto pushChild = [&](int start, int count, int node, const float* lo, const float* hi, int buf) {
        if (count <= kSerial) {
            serial.push_back({start, count, node, buf});
        } else {
            HostTask h{start, count, node, {lo[0], lo[1], lo[2]}, {hi[0], hi[1], hi[2]}};
            next.push_back(h);
        }
    };
    {
        float lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = fromOrderedBits(cbh[a]);
            hi[a] = fromOrderedBits(cbh[3 + a]);
        }
        pushChild(0, n, 0, lo, hi, 0);
        level.swap(next);
    }

    double binMs = 0.0, splitMs = 0.0, partMs = 0.0;
    SahBuildStats st{};
    int cur = 0;
    for (int lvl = 0; !level.empty(); ++lvl) {
        next.clear();
        const int nt = static_cast<int>(level.size());
        long long slots = 0;
        for (const HostTask& h : level) slots += h.count;
        ++st.levels;
        st.partitionedSlots += slots;
        if (lvl < kMedianLevel) st.binnedSlots += slots;
        if (nt > maxTasks) throw DeviceError("sah: task table overflow");
        std::fill(hRes, hRes + nt, SahResult{});
        if (lvl >= kMedianLevel) {
            for (int i = 0; i < nt; ++i) {
                SahResult& r = hRes[static_cast<size_t>(i)];
                r.axis = -1;
                r.nL = level[static_cast<size_t>(i)].count / 2;
                for (int k = 0; k < 3; ++k) {
                    r.lLo[k] = r.rLo[k] = level[static_cast<size_t>(i)].cLo[k];
                    r.lHi[k] = r.rHi[k] = level[static_cast<size_t>(i)].cHi[k];
                }
            }
        } else {
            // large tasks first, then local ones
            std::stable_partition(level.begin(), level.end(), [](const HostTask& h) { return h.count > kChunk; });
            int nBig = 0;
            while (nBig < nt && level[static_cast<size_t>(nBig)].count > kChunk) ++nBig;
            if (nBig > maxBig) throw DeviceError("sah: task table overflow");
            for (int i = 0; i < nt; ++i) {
                const HostTask& h = level[static_cast<size_t>(i)];
                SahTask& d = hTasks[static_cast<size_t>(i)];
                d.start = h.start;
                d.count = h.count;
                d.pad0 = d.pad1 = 0;
                for (int k = 0; k < 3; ++k) {
                    d.cLo[k] = h.cLo[k];
                    d.cHi[k] = h.cHi[k];
                }
                d.cLo[3] = d.cHi[3] = 0.0f;
            }
            int nChunks = 0;
            for (int i = 0; i < nBig; ++i) {
                hFirst[i] = nChunks;
                const HostTask& h = level[static_cast<size_t>(i)];
                for (int off = 0; off < h.count; off += kChunk) {
                    if (nChunks >= maxChunks) throw DeviceError("sah: task table overflow");
                    hChunks[nChunks++] = {i, h.start + off, std::min(kChunk, h.count - off), 0};
                }
            }
            hFirst[nBig] = nChunks;
            hipCheck(hipMemcpyAsync(dTasks.p, hTasks, sizeof(SahTask) * static_cast<size_t>(nt),
                                    hipMemcpyHostToDevice, s), "sah tasks");
            EventSet<4> lev;
            lev.record(0, s);
            if (nBig > 0) {
                hipCheck(hipMemcpyAsync(dChunks.p, hChunks, sizeof(Chunk) * static_cast<size_t>(nChunks),
                                        hipMemcpyHostToDevice, s), "sah chunks");
                hipCheck(hipMemcpyAsync(dFirst.p, hFirst, sizeof(int) * static_cast<size_t>(nBig + 1),
                                        hipMemcpyHostToDevice, s), "sah chunk index");
                sahClearBinsKernel<<<blocksFor(static_cast<long long>(nBig) * kHist, kBlock), kBlock, 0, s>>>(dBins.p, nBig);
                sahBinKernel<<<nChunks, kBlock, 0, s>>>(rec[cur].p, dTasks.p, dChunks.p, dBins.p);
            }
            if (nt > nBig)
                sahLocalKernel<<<nt - nBig, kBlock, 0, s>>>(rec[cur].p, rec[1 - cur].p, dTasks.p, dRes.p, nBig, errDev);
            lev.record(1, s);
            if (nBig > 0) sahSplitKernel<<<nBig, 64, 0, s>>>(dBins.p, dTasks.p, dRes.p);
            lev.record(2, s);
            if (nBig > 0) {
                sahPartCountKernel<<<nChunks, kBlock, 0, s>>>(rec[cur].p, dTasks.p, dRes.p, dChunks.p, dChunkLeft.p);
                sahChunkScanKernel<<<blocksFor(nBig, kBlock), kBlock, 0, s>>>(dFirst.p, nBig, dChunkLeft.p, dChunkOff.p);
                sahPartScatterKernel<<<nChunks, kBlock, 0, s>>>(rec[cur].p, rec[1 - cur].p, dTasks.p, dRes.p,
                                                               dChunks.p, dChunkOff.p, errDev);
            }
            hipCheck(hipGetLastError(), "sah level");
            lev.record(3, s);
            hipCheck(hipMemcpyAsync(hRes, dRes.p, sizeof(SahResult) * static_cast<size_t>(nt),
                                    hipMemcpyDeviceToHost, s), "sah results");
            hipCheck(hipMemcpyAsync(hErr, errDev, sizeof(int), hipMemcpyDeviceToHost, s), "sah check");
            hipCheck(hipStreamSynchronize(s), "sah level sync");
            if (*hErr != 0) throw DeviceError("sah: partition counts disagree with the bin histogram");
            binMs += lev.ms(0, 1);
            splitMs += lev.ms(1, 2);
            partMs += lev.ms(2, 3);
            cur = 1 - cur;
        }
        for (int i = 0; i < nt; ++i) {
            const HostTask& h = level[static_cast<size_t>(i)];
            const SahResult& r = hRes[static_cast<size_t>(i)];
            int nL = r.nL;
            if (nL < 1 || nL >= h.count) throw DeviceError("sah: empty side in a split");
            const int g = h.start + nL - 1;
            const int e = h.start + h.count - 1;
            const int left = (g == h.start) ? n - 1 + g : g;
            const int right = (g + 1 == e) ? n - 1 + e : g + 1;
            triples.push_back({h.node, left, right, h.node == h.start ? e : h.start});
            pushChild(h.start, nL, left, r.lLo, r.lHi, cur);
            pushChild(g + 1, h.count - nL, right, r.rLo, r.rHi, cur);
        }
        level.swap(next);
        if (lvl > 2 * kMedianLevel + 64) throw DeviceError("sah: level limit exceeded");
    }

    for (const SerialTask& sk : serial) st.serialSlots += sk.count;
    tSahStats = st;
    // node records decided on the host, then the serial tails
    DevBuf<NodeTriple> dTri(triples.size() + 1);
    DevBuf<SerialTask> dSerial(serial.size() + 1);
    if (!triples.empty()) {
        hipCheck(hipMemcpyAsync(dTri.p, triples.data(), sizeof(NodeTriple) * triples.size(), hipMemcpyHostToDevice, s),
                 "sah nodes");
        sahWriteNodesKernel<<<blocksFor(static_cast<long long>(triples.size()), kBlock), kBlock, 0, s>>>(
            dTri.p, static_cast<int>(triples.size()), b.nodes.p, b.parent.p, b.farEnd.p);
    }
    hipCheck(hipMemcpyAsync(dSerial.p, serial.data(), sizeof(SerialTask) * serial.size(), hipMemcpyHostToDevice, s),
             "sah serial");
    sahSerialKernel<<<blocksFor(static_cast<long long>(serial.size()), kBlock), kBlock, 0, s>>>(
        rec[0].p, rec[1].p, dSerial.p, static_cast<int>(serial.size()), n, order.p, b.nodes.p, b.parent.p,
        b.farEnd.p);
    hipCheck(hipGetLastError(), "sah serial");
    ev.record(2, s);

    finishTopology(b, dSrc, order.p, s);
    int depth = 0;
    hipCheck(hipMemcpyAsync(&depth, b.height.p, sizeof(int), hipMemcpyDeviceToHost, s), "depth");
    ev.record(3, s);
    ev.sync(3);
    b.depth = depth;
    if (t) {
        t->mortonMs = ev.ms(0, 1);
        t->sortMs = 0.0;
        t->treeMs = ev.ms(1, 2);
        t->scatterMs = 0.0;
        t->refitMs = ev.ms(2, 3);
        t->sahBinMs = binMs;
        t->sahSplitMs = splitMs;
        t->sahPartMs = partMs;
        t->totalMs = ev.ms(0, 3);
    }
    b.primSlotValid = false;
    b.depthValid = false;
    b.wideValid = false;
    b.wideCount = 0;
}

void buildSahDevice(BvhState& b, int kind, const void* dItems, int n, hipStream_t s, BuildTimingsC* t) {
    if (n < 2) throw ArgError("build: at least 2 primitives required");
    b.kind = kind;
    b.builder = kBuildSah;
    switch (kind) {
        case kPrimTri:
            buildSahT<TriLeaf>(b, static_cast<const Tri*>(dItems), n, s, t);
            break;
        case kPrimPoint:
            buildSahT<PointLeaf>(b, static_cast<const V3*>(dItems), n, s, t);
            break;
        case kPrimGauss:
            buildSahT<GaussLeaf>(b, static_cast<const Gaussian*>(dItems), n, s, t);
            break;
        default:
            throw ArgError("build: unknown primitive kind");
    }
}

}  // namespace qr
