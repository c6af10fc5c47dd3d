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
    int start, count, node, meta;
    float cLo[4], cHi[4];
};

struct SahResult {
    int axis, plane, nL, pad;
    float lLo[4], lHi[4], rLo[4], rHi[4];
};

struct Chunk {
    int task, start, count, pad;
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

template <int kBT = kBlock>
__device__ __forceinline__ void histAddRange(uint32_t* h, const PrimRec* __restrict__ rec, int count,
                                             const AxisMap& m) {
    const int per = (count + kBT - 1) / kBT;
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

// Plane choice by one wave: lane b holds bin b (lanes 32..63 repeat lanes 0..31);
// inclusive prefix and suffix unions of the bins' counts and boxes (five shuffle
// steps each) give every plane p = 1..31 its left side (bins 0..p-1) and right side
// (bins p..31) at lane p, which scores it on each usable axis; lane 0 writes the
// result.  Counts and min / max are exact, so the costs are the same bits as summing
// the bins plane by plane.  h is the task's histogram (LDS or global).
__device__ void chooseSplit(const uint32_t* h, const SahTask& t, const AxisMap& m, SahResult* out) {
    const int lane = threadIdx.x & 63;
    const int bin = lane & (kBins - 1);
    const int count = t.count;
    const int minSide = count / 8 > 1 ? count / 8 : 1;
    float best = kInf;
    int bestKey = 0x7fffffff;  // axis * 32 + plane
    for (int a = 0; a < 3; ++a) {
        if (!m.use[a]) continue;
        const int c = static_cast<int>(h[hidx(a, bin, 0)]);
        float v[6];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            v[k] = fromOrderedBits(h[hidx(a, bin, 1 + k)]);
            v[3 + k] = fromOrderedBits(h[hidx(a, bin, 4 + k)]);
        }
        int pc = c, sc = c;
        float pv[6], sv[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) pv[k] = sv[k] = v[k];
#pragma unroll
        for (int off = 1; off < kBins; off <<= 1) {
            const int oc = __shfl_up(pc, off, kBins);
            const int uc = __shfl_down(sc, off, kBins);
            const bool up = bin >= off, down = bin + off < kBins;
            pc += up ? oc : 0;
            sc += down ? uc : 0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const float op = __shfl_up(pv[k], off, kBins);
                const float os = __shfl_down(sv[k], off, kBins);
                if (k < 3) {
                    pv[k] = up ? fminf(pv[k], op) : pv[k];
                    sv[k] = down ? fminf(sv[k], os) : sv[k];
                } else {
                    pv[k] = up ? fmaxf(pv[k], op) : pv[k];
                    sv[k] = down ? fmaxf(sv[k], os) : sv[k];
                }
            }
        }
        // the left side of plane `bin` is the prefix through bin - 1
        const int nL = __shfl_up(pc, 1, kBins);
        float l[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) l[k] = __shfl_up(pv[k], 1, kBins);
        const int nR = sc;
        if (bin >= 1 && nL >= minSide && nR >= minSide) {
            const float cost = areaOf(l[0], l[1], l[2], l[3], l[4], l[5]) * static_cast<float>(nL) +
                               areaOf(sv[0], sv[1], sv[2], sv[3], sv[4], sv[5]) * static_cast<float>(nR);
            const int key = a * kBins + bin;
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
    SahResult res;
    res.pad = 0;
    if (bestKey == 0x7fffffff) {
        if (lane != 0) return;
        res.axis = -1;
        res.plane = 0;
        res.nL = count / 2;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            res.lLo[k] = res.rLo[k] = t.cLo[k];
            res.lHi[k] = res.rHi[k] = t.cHi[k];
        }
        *out = res;
        return;
    }
    // the chosen plane's two sides: counts and centroid boxes of the bins left and
    // right of it, reduced across the 32 bin lanes (a min / max of ordered bits is
    // exact and order-free; empty bins take the identities)
    const int a = bestKey / kBins, plane = bestKey % kBins;
    const uint32_t c = h[hidx(a, bin, 0)];
    const bool isLeft = bin < plane;
    int nL = isLeft ? static_cast<int>(c) : 0;
    uint32_t v[12];  // left min xyz, left max xyz, right min xyz, right max xyz
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const uint32_t lo = c ? h[hidx(a, bin, 7 + k)] : 0xFFFFFFFFu;
        const uint32_t hi = c ? h[hidx(a, bin, 10 + k)] : 0u;
        v[k] = isLeft ? lo : 0xFFFFFFFFu;
        v[3 + k] = isLeft ? hi : 0u;
        v[6 + k] = isLeft ? 0xFFFFFFFFu : lo;
        v[9 + k] = isLeft ? 0u : hi;
    }
#pragma unroll
    for (int off = kBins / 2; off > 0; off >>= 1) {
        nL += __shfl_xor(nL, off, 64);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            const uint32_t o = __shfl_xor(v[k], off, 64);
            const bool isMin = k < 3 || (k >= 6 && k < 9);
            v[k] = isMin ? (o < v[k] ? o : v[k]) : (o > v[k] ? o : v[k]);
        }
    }
    if (lane != 0) return;
    res.axis = a;
    res.plane = plane;
    res.nL = nL;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        res.lLo[k] = fromOrderedBits(v[k]);
        res.lHi[k] = fromOrderedBits(v[3 + k]);
        res.rLo[k] = fromOrderedBits(v[6 + k]);
        res.rHi[k] = fromOrderedBits(v[9 + k]);
    }
    res.lLo[3] = res.lHi[3] = res.rLo[3] = res.rHi[3] = 0.0f;
    *out = res;
}

// Coordinate `axis` of a record held in registers, by selection (a dynamic subscript
// into a register copy would go through scratch).
__device__ __forceinline__ float axisCoord(const PrimRec& r, int axis) {
    return axis == 0 ? r.c[0] : axis == 1 ? r.c[1] : r.c[2];
}

__device__ __forceinline__ float axisSel(const float* v, int axis) {
    return axis == 0 ? v[0] : axis == 1 ? v[1] : v[2];
}

// Side test of one record for split s of task t (by plane, or by position for a
// split without a plane); c is the record's coordinate on the split axis.
__device__ __forceinline__ bool goesLeft(float c, int pos, const SahTask& t, const SahResult& s, const AxisMap& m) {
    if (s.axis < 0) return pos - t.start < s.nL;
    return binOf(c, axisSel(m.lo, s.axis), axisSel(m.scale, s.axis)) < s.plane;
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

// Stable partition of the records [begin, begin + cnt) of task t by split s into
// dst: the range's left records go to leftDst onwards, its right records to rightDst
// onwards, each side in record order.  Wave w takes the w-th quarter of the range and
// reads it lane-strided (coalesced): one ballot pass counts its lefts, the waves'
// counts meet in LDS (one barrier), and a second pass writes every record at its
// rank.  Called by the whole workgroup.  A destination outside the task's side
// (counts that disagree with the histogram) is reported through err, never written.
template <int kBT = kBlock>
__device__ void partitionRange(const PrimRec* __restrict__ src, PrimRec* __restrict__ dst, int begin, int cnt,
                               int leftDst, int rightDst, const SahTask& t, const SahResult& s, const AxisMap& m,
                               int* ldsWave, int* err) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    constexpr int kWaves = kBT / 64;
    const int per = (cnt + kWaves - 1) / kWaves;
    const int wBeg = w * per < cnt ? w * per : cnt;
    const int wEnd = wBeg + per < cnt ? wBeg + per : cnt;
    const unsigned long long below = (1ull << lane) - 1ull;
    int lefts = 0;
    for (int r = wBeg; r < wEnd; r += 64) {
        const int i = r + lane;
        const bool left = i < wEnd && goesLeft(axisCoord(src[begin + i], s.axis), begin + i, t, s, m);
        lefts += __popcll(__ballot(left));
    }
    if (lane == 0) ldsWave[w] = lefts;
    __syncthreads();
    int leftBefore = 0;
    for (int v = 0; v < w; ++v) leftBefore += ldsWave[v];
    __syncthreads();  // ldsWave is free for the caller's next use
    int runL = leftDst + leftBefore;
    int runR = rightDst + (wBeg - leftBefore);
    const int leftEnd = t.start + s.nL, rightEnd = t.start + t.count;
    bool bad = false;
    for (int r = wBeg; r < wEnd; r += 64) {
        const int i = r + lane;
        const bool valid = i < wEnd;
        PrimRec rec{};
        bool left = false;
        if (valid) {
            rec = src[begin + i];
            left = goesLeft(axisCoord(rec, s.axis), begin + i, t, s, m);
        }
        const unsigned long long lm = __ballot(valid && left);
        const unsigned long long rm = __ballot(valid && !left);
        if (valid) {
            const int d = left ? runL + __popcll(lm & below) : runR + __popcll(rm & below);
            const bool ok = left ? (d >= t.start && d < leftEnd) : (d >= leftEnd && d < rightEnd);
            if (ok) {
                dst[d] = rec;
            } else {
                bad = true;
            }
        }
        runL += __popcll(lm);
        runR += __popcll(rm);
    }
    if (bad) atomicOr(err, 2);
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

// ---- device-driven split loop ------------------------------------------------------------
// Control words of the loop (device memory): the big-task counts of the two level
// buffers, the local and serial list lengths, the current level's chunk count, an
// error flag and the work counters.
enum : int {
    kCtlBig0 = 0,       // tasks in level buffer 0 (1: buffer 1)
    kCtlLocal = 2,      // tasks of kSerial+1 .. kChunk primitives, for the subtree kernel
    kCtlSerial = 3,     // ranges of at most kSerial primitives, for the serial kernel
    kCtlChunks = 4,     // chunks of the current level's big tasks
    kCtlError = 5,      // set when a task table overflows or a split is empty
    kCtlBinned = 6,     // primitive slots binned (summed over levels)
    kCtlParted = 7,     // primitive slots partitioned
    kCtlSerialSlots = 8,
    kCtlMaxLevel = 9,   // deepest level that split a task
    kCtlSmall = 10,     // tasks of kSerial+1 .. kSmallTask primitives from the top subtree tier
    kCtlWords = 16,
};

struct SerialTask {
    int start, count, node, buf;
};

// SahTask.meta: bit 0 = the record buffer holding the task's range, bits 1.. = level.
__device__ __forceinline__ int metaOf(int buf, int level) { return buf | (level << 1); }

// Where the split loop's tables live (all sized for the worst case of n primitives:
// big tasks of one level and local tasks are disjoint ranges of more than kChunk /
// kSerial primitives, serial ranges are disjoint and non-empty).
struct SahTables {
    PrimRec* rec[2];
    SahTask* big[2];
    SahTask* local;
    SahTask* small;
    SerialTask* serial;
    Chunk* chunks;
    int* first;
    int* chunkBins;  // per chunk: its primitives' counts per axis and bin (3 x kBins)
    int* chunkOff;
    uint32_t* bins;
    SahResult* results;
    int* ctl;
    int maxBig, maxLocal, maxSerial, maxChunks;
};

// Records the split of task (start, count, node) after its first nL positions: the
// node's children in the LBVH numbering, parent links and the node's far end.
__device__ __forceinline__ void writeSplit(int n, int start, int count, int node, int nL, Node2* __restrict__ nodes,
                                           int* __restrict__ parent, int* __restrict__ farEnd, int* lId, int* rId) {
    const int g = start + nL - 1;
    const int e = start + count - 1;
    const int l = (g == start) ? n - 1 + g : g;
    const int r = (g + 1 == e) ? n - 1 + e : g + 1;
    nodes[node].left = l;
    nodes[node].right = r;
    parent[l] = node;
    parent[r] = node;
    farEnd[node] = node == start ? e : start;
    *lId = l;
    *rId = r;
}

// The tables' two-entry arrays indexed by a selection, not a dynamic subscript (the
// tables are a kernel argument; a subscript would copy them to scratch).
__device__ __forceinline__ PrimRec* recOf(const SahTables& T, int buf) { return buf ? T.rec[1] : T.rec[0]; }
__device__ __forceinline__ SahTask* bigOf(const SahTables& T, int cur) { return cur ? T.big[1] : T.big[0]; }

__device__ __forceinline__ void pushSerial(const SahTables& T, int start, int count, int node, int buf) {
    const int at = atomicAdd(&T.ctl[kCtlSerial], 1);
    if (at < T.maxSerial) {
        T.serial[at] = {start, count, node, buf};
    } else {
        atomicOr(&T.ctl[kCtlError], 1);
    }
}

// A split's child task: side 0 takes the left centroid box, side 1 the right.
__device__ __forceinline__ SahTask childTask(int start, int count, int node, int meta, const SahResult& s, int side) {
    SahTask t;
    t.start = start;
    t.count = count;
    t.node = node;
    t.meta = meta;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        t.cLo[k] = side ? s.rLo[k] : s.lLo[k];
        t.cHi[k] = side ? s.rHi[k] : s.lHi[k];
    }
    t.cLo[3] = t.cHi[3] = 0.0f;
    return t;
}

__device__ __forceinline__ SahTask makeTask(int start, int count, int node, int meta, const float* lo, const float* hi) {
    SahTask t;
    t.start = start;
    t.count = count;
    t.node = node;
    t.meta = meta;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        t.cLo[k] = lo[k];
        t.cHi[k] = hi[k];
    }
    t.cLo[3] = t.cHi[3] = 0.0f;
    return t;
}

// Local tasks of at most this many primitives go to the one-wave subtree tier.
constexpr int kSmallTask = 512;

// A child of a big task: serial range, small or local task, or next-level big task.
__device__ __forceinline__ void routeChild(const SahTables& T, int nextBuf, int start, int count, int node, int buf,
                                           int level, const float* lo, const float* hi) {
    if (count <= kSerial) {
        pushSerial(T, start, count, node, buf);
    } else if (count <= kSmallTask) {
        const int at = atomicAdd(&T.ctl[kCtlSmall], 1);
        if (at < T.maxLocal) {
            T.small[at] = makeTask(start, count, node, metaOf(buf, level), lo, hi);
        } else {
            atomicOr(&T.ctl[kCtlError], 1);
        }
    } else if (count <= kChunk) {
        const int at = atomicAdd(&T.ctl[kCtlLocal], 1);
        if (at < T.maxLocal) {
            T.local[at] = makeTask(start, count, node, metaOf(buf, level), lo, hi);
        } else {
            atomicOr(&T.ctl[kCtlError], 1);
        }
    } else {
        const int at = atomicAdd(&T.ctl[kCtlBig0 + nextBuf], 1);
        if (at < T.maxBig) {
            bigOf(T, nextBuf)[at] = makeTask(start, count, node, metaOf(buf, level), lo, hi);
        } else {
            atomicOr(&T.ctl[kCtlError], 1);
        }
    }
}

// The root: one task over [0, n) with the centroid bounds of the prep kernel.
__global__ void sahRootKernel(SahTables T, const uint32_t* __restrict__ cb, int n) {
    if (threadIdx.x != 0) return;
    for (int i = 0; i < kCtlWords; ++i) T.ctl[i] = 0;
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = fromOrderedBits(cb[a]);
        hi[a] = fromOrderedBits(cb[3 + a]);
    }
    routeChild(T, 0, 0, n, 0, 0, 0, lo, hi);
}

// Level plan (one workgroup): the current big tasks' chunk list (chunks of one task
// consecutive, first[task] .. first[task + 1]); clears the count of the buffer this
// level's children go to.  Per group of kBlock tasks, an exclusive scan of their chunk
// counts, then every thread writes chunks, each finding its task by binary search of
// the group's chunk offsets in LDS.
__global__ __launch_bounds__(kBlock) void sahPlanKernel(SahTables T, int cur, int median) {
    __shared__ int lds[4];
    __shared__ int sEx[kBlock], sStart[kBlock], sCount[kBlock];
    const int nt = T.ctl[kCtlBig0 + cur];
    int carry = 0;
    for (int base = 0; base < nt && !median; base += kBlock) {
        const int i = base + threadIdx.x;
        int c = 0;
        if (i < nt) {
            const SahTask t = bigOf(T, cur)[i];
            c = (t.count + kChunk - 1) / kChunk;
            sStart[threadIdx.x] = t.start;
            sCount[threadIdx.x] = t.count;
        }
        int total = 0;
        const int ex = blockExclusive(c, lds, &total);
        if (i < nt) {
            T.first[i] = carry + ex;
            sEx[threadIdx.x] = ex;
        }
        __syncthreads();
        const int g = min(kBlock, nt - base);
        for (int k = threadIdx.x; k < total && carry + k < T.maxChunks; k += kBlock) {
            int lo = 0, hi = g - 1;  // the last task of the group whose first chunk is <= k
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (sEx[mid] <= k) lo = mid; else hi = mid - 1;
            }
            const int j = k - sEx[lo];
            T.chunks[carry + k] = {base + lo, sStart[lo] + j * kChunk, min(kChunk, sCount[lo] - j * kChunk), 0};
        }
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (carry > T.maxChunks) {
            atomicOr(&T.ctl[kCtlError], 1);
            carry = 0;
        }
        if (!median) T.first[nt] = carry;
        T.ctl[kCtlChunks] = median ? 0 : carry;
        T.ctl[kCtlBig0 + (1 - cur)] = 0;
    }
}

__global__ void sahClearBinsKernel(SahTables T, int cur) {
    const long long words = static_cast<long long>(T.ctl[kCtlBig0 + cur]) * kHist;
    for (long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; i < words;
         i += static_cast<long long>(gridDim.x) * kBlock) {
        const int v = static_cast<int>(i % kHist) % kBinVals;
        T.bins[i] = v == 0 ? 0u : ((v >= 1 && v <= 3) || (v >= 7 && v <= 9)) ? orderedBits(kInf) : orderedBits(-kInf);
    }
}

// One workgroup per chunk of a large task: LDS histogram, then one global flush.
__global__ __launch_bounds__(kBlock) void sahBinKernel(SahTables T, int cur) {
    if (static_cast<int>(blockIdx.x) >= T.ctl[kCtlChunks]) return;
    __shared__ uint32_t h[kHist];
    const Chunk ch = T.chunks[blockIdx.x];
    const SahTask t = bigOf(T, cur)[ch.task];
    const AxisMap m = axisMap(t);
    histInit(h);
    __syncthreads();
    histAddRange(h, recOf(T, t.meta & 1) + ch.start, ch.count, m);
    __syncthreads();
    uint32_t* g = T.bins + static_cast<size_t>(ch.task) * kHist;
    // the chunk's own bin counts: its left count for any plane, without a re-read
    for (int i = threadIdx.x; i < 3 * kBins; i += kBlock)
        T.chunkBins[static_cast<size_t>(blockIdx.x) * 3 * kBins + i] = static_cast<int>(h[i * kBinVals]);
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
    if (threadIdx.x == 0) atomicAdd(&T.ctl[kCtlBinned], ch.count);
}

__device__ __forceinline__ SahResult medianResult(const SahTask& t) {
    SahResult res;
    res.axis = -1;
    res.plane = 0;
    res.nL = t.count / 2;
    res.pad = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        res.lLo[k] = res.rLo[k] = t.cLo[k];
        res.lHi[k] = res.rHi[k] = t.cHi[k];
    }
    return res;
}

// One wave per large task.
__global__ __launch_bounds__(64) void sahSplitKernel(SahTables T, int cur, int median) {
    if (static_cast<int>(blockIdx.x) >= T.ctl[kCtlBig0 + cur]) return;
    const SahTask t = bigOf(T, cur)[blockIdx.x];
    if (median) {
        if (threadIdx.x == 0) T.results[blockIdx.x] = medianResult(t);
        return;
    }
    chooseSplit(T.bins + static_cast<size_t>(blockIdx.x) * kHist, t, axisMap(t), &T.results[blockIdx.x]);
}

// One wave per large task: exclusive offsets of its chunks' left counts, from the
// chunks' bin counts (the bins left of the chosen plane on its axis) or, for a split
// by position, from the chunk's place in the task; 64 chunks per step, a wave scan.
__global__ __launch_bounds__(64) void sahChunkScanKernel(SahTables T, int cur) {
    const int ti = blockIdx.x;
    if (ti >= T.ctl[kCtlBig0 + cur] || T.ctl[kCtlChunks] == 0) return;
    const int lane = threadIdx.x;
    const SahTask t = bigOf(T, cur)[ti];
    const SahResult r = T.results[ti];
    const int c0 = T.first[ti], c1 = T.first[ti + 1];
    int acc = 0;
    for (int base = c0; base < c1; base += 64) {
        const int c = base + lane;
        int left = 0;
        if (c < c1) {
            if (r.axis < 0) {
                const Chunk ch = T.chunks[c];
                const int before = t.start + r.nL - ch.start;
                left = before < 0 ? 0 : (before > ch.count ? ch.count : before);
            } else {
                const int* cnt = T.chunkBins + static_cast<size_t>(c) * 3 * kBins + r.axis * kBins;
                for (int b = 0; b < r.plane; ++b) left += cnt[b];
            }
        }
        int x = left;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (c < c1) T.chunkOff[c] = acc + x - left;
        acc += __shfl(x, 63, 64);
    }
    if (lane == 0 && acc != r.nL) atomicOr(&T.ctl[kCtlError], 2);  // the chunks disagree with the task's histogram
}

__global__ __launch_bounds__(kBlock) void sahPartScatterKernel(SahTables T, int cur) {
    if (static_cast<int>(blockIdx.x) >= T.ctl[kCtlChunks]) return;
    __shared__ int lds[4];
    const Chunk ch = T.chunks[blockIdx.x];
    const SahTask t = bigOf(T, cur)[ch.task];
    const SahResult s = T.results[ch.task];
    const AxisMap m = axisMap(t);
    const int chunkLeft = T.chunkOff[blockIdx.x];  // lefts of the task's earlier chunks
    const int chunkPos = ch.start - t.start;
    partitionRange(recOf(T, t.meta & 1), recOf(T, 1 - (t.meta & 1)), ch.start, ch.count, t.start + chunkLeft,
                   t.start + s.nL + (chunkPos - chunkLeft), t, s, m, lds, &T.ctl[kCtlError]);
    if (threadIdx.x == 0) atomicAdd(&T.ctl[kCtlParted], ch.count);
}

// One thread per large task: its node, and its children routed by size.
__global__ void sahAssignKernel(SahTables T, int cur, int median, int n, Node2* __restrict__ nodes,
                                int* __restrict__ parent, int* __restrict__ farEnd) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= T.ctl[kCtlBig0 + cur]) return;
    const SahTask t = bigOf(T, cur)[i];
    const SahResult r = T.results[i];
    int nL = r.nL;
    if (nL < 1 || nL >= t.count) {
        atomicOr(&T.ctl[kCtlError], 4);
        return;
    }
    const int level = t.meta >> 1;
    const int buf = median ? (t.meta & 1) : 1 - (t.meta & 1);
    atomicMax(&T.ctl[kCtlMaxLevel], level);
    int l, rr;
    writeSplit(n, t.start, t.count, t.node, nL, nodes, parent, farEnd, &l, &rr);
    routeChild(T, 1 - cur, t.start, nL, l, buf, level + 1, r.lLo, r.lHi);
    routeChild(T, 1 - cur, t.start + nL, t.count - nL, rr, buf, level + 1, r.rLo, r.rHi);
}

// Local tasks (kSerial+1 .. kChunk primitives) build their whole subtrees without a
// return to the host, in two tiers.  The top tier takes one local task per 256-lane
// workgroup and splits it from an LDS stack (histogram, plane choice and stable
// partition per task inside the workgroup) until its pieces have at most kSmallTask
// primitives; those go to the small list.  The bottom tier (sahSmallKernel) takes one
// small task per one-wave workgroup (many per CU, no cross-wave barriers) and splits
// it in LDS down to ranges of kSerial: the many small tasks of a subtree's lower
// levels run side by side instead of one after another in the workgroup that split
// their ancestors.
constexpr int kSubtreeStack = 64;  // a chain of count/8 sides from 4096 to 64 is 32 deep
constexpr int kSerialBatch = 64;

// Appends count buffered serial ranges to the global list with one counter bump.
__device__ __forceinline__ void flushSerial(const SahTables& T, const SerialTask* buf, int count) {
    if (count == 0) return;
    const int at = atomicAdd(&T.ctl[kCtlSerial], count);
    for (int i = 0; i < count; ++i) {
        if (at + i < T.maxSerial) {
            T.serial[at + i] = buf[i];
        } else {
            atomicOr(&T.ctl[kCtlError], 1);
        }
    }
}
template <int kBT>
__global__ __launch_bounds__(kBT) void sahSubtreeKernel(SahTables T, int n, Node2* __restrict__ nodes,
                                                        int* __restrict__ parent, int* __restrict__ farEnd) {
    __shared__ uint32_t h[kHist];
    __shared__ SahResult sres;
    __shared__ SahTask stk[kSubtreeStack];
    __shared__ SahTask curTask;
    __shared__ int ssp;
    __shared__ int lds[4];
    __shared__ SerialTask serBuf[kSerialBatch];
    int serN = 0;  // thread 0's count of buffered serial ranges
    int binned = 0, parted = 0, deepest = 0;
    // local tasks blockIdx.x, + gridDim.x, ...
    const int listLen = T.ctl[kCtlLocal];
    for (int task = blockIdx.x; task < listLen; task += gridDim.x) {
    if (threadIdx.x == 0) {
        stk[0] = T.local[task];
        ssp = 1;
    }
    __syncthreads();
    while (true) {
        if (threadIdx.x == 0) {
            if (ssp > 0) {
                curTask = stk[--ssp];
            } else {
                curTask.count = -1;
            }
        }
        __syncthreads();
        const SahTask t = curTask;
        if (t.count < 0) break;
        const int buf = t.meta & 1, level = t.meta >> 1;
        const bool median = level >= kMedianLevel;
        const PrimRec* src = recOf(T, buf);
        PrimRec* dst = recOf(T, 1 - buf);
        const AxisMap m = axisMap(t);
        if (!median) {
            histInit(h);
            __syncthreads();
            histAddRange<kBT>(h, src + t.start, t.count, m);
            __syncthreads();
            if (threadIdx.x < 64) chooseSplit(h, t, m, &sres);
        } else if (threadIdx.x == 0) {
            sres = medianResult(t);
        }
        __syncthreads();
        const SahResult s = sres;
        if (!median) {
            partitionRange<kBT>(src, dst, t.start, t.count, t.start, t.start + s.nL, t, s, m, lds,
                                &T.ctl[kCtlError]);
            binned += t.count;
            parted += t.count;
        }
        if (threadIdx.x == 0) {
            deepest = level > deepest ? level : deepest;
            const int nL = s.nL;
            if (nL < 1 || nL >= t.count) {
                atomicOr(&T.ctl[kCtlError], 4);
            } else {
                const int childBuf = median ? buf : 1 - buf;
                int l, r;
                writeSplit(n, t.start, t.count, t.node, nL, nodes, parent, farEnd, &l, &r);
                const int cs[2] = {t.start, t.start + nL}, cc[2] = {nL, t.count - nL}, ci[2] = {l, r};
#pragma unroll
                for (int side = 1; side >= 0; --side) {
                    if (cc[side] <= kSerial) {
                        // serial ranges collect in LDS and go out in batches: one global
                        // counter bump per batch instead of one per range
                        if (serN == kSerialBatch) {
                            flushSerial(T, serBuf, serN);
                            serN = 0;
                        }
                        serBuf[serN++] = {cs[side], cc[side], ci[side], childBuf};
                    } else if (cc[side] <= kSmallTask) {
                        const int at = atomicAdd(&T.ctl[kCtlSmall], 1);
                        if (at < T.maxLocal) {
                            T.small[at] = childTask(cs[side], cc[side], ci[side], metaOf(childBuf, level + 1), s, side);
                        } else {
                            atomicOr(&T.ctl[kCtlError], 1);
                        }
                    } else if (ssp < kSubtreeStack) {
                        stk[ssp++] = childTask(cs[side], cc[side], ci[side], metaOf(childBuf, level + 1), s, side);
                    } else {
                        atomicOr(&T.ctl[kCtlError], 1);
                    }
                }
            }
        }
        __syncthreads();
    }
    }
    if (threadIdx.x == 0) {
        flushSerial(T, serBuf, serN);
        atomicAdd(&T.ctl[kCtlBinned], binned);
        atomicAdd(&T.ctl[kCtlParted], parted);
        atomicMax(&T.ctl[kCtlMaxLevel], deepest);
    }
}

// Bottom tier: one small task (kSerial+1 .. kSmallTask primitives) per one-wave
// workgroup, grid-strided over the small list.  The task's records are read into LDS
// once; its subtree is then split entirely in LDS -- histograms read the records
// through a position -> record index, and each stable partition moves 16-bit indices
// between two LDS index buffers -- so the lower levels cost no HBM traffic.  A range
// that reaches kSerial or fewer primitives writes its leaf order (scene ids) at once
// and goes to the serial list marked kOrderWritten for its balanced splits.
constexpr int kOrderWritten = 2;  // SerialTask.buf: the range's leaf order is already in place
constexpr int kSmallBT = 128;     // threads per small task

template <int kBT>
__device__ __forceinline__ void histAddLds(uint32_t* h, const PrimRec* recL, const unsigned short* idx, int count,
                                           const AxisMap& m) {
    const int per = (count + kBT - 1) / kBT;
    const int beg = threadIdx.x * per;
    const int end = beg + per < count ? beg + per : count;
    BinRun run[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        run[a].bin = -1;
        run[a].cnt = 0u;
    }
    for (int i = beg; i < end; ++i) {
        const PrimRec r = recL[idx[i]];
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

template <int kBT>
__global__ __launch_bounds__(kBT) void sahSmallKernel(SahTables T, int n, uint32_t* __restrict__ order,
                                                      Node2* __restrict__ nodes, int* __restrict__ parent,
                                                      int* __restrict__ farEnd) {
    constexpr int kWaves = kBT / 64;
    __shared__ PrimRec recL[kSmallTask];
    __shared__ unsigned short idxL[2][kSmallTask];
    __shared__ uint32_t h[kHist];
    __shared__ SahResult sres;
    __shared__ SahTask stk[kSubtreeStack];
    __shared__ SerialTask serBuf[kSerialBatch];
    __shared__ int waveLefts[kWaves];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    int serN = 0;  // thread 0's count of buffered serial ranges
    int binned = 0, parted = 0, deepest = 0;
    int* err = &T.ctl[kCtlError];
    const int listLen = T.ctl[kCtlSmall];
    for (int task = blockIdx.x; task < listLen; task += gridDim.x) {
        const SahTask root = T.small[task];
        if (root.count > kSmallTask) {  // the list holds tasks of at most kSmallTask primitives
            if (threadIdx.x == 0) atomicOr(err, 4);
            continue;
        }
        const int base = root.start;
        const PrimRec* src = recOf(T, root.meta & 1) + base;
        for (int i = threadIdx.x; i < root.count; i += kBT) {
            recL[i] = src[i];
            idxL[0][i] = static_cast<unsigned short>(i);
        }
        int sp = 1;  // uniform: every thread keeps the stack depth
        if (threadIdx.x == 0)
            stk[0] = makeTask(root.start, root.count, root.node, metaOf(0, root.meta >> 1), root.cLo, root.cHi);
        __syncthreads();
        while (sp > 0) {
            --sp;
            const SahTask t = stk[sp];
            const int buf = t.meta & 1, level = t.meta >> 1;
            const bool median = level >= kMedianLevel;
            const int ls = t.start - base;  // the range's first position in LDS
            const unsigned short* idxS = idxL[buf] + ls;
            const AxisMap m = axisMap(t);
            if (!median) {
                histInit(h);
                __syncthreads();
                histAddLds<kBT>(h, recL, idxS, t.count, m);
                __syncthreads();
                if (threadIdx.x < 64) chooseSplit(h, t, m, &sres);
            } else if (threadIdx.x == 0) {
                sres = medianResult(t);
            }
            __syncthreads();
            const SahResult s = sres;
            const int nL = s.nL;
            if (nL < 1 || nL >= t.count) {
                if (threadIdx.x == 0) atomicOr(err, 4);
                __syncthreads();
                continue;
            }
            const int childBuf = median ? buf : 1 - buf;
            if (!median) {
                // stable partition of the range's indices into the other buffer: wave w
                // takes the w-th part of the range (lefts of the earlier parts counted
                // first when there are several waves)
                unsigned short* idxD = idxL[childBuf] + ls;
                const int axis = s.axis < 0 ? 0 : s.axis;
                const int per = (t.count + kWaves - 1) / kWaves;
                const int wBeg = w * per < t.count ? w * per : t.count;
                const int wEnd = wBeg + per < t.count ? wBeg + per : t.count;
                int leftBefore = 0;
                if constexpr (kWaves > 1) {
                    int lefts = 0;
                    for (int r = wBeg; r < wEnd; r += 64) {
                        const int i = r + lane;
                        const bool left = i < wEnd && goesLeft(recL[idxS[i]].c[axis], t.start + i, t, s, m);
                        lefts += __popcll(__ballot(left));
                    }
                    if (lane == 0) waveLefts[w] = lefts;
                    __syncthreads();
                    for (int v = 0; v < w; ++v) leftBefore += waveLefts[v];
                }
                int runL = leftBefore, runR = nL + (wBeg - leftBefore);
                bool bad = false;
                for (int r = wBeg; r < wEnd; r += 64) {
                    const int i = r + lane;
                    const bool valid = i < wEnd;
                    unsigned short id = 0;
                    bool left = false;
                    if (valid) {
                        id = idxS[i];
                        left = goesLeft(recL[id].c[axis], t.start + i, t, s, m);
                    }
                    const unsigned long long lm = __ballot(valid && left);
                    const unsigned long long rm = __ballot(valid && !left);
                    if (valid) {
                        const int d = left ? runL + __popcll(lm & below) : runR + __popcll(rm & below);
                        const bool ok = left ? (d >= 0 && d < nL) : (d >= nL && d < t.count);
                        if (ok) {
                            idxD[d] = id;
                        } else {
                            bad = true;
                        }
                    }
                    runL += __popcll(lm);
                    runR += __popcll(rm);
                }
                if (bad) atomicOr(err, 2);
                binned += t.count;
                parted += t.count;
                __syncthreads();
            }
            deepest = level > deepest ? level : deepest;
            // every thread computes the split's ids; thread 0 writes the node
            const int g = t.start + nL - 1, e = t.start + t.count - 1;
            const int l = (g == t.start) ? n - 1 + g : g;
            const int r = (g + 1 == e) ? n - 1 + e : g + 1;
            if (threadIdx.x == 0) {
                nodes[t.node].left = l;
                nodes[t.node].right = r;
                parent[l] = t.node;
                parent[r] = t.node;
                farEnd[t.node] = t.node == t.start ? e : t.start;
            }
            const int cs[2] = {t.start, t.start + nL}, cc[2] = {nL, t.count - nL}, ci[2] = {l, r};
#pragma unroll
            for (int side = 1; side >= 0; --side) {
                if (cc[side] <= kSerial) {
                    // the range's leaf order, one thread per position (at most kSerial)
                    if (threadIdx.x < cc[side])
                        order[cs[side] + threadIdx.x] =
                            static_cast<uint32_t>(recL[idxL[childBuf][cs[side] - base + threadIdx.x]].id);
                    if (threadIdx.x == 0) {
                        if (serN == kSerialBatch) {
                            flushSerial(T, serBuf, serN);
                            serN = 0;
                        }
                        serBuf[serN++] = {cs[side], cc[side], ci[side], kOrderWritten};
                    }
                } else if (sp < kSubtreeStack) {
                    if (threadIdx.x == 0)
                        stk[sp] = childTask(cs[side], cc[side], ci[side], metaOf(childBuf, level + 1), s, side);
                    ++sp;
                } else if (threadIdx.x == 0) {
                    atomicOr(err, 1);
                }
            }
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        flushSerial(T, serBuf, serN);
        atomicAdd(&T.ctl[kCtlBinned], binned);
        atomicAdd(&T.ctl[kCtlParted], parted);
        atomicMax(&T.ctl[kCtlMaxLevel], deepest);
    }
}

// One thread per final range of at most kSerial primitives (grid-stride over the
// serial list): its leaf order (the scene indices, read from the record buffer that
// holds the range) and, for two or more primitives, a balanced subtree of count/2
// splits.
__global__ void sahSerialKernel(SahTables T, int n, uint32_t* __restrict__ order, Node2* __restrict__ nodes,
                                int* __restrict__ parent, int* __restrict__ farEnd) {
    const int count = T.ctl[kCtlSerial] < T.maxSerial ? T.ctl[kCtlSerial] : T.maxSerial;
    int slots = 0;
    for (int ti = blockIdx.x * kBlock + threadIdx.x; ti < count; ti += kBlock * gridDim.x) {
        const SerialTask t = T.serial[ti];
        const PrimRec* src = recOf(T, t.buf);
        slots += t.count;
        if (t.buf != kOrderWritten)
            for (int i = 0; i < t.count; ++i) order[t.start + i] = static_cast<uint32_t>(src[t.start + i].id);
        if (t.count < 2) continue;
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
    if (slots) atomicAdd(&T.ctl[kCtlSerialSlots], slots);
}

// Pinned host words for the loop's control reads (per host thread, kept for the
// process): a read into pageable memory would go through a driver bounce buffer.
int* pinnedWords() {
    static thread_local int* w = nullptr;
    if (!w) hipCheck(hipHostMalloc(reinterpret_cast<void**>(&w), kCtlWords * sizeof(int), hipHostMallocDefault),
                     "sah control words");
    return w;
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

    DevBuf<PrimRec> rec[2];
    rec[0].alloc(static_cast<size_t>(n));
    rec[1].alloc(static_cast<size_t>(n));
    DevBuf<uint32_t> order(static_cast<size_t>(n));

    SahTables T{};
    T.maxBig = n / (kChunk + 1) + 2;
    T.maxLocal = n / (kSerial + 1) + 2;
    T.maxSerial = n + 2;
    T.maxChunks = n / kChunk + T.maxBig + 2;
    DevBuf<SahTask> dBig(2 * static_cast<size_t>(T.maxBig)), dLocal(static_cast<size_t>(T.maxLocal)),
        dSmall(static_cast<size_t>(T.maxLocal));
    DevBuf<SerialTask> dSerial(static_cast<size_t>(T.maxSerial));
    DevBuf<Chunk> dChunks(static_cast<size_t>(T.maxChunks));
    DevBuf<int> dFirst(static_cast<size_t>(T.maxBig) + 1), dChunkBins(static_cast<size_t>(T.maxChunks) * 3 * kBins),
        dChunkOff(static_cast<size_t>(T.maxChunks)), dCtl(kCtlWords);
    DevBuf<uint32_t> dBins(static_cast<size_t>(T.maxBig) * kHist);
    DevBuf<SahResult> dRes(static_cast<size_t>(T.maxBig));
    T.rec[0] = rec[0].p;
    T.rec[1] = rec[1].p;
    T.big[0] = dBig.p;
    T.big[1] = dBig.p + T.maxBig;
    T.local = dLocal.p;
    T.small = dSmall.p;
    T.serial = dSerial.p;
    T.chunks = dChunks.p;
    T.first = dFirst.p;
    T.chunkBins = dChunkBins.p;
    T.chunkOff = dChunkOff.p;
    T.bins = dBins.p;
    T.results = dRes.p;
    T.ctl = dCtl.p;
    int* hw = pinnedWords();

    EventSet<6> ev;
    ev.record(0, s);
    sahResetKernel<<<blocksFor(n - 1 > 256 ? n - 1 : 256, kBlock), kBlock, 0, s>>>(cb, b.arrive.p, n - 1);
    int grid = blocksFor(n, kBlock * 4);
    if (grid > 512) grid = 512;  // two workgroups per CU stream the items; few bounds atomics
    sahPrepKernel<L><<<grid, kBlock, 0, s>>>(dSrc, n, rec[0].p, cb, b.parent.p);
    sahRootKernel<<<1, 64, 0, s>>>(T, cb, n);
    hipCheck(hipGetLastError(), "sah prep");
    ev.record(1, s);

    // Large tasks, level by level, enqueued in batches with no host work between the
    // levels of a batch; after each batch one read of the next level's task count
    // says whether to go on.  The first batch covers the levels a balanced split
    // would need to bring every task under kChunk.
    std::vector<hipEvent_t> lev;
    auto mark = [&](hipStream_t st) {
        const hipEvent_t e = takeEvent();
        hipCheck(hipEventRecord(e, st), "event record");
        lev.push_back(e);
    };
    int level = 0;
    if (n > kChunk) {
        int batch = 1;
        while ((static_cast<long long>(kChunk) << batch) < n) ++batch;
        const int chunkGrid = T.maxChunks, taskGrid = blocksFor(T.maxBig, kBlock);
        const int clearGrid = std::min(blocksFor(static_cast<long long>(T.maxBig) * kHist, kBlock), 1024);
        while (true) {
            for (int j = 0; j < batch; ++j, ++level) {
                const int cur = level & 1;
                const int median = level >= kMedianLevel ? 1 : 0;
                mark(s);
                sahPlanKernel<<<1, kBlock, 0, s>>>(T, cur, median);
                sahClearBinsKernel<<<clearGrid, kBlock, 0, s>>>(T, cur);
                sahBinKernel<<<chunkGrid, kBlock, 0, s>>>(T, cur);
                mark(s);
                sahSplitKernel<<<T.maxBig, 64, 0, s>>>(T, cur, median);
                mark(s);
                sahChunkScanKernel<<<T.maxBig, 64, 0, s>>>(T, cur);
                sahPartScatterKernel<<<chunkGrid, kBlock, 0, s>>>(T, cur);
                sahAssignKernel<<<taskGrid, kBlock, 0, s>>>(T, cur, median, n, b.nodes.p, b.parent.p, b.farEnd.p);
                mark(s);
            }
            hipCheck(hipGetLastError(), "sah level");
            hipCheck(hipMemcpyAsync(hw, T.ctl, kCtlWords * sizeof(int), hipMemcpyDeviceToHost, s), "sah control");
            hipCheck(hipStreamSynchronize(s), "sah level sync");
            if (hw[kCtlError] != 0) break;
            if (hw[kCtlBig0 + (level & 1)] == 0) break;
            if (level > 2 * kMedianLevel + 64) throw DeviceError("sah: level limit exceeded");
            batch = 2;
        }
    } else {
        hipCheck(hipMemcpyAsync(hw, T.ctl, kCtlWords * sizeof(int), hipMemcpyDeviceToHost, s), "sah control");
        hipCheck(hipStreamSynchronize(s), "sah sync");
    }
    auto failed = [&]() {
        for (hipEvent_t e : lev) giveEvent(e);
        if (hw[kCtlError] & 1) throw DeviceError("sah: task table overflow");
        if (hw[kCtlError] & 2) throw DeviceError("sah: partition counts disagree with the bin histogram");
        throw DeviceError("sah: empty side in a split");
    };
    if (hw[kCtlError] != 0) failed();
    ev.record(2, s);
    // local subtrees (the host knows their number from the last read), then the
    // serial tails; the topology is checked complete before the bottom-up pass
    // climbs it
    if (hw[kCtlLocal] > 0)
        sahSubtreeKernel<kBlock><<<std::min(hw[kCtlLocal], T.maxLocal), kBlock, 0, s>>>(
            T, n, b.nodes.p, b.parent.p, b.farEnd.p);
    if (hw[kCtlLocal] > 0 || hw[kCtlSmall] > 0) {
        // the top tier adds small tasks: their count is known only on the device, so a
        // fixed grid strides the list
        const int smallGrid = std::min(T.maxLocal, 8192);
        sahSmallKernel<kSmallBT><<<smallGrid, kSmallBT, 0, s>>>(T, n, order.p, b.nodes.p, b.parent.p, b.farEnd.p);
    }
    ev.record(3, s);
    sahSerialKernel<<<std::min(blocksFor(T.maxSerial, kBlock), 1024), kBlock, 0, s>>>(T, n, order.p, b.nodes.p,
                                                                                     b.parent.p, b.farEnd.p);
    hipCheck(hipGetLastError(), "sah subtrees");
    hipCheck(hipMemcpyAsync(hw, T.ctl, kCtlWords * sizeof(int), hipMemcpyDeviceToHost, s), "sah control");
    ev.record(4, s);
    ev.sync(4);
    if (hw[kCtlError] != 0) failed();

    finishTopology(b, dSrc, order.p, s);
    hipCheck(hipMemcpyAsync(hw, T.ctl, kCtlWords * sizeof(int), hipMemcpyDeviceToHost, s), "sah control");
    int depth = 0;
    hipCheck(hipMemcpyAsync(&depth, b.height.p, sizeof(int), hipMemcpyDeviceToHost, s), "depth");
    ev.record(5, s);
    ev.sync(5);
    b.depth = depth;
    SahBuildStats st;
    st.levels = hw[kCtlMaxLevel] + 1;
    st.binnedSlots = static_cast<unsigned int>(hw[kCtlBinned]);
    st.partitionedSlots = static_cast<unsigned int>(hw[kCtlParted]);
    st.serialSlots = static_cast<unsigned int>(hw[kCtlSerialSlots]);
    tSahStats = st;
    if (t) {
        double binMs = 0.0, splitMs = 0.0, partMs = 0.0;
        for (size_t i = 0; i + 3 < lev.size(); i += 4) {
            float v = 0.0f;
            hipCheck(hipEventElapsedTime(&v, lev[i], lev[i + 1]), "event elapsed");
            binMs += v;
            hipCheck(hipEventElapsedTime(&v, lev[i + 1], lev[i + 2]), "event elapsed");
            splitMs += v;
            hipCheck(hipEventElapsedTime(&v, lev[i + 2], lev[i + 3]), "event elapsed");
            partMs += v;
        }
        t->mortonMs = ev.ms(0, 1);
        t->sortMs = 0.0;
        t->treeMs = ev.ms(1, 4);
        t->scatterMs = ev.ms(2, 4);  // local subtrees and serial tails
        t->refitMs = ev.ms(4, 5);
        t->sahBinMs = binMs;
        t->sahSplitMs = splitMs;
        t->sahPartMs = partMs;
        t->totalMs = ev.ms(0, 5);
    }
    for (hipEvent_t e : lev) giveEvent(e);
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
