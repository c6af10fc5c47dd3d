// Binned SAH builder for gfx950: top-down, level-synchronous, 32 bins per axis.
//
// Reference: src/gpu/lbvh.cu:827-1291 (binned SAH kernels), :1777-1952 (level
// driver), :2043-2122 (entry points).  The reference bins 16 ways with one global
// bin table per task; this build bins 32 ways and keeps the histograms in LDS:
//   * tasks above kChunk primitives are cut into kChunk-position chunks; each
//     workgroup builds its chunk's histogram (3 axes x 32 bins x {count, box,
//     centroid box}) with LDS atomics and flushes it once with global atomics;
//     one wave per task then scores the 3 x 31 planes, and a one-pass stable
//     partition moves the primitive records (the bin pass also counts each
//     quarter-chunk part's bins, a per-task scan of the parts' left counts gives
//     every partition wave its destinations);
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
#include <cstdint>
#include <type_traits>
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
constexpr int kParts = 4;                   // a chunk's parts: one partition wave each
constexpr int kPartLen = kChunk / kParts;   // records per part

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
// Histograms that keep no centroid bounds (kCent = false: the subtree tiers, which
// take their children's centroid boxes from the records) store 7 values per bin
// (count, box lo xyz, box hi xyz) in a smaller LDS array.
template <bool kCent>
constexpr int binVals() { return kCent ? kBinVals : 7; }
template <bool kCent>
constexpr int histWords() { return 3 * kBins * binVals<kCent>(); }
template <bool kCent>
__device__ __forceinline__ int hidxT(int a, int b, int v) { return (a * kBins + b) * binVals<kCent>() + v; }

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

// The empty value of histogram word i: count 0, box and centroid bounds inverted.
__device__ __forceinline__ uint32_t binInit(int i) {
    const int v = i % kBinVals;
    return v == 0 ? 0u : ((v >= 1 && v <= 3) || (v >= 7 && v <= 9)) ? orderedBits(kInf) : orderedBits(-kInf);
}

template <bool kCent = true>
__device__ __forceinline__ void histInit(uint32_t* h) {
    // (word v of a bin: the same empty values for both layouts, v < 7 being shared)
    for (int i = threadIdx.x; i < histWords<kCent>(); i += blockDim.x) h[i] = binInit(i % binVals<kCent>());
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

// Adds a finished run to the histogram (and its count to part[a * kBins + bin] when
// part is given: the counts of the current part of a chunk).
// kCent = false leaves the centroid bounds out (a caller that takes its children's
// centroid boxes from the records themselves).
template <bool kCent = true>
__device__ __forceinline__ void binRunFlush(uint32_t* h, int a, const BinRun& r, uint32_t* part = nullptr) {
    if (r.cnt == 0u) return;
    atomicAdd(&h[hidxT<kCent>(a, r.bin, 0)], r.cnt);
    if (part) atomicAdd(&part[a * kBins + r.bin], r.cnt);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        atomicMin(&h[hidxT<kCent>(a, r.bin, 1 + k)], r.v[k]);
        atomicMax(&h[hidxT<kCent>(a, r.bin, 4 + k)], r.v[3 + k]);
        if constexpr (kCent) {
            atomicMin(&h[hidxT<kCent>(a, r.bin, 7 + k)], r.v[6 + k]);
            atomicMax(&h[hidxT<kCent>(a, r.bin, 10 + k)], r.v[9 + k]);
        }
    }
}

__device__ __forceinline__ void binRunsInit(BinRun* run) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        run[a].bin = -1;
        run[a].cnt = 0u;
    }
}

// One record into the thread's runs (a run is flushed when its bin changes); one
// instantiation per axis, so that the runs and the axis map stay in registers.
template <int a, bool kCent, typename R>
__device__ __forceinline__ void binRunAddAxis(uint32_t* h, BinRun& u, const R& r, const AxisMap& m,
                                              uint32_t* part) {
    if (!m.use[a]) return;
    const int b = binOf(r.c[a], m.lo[a], m.scale[a]);
    if (b != u.bin) {
        binRunFlush<kCent>(h, a, u, part);
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

// r is taken by value: loaded once, not re-read after the flushes' LDS atomics.
template <bool kCent = true, typename R = PrimRec>
__device__ __forceinline__ void binRunAdd(uint32_t* h, BinRun* run, const R r, const AxisMap& m,
                                          uint32_t* part = nullptr) {
    binRunAddAxis<0, kCent>(h, run[0], r, m, part);
    binRunAddAxis<1, kCent>(h, run[1], r, m, part);
    binRunAddAxis<2, kCent>(h, run[2], r, m, part);
}

template <bool kCent = true>
__device__ __forceinline__ void binRunsFlush(uint32_t* h, const BinRun* run, const AxisMap& m,
                                             uint32_t* part = nullptr) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (m.use[a]) binRunFlush<kCent>(h, a, run[a], part);
}

// Bins a contiguous range of records: each of the kBT threads takes a run of
// consecutive records (neighbouring records mostly share a bin, so the runs flush
// rarely).
template <int kBT, bool kCent = true>
__device__ __forceinline__ void histAddRange(uint32_t* h, const PrimRec* __restrict__ rec, int count,
                                             const AxisMap& m) {
    const int per = (count + kBT - 1) / kBT;
    const int beg = threadIdx.x * per;
    const int end = beg + per < count ? beg + per : count;
    BinRun run[3];
    binRunsInit(run);
    if (beg < end) {  // the next record is requested before the current one is binned
        PrimRec cur = rec[beg];
        for (int i = beg; i < end; ++i) {
            const PrimRec nxt = rec[i + 1 < end ? i + 1 : i];
            binRunAdd<kCent>(h, run, cur, m);
            cur = nxt;
        }
    }
    binRunsFlush<kCent>(h, run, m);
}

__device__ __forceinline__ float areaOf(float lx, float ly, float lz, float hx, float hy, float hz) {
    const float dx = hx - lx, dy = hy - ly, dz = hz - lz;
    return dx * dy + dy * dz + dz * dx;
}

// Plane choice by one wave (called by all 64 lanes): lane b holds bin b (lanes
// 32..63 repeat lanes 0..31); inclusive prefix and suffix unions of the bins' counts
// and boxes give every plane p = 1..31 its left side (bins 0..p-1, the prefix one
// lane down) and right side (bins p..31) at lane p, which scores it on each usable
// axis.  Counts and min / max are exact and order-free, so the costs are the same
// bits as summing the bins plane by plane.  The (cost, axis * 32 + plane) minimum
// is reduced across the lanes, then the chosen plane's side counts and centroid
// boxes; lane 0 writes the result.  h is the task's histogram (LDS or global).
// kCent = false: a histogram without centroid bounds; the chosen plane's side
// centroid boxes are then left 0 for the caller to fill from the records.
template <bool kCent = true>
__device__ void chooseSplit(const uint32_t* h, const SahTask& t, const AxisMap& m, SahResult* out) {
    const int lane = threadIdx.x & 63;
    const int bin = lane & (kBins - 1);
    const int count = t.count;
    const int minSide = count / 8 > 1 ? count / 8 : 1;
    float best = kInf;
    int bestKey = 0x7fffffff;  // axis * 32 + plane
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (!m.use[a]) continue;
        const int c = static_cast<int>(h[hidxT<kCent>(a, bin, 0)]);
        float v[6];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            v[k] = fromOrderedBits(h[hidxT<kCent>(a, bin, 1 + k)]);
            v[3 + k] = fromOrderedBits(h[hidxT<kCent>(a, bin, 4 + k)]);
        }
        const int pc = prefixAdd32(c, lane);
        const int total = __builtin_amdgcn_readlane(pc, 31);
        const int nL = dppI<0x138>(pc, 0);  // the prefix through bin - 1 (wave_shr:1)
        const int nR = total - nL;
        float l[6], r[6];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            l[k] = dppF<0x138>(prefix32<MinF>(v[k], lane), kInf);
            l[3 + k] = dppF<0x138>(prefix32<MaxF>(v[3 + k], lane), -kInf);
            r[k] = suffix32<MinF>(v[k], lane);
            r[3 + k] = suffix32<MaxF>(v[3 + k], lane);
        }
        if (bin >= 1 && nL >= minSide && nR >= minSide) {
            const float cost = areaOf(l[0], l[1], l[2], l[3], l[4], l[5]) * static_cast<float>(nL) +
                               areaOf(r[0], r[1], r[2], r[3], r[4], r[5]) * static_cast<float>(nR);
            const int key = a * kBins + bin;
            if (cost < best || (cost == best && key < bestKey)) {
                best = cost;
                bestKey = key;
            }
        }
    }
    {
        // (cost, key) minimum over the half, then over the two rows' results
        auto take = [&](float oc, int ok) {
            if (oc < best || (oc == best && ok < bestKey)) {
                best = oc;
                bestKey = ok;
            }
        };
        take(dppF<0xB1>(best, kInf), dppI<0xB1>(bestKey, 0x7fffffff));
        take(dppF<0x4E>(best, kInf), dppI<0x4E>(bestKey, 0x7fffffff));
        take(dppF<0x141>(best, kInf), dppI<0x141>(bestKey, 0x7fffffff));
        take(dppF<0x140>(best, kInf), dppI<0x140>(bestKey, 0x7fffffff));
        const float c0 = readlaneF(best, 0), c1 = readlaneF(best, 16);
        const int k0 = __builtin_amdgcn_readlane(bestKey, 0), k1 = __builtin_amdgcn_readlane(bestKey, 16);
        best = c0;
        bestKey = k0;
        take(c1, k1);
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
    // right of it (a min / max of ordered bits is exact and order-free; empty bins
    // take the identities)
    const int a = bestKey / kBins, plane = bestKey % kBins;
    const uint32_t c = h[hidxT<kCent>(a, bin, 0)];
    const bool isLeft = bin < plane;
    const uint32_t nL = reduce32<AddU>(isLeft ? c : 0u);
    uint32_t v[12];  // left min xyz, left max xyz, right min xyz, right max xyz
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = 0u;
    if constexpr (kCent) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint32_t lo = c ? h[hidxT<kCent>(a, bin, 7 + k)] : 0xFFFFFFFFu;
            const uint32_t hi = c ? h[hidxT<kCent>(a, bin, 10 + k)] : 0u;
            v[k] = reduce32<MinU>(isLeft ? lo : 0xFFFFFFFFu);
            v[3 + k] = reduce32<MaxU>(isLeft ? hi : 0u);
            v[6 + k] = reduce32<MinU>(isLeft ? 0xFFFFFFFFu : lo);
            v[9 + k] = reduce32<MaxU>(isLeft ? 0u : hi);
        }
    }
    if (lane != 0) return;
    res.axis = a;
    res.plane = plane;
    res.nL = static_cast<int>(nL);
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

// Element `axis` of a three-element array held in registers, by selection (a dynamic
// subscript into a register copy would go through scratch).
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
template <int kBT = kBlock, bool kBoxes = false>
__device__ __forceinline__ void partitionRange(const PrimRec* __restrict__ src, PrimRec* __restrict__ dst, int begin, int cnt,
                               int leftDst, int rightDst, const SahTask& t, const SahResult& s, const AxisMap& m,
                               int* ldsWave, int* err, uint32_t* waveBox = nullptr) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ax = s.axis < 0 ? 0 : s.axis;  // the split axis (any for a split by position)
    constexpr int kWaves = kBT / 64;
    const int per = (cnt + kWaves - 1) / kWaves;
    const int wBeg = w * per < cnt ? w * per : cnt;
    const int wEnd = wBeg + per < cnt ? wBeg + per : cnt;
    const unsigned long long below = (1ull << lane) - 1ull;
    int lefts = 0;
    for (int r = wBeg; r < wEnd; r += 64) {
        const int i = r + lane;
        const bool left = i < wEnd && goesLeft(src[begin + i].c[ax], begin + i, t, s, m);
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
    // kBoxes: the sides' centroid boxes as ordered bits (left min xyz, max xyz, right
    // min xyz, max xyz), reduced per wave into waveBox[w][12]
    uint32_t bx[12];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        bx[k] = bx[6 + k] = 0xFFFFFFFFu;
        bx[3 + k] = bx[9 + k] = 0u;
    }
    for (int r = wBeg; r < wEnd; r += 64) {
        const int i = r + lane;
        const bool valid = i < wEnd;
        // the record as five 8-byte words (plain values: a struct copy here would be
        // kept in scratch once the split coordinate is read out of it by index)
        uint2 w0{}, w1{}, w2{}, w3{}, w4{};
        bool left = false;
        if (valid) {
            const uint2* sw = reinterpret_cast<const uint2*>(src + begin + i);
            w0 = sw[0];
            w1 = sw[1];
            w2 = sw[2];
            w3 = sw[3];
            w4 = sw[4];
            const float c = ax == 0 ? __uint_as_float(w0.x) : ax == 1 ? __uint_as_float(w0.y) : __uint_as_float(w1.x);
            left = goesLeft(c, begin + i, t, s, m);
            if (kBoxes) {
                const uint32_t u[3] = {orderedBits(__uint_as_float(w0.x)), orderedBits(__uint_as_float(w0.y)),
                                       orderedBits(__uint_as_float(w1.x))};
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const int o = left ? 0 : 6;
                    const uint32_t lo = left ? bx[k] : bx[6 + k], hi = left ? bx[3 + k] : bx[9 + k];
                    const uint32_t nlo = u[k] < lo ? u[k] : lo, nhi = u[k] > hi ? u[k] : hi;
                    bx[k] = o == 0 ? nlo : bx[k];
                    bx[6 + k] = o == 6 ? nlo : bx[6 + k];
                    bx[3 + k] = o == 0 ? nhi : bx[3 + k];
                    bx[9 + k] = o == 6 ? nhi : bx[9 + k];
                }
            }
        }
        const unsigned long long lm = __ballot(valid && left);
        const unsigned long long rm = __ballot(valid && !left);
        if (valid) {
            const int d = left ? runL + __popcll(lm & below) : runR + __popcll(rm & below);
            const bool ok = left ? (d >= t.start && d < leftEnd) : (d >= leftEnd && d < rightEnd);
            if (ok) {
                uint2* dw = reinterpret_cast<uint2*>(dst + d);
                dw[0] = w0;
                dw[1] = w1;
                dw[2] = w2;
                dw[3] = w3;
                dw[4] = w4;
            } else {
                bad = true;
            }
        }
        runL += __popcll(lm);
        runR += __popcll(rm);
    }
    if (bad) atomicOr(err, 2);
    if (kBoxes) {
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            const bool isMin = k < 3 || (k >= 6 && k < 9);
            bx[k] = isMin ? reduce64<MinU>(bx[k]) : reduce64<MaxU>(bx[k]);
        }
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 12; ++k) waveBox[w * 12 + k] = bx[k];
        }
    }
}

// Stable partition of the records [begin, begin + cnt) by one wave whose left and
// right destinations are known (from the counts of the range's bins): one pass,
// lane-strided reads, ballot ranks.  Destinations outside the task's sides are
// reported through err, never written.
__device__ __forceinline__ void partitionWaveRange(const PrimRec* __restrict__ src, PrimRec* __restrict__ dst,
                                                   int begin, int cnt, int leftDst, int rightDst, const SahTask& t,
                                                   const SahResult& s, const AxisMap& m, int* err) {
    const int lane = threadIdx.x & 63;
    const int ax = s.axis < 0 ? 0 : s.axis;
    const unsigned long long below = (1ull << lane) - 1ull;
    const int leftEnd = t.start + s.nL, rightEnd = t.start + t.count;
    int runL = leftDst, runR = rightDst;
    bool bad = false;
    for (int r = 0; r < cnt; r += 64) {
        const int i = r + lane;
        const bool valid = i < cnt;
        uint2 w0{}, w1{}, w2{}, w3{}, w4{};
        bool left = false;
        if (valid) {
            const uint2* sw = reinterpret_cast<const uint2*>(src + begin + i);
            w0 = sw[0];
            w1 = sw[1];
            w2 = sw[2];
            w3 = sw[3];
            w4 = sw[4];
            const float c = ax == 0 ? __uint_as_float(w0.x) : ax == 1 ? __uint_as_float(w0.y) : __uint_as_float(w1.x);
            left = goesLeft(c, begin + i, t, s, m);
        }
        const unsigned long long lm = __ballot(valid && left);
        const unsigned long long rm = __ballot(valid && !left);
        if (valid) {
            const int d = left ? runL + __popcll(lm & below) : runR + __popcll(rm & below);
            const bool ok = left ? (d >= t.start && d < leftEnd) : (d >= leftEnd && d < rightEnd);
            if (ok) {
                uint2* dw = reinterpret_cast<uint2*>(dst + d);
                dw[0] = w0;
                dw[1] = w1;
                dw[2] = w2;
                dw[3] = w3;
                dw[4] = w4;
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
    int* chunkBins;  // per chunk part: its primitives' counts per axis and bin (3 x kBins)
    int* chunkOff;   // per chunk part: lefts of the task's earlier parts
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
// Build set-up: centroid bounds to empty, arrival counters to zero, and (given the
// loop's tables) the control words to zero and both histogram buffers to empty.
__global__ void sahResetKernel(uint32_t* __restrict__ cb, int* __restrict__ arrive, int internal, SahTables T,
                               int clearBins) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int stride = kBlock * gridDim.x;
    if (i == 8) cb[8] = 0u;  // partition error counter
    if (i < 3) cb[i] = orderedBits(kInf);
    if (i >= 3 && i < 6) cb[i] = orderedBits(-kInf);
    if (i < kCtlWords) T.ctl[i] = 0;
    for (int j = i; j < internal; j += stride) arrive[j] = 0;
    if (clearBins) {
        const long long words = 2LL * T.maxBig * kHist;
        for (long long j = i; j < words; j += stride) T.bins[j] = binInit(static_cast<int>(j % kHist));
    }
}

// The root task, from the scene's centroid bounds (the control words are zero).
__device__ __forceinline__ void routeRoot(const SahTables& T, const uint32_t* __restrict__ cb, int n) {
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = fromOrderedBits(cb[a]);
        hi[a] = fromOrderedBits(cb[3 + a]);
    }
    routeChild(T, 0, 0, n, 0, 0, 0, lo, hi);
}

__global__ void sahRootKernel(SahTables T, const uint32_t* __restrict__ cb, int n) {
    if (threadIdx.x == 0) routeRoot(T, cb, n);
}

// Level plan (one workgroup): the current big tasks' chunk list (chunks of one task
// consecutive, first[task] .. first[task + 1]); clears the count of the buffer this
// level's children go to.  Per group of kBlock tasks, an exclusive scan of their chunk
// counts, then every thread writes chunks, each finding its task by binary search of
// the group's chunk offsets in LDS.
__global__ __launch_bounds__(kBlock) void sahPlanKernel(SahTables T, int cur, int median, const uint32_t* __restrict__ cb,
                                                        int rootN) {
    __shared__ int lds[4];
    __shared__ int sEx[kBlock], sStart[kBlock], sCount[kBlock];
    if (rootN > 0) {  // the first level routes the root task first
        if (threadIdx.x == 0) routeRoot(T, cb, rootN);
        __syncthreads();
    }
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

// The histogram buffer of a level's tasks (two buffers, by level parity).
__device__ __forceinline__ uint32_t* binsOf(const SahTables& T, int cur) {
    return T.bins + (cur ? static_cast<size_t>(T.maxBig) * kHist : size_t(0));
}

// One workgroup per chunk of a large task: LDS histogram, then one global flush.
// Wave w bins the chunk's part w (kPartLen records); besides the task's histogram it
// records each part's bin counts, from which the chunk scan gives every part its
// partition offsets without a counting pass.
__global__ __launch_bounds__(kBlock) void sahBinKernel(SahTables T, int cur) {
    static_assert(kChunk == kParts * kPartLen, "a chunk is its parts");
    static_assert(kBlock / 64 == kParts, "one wave per part");
    if (static_cast<int>(blockIdx.x) >= T.ctl[kCtlChunks]) return;
    __shared__ uint32_t h[kHist];
    __shared__ uint32_t part[kParts][3 * kBins];
    const Chunk ch = T.chunks[blockIdx.x];
    const SahTask t = bigOf(T, cur)[ch.task];
    const AxisMap m = axisMap(t);
    histInit(h);
    for (int i = threadIdx.x; i < kParts * 3 * kBins; i += kBlock) (&part[0][0])[i] = 0u;
    __syncthreads();
    // wave w bins part w, each lane a run of consecutive records
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int pBeg = w * kPartLen;
    const int pCnt = ch.count - pBeg < kPartLen ? ch.count - pBeg : kPartLen;
    if (pCnt > 0) {
        const PrimRec* src = recOf(T, t.meta & 1) + ch.start + pBeg;
        const int per = (pCnt + 63) / 64;
        const int beg = lane * per;
        const int end = beg + per < pCnt ? beg + per : pCnt;
        BinRun run[3];
        binRunsInit(run);
        if (beg < end) {  // the next record is requested before the current one is binned
            PrimRec cur = src[beg];
            for (int i = beg; i < end; ++i) {
                const PrimRec nxt = src[i + 1 < end ? i + 1 : i];
                binRunAdd(h, run, cur, m, part[w]);
                cur = nxt;
            }
        }
        binRunsFlush(h, run, m, part[w]);
    }
    __syncthreads();
    uint32_t* g = binsOf(T, cur) + static_cast<size_t>(ch.task) * kHist;
    // the parts' bin counts: each part's left count for any plane, without a re-read
    for (int i = threadIdx.x; i < kParts * 3 * kBins; i += kBlock)
        T.chunkBins[static_cast<size_t>(blockIdx.x) * kParts * 3 * kBins + i] = static_cast<int>((&part[0][0])[i]);
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


// One workgroup per large task: the plane choice (wave 0), then the exclusive
// offsets of the task's chunk parts' left counts (part e = kParts * chunk + p, from
// the parts' bin counts -- the bins left of the chosen plane on its axis, read as
// eight 16-byte words -- or, for a split by position, from the part's place in the
// task; kBlock parts per step, a block scan), the task's node and its children
// routed by size.  The task's histogram is left empty for the level two on.
__global__ __launch_bounds__(kBlock) void sahSplitKernel(SahTables T, int cur, int median, int n,
                                                         Node2* __restrict__ nodes, int* __restrict__ parent,
                                                         int* __restrict__ farEnd) {
    __shared__ int lds[4];
    __shared__ SahResult sres;
    const int ti = blockIdx.x;
    if (ti >= T.ctl[kCtlBig0 + cur]) return;
    const SahTask t = bigOf(T, cur)[ti];
    uint32_t* h = binsOf(T, cur) + static_cast<size_t>(ti) * kHist;
    if (median) {
        if (threadIdx.x == 0) sres = medianResult(t);
    } else if (threadIdx.x < 64) {
        chooseSplit(h, t, axisMap(t), &sres);
    }
    __syncthreads();
    const SahResult r = sres;
    if (threadIdx.x == 0) T.results[ti] = r;
    if (!median)
        for (int i = threadIdx.x; i < kHist; i += kBlock) h[i] = binInit(i);

    if (T.ctl[kCtlChunks] != 0) {
        const int e0 = T.first[ti] * kParts, e1 = T.first[ti + 1] * kParts;
        int acc = 0;
        for (int base = e0; base < e1; base += kBlock) {
            const int e = base + threadIdx.x;
            int left = 0;
            if (e < e1) {
                if (r.axis < 0) {
                    const Chunk ch = T.chunks[e / kParts];
                    const int pBeg = (e % kParts) * kPartLen;
                    int pCnt = ch.count - pBeg;
                    pCnt = pCnt < 0 ? 0 : (pCnt > kPartLen ? kPartLen : pCnt);
                    const int before = t.start + r.nL - (ch.start + pBeg);
                    left = before < 0 ? 0 : (before > pCnt ? pCnt : before);
                } else {
                    const int4* cnt = reinterpret_cast<const int4*>(T.chunkBins + static_cast<size_t>(e) * 3 * kBins +
                                                                    r.axis * kBins);
#pragma unroll
                    for (int q = 0; q < kBins / 4; ++q) {
                        const int4 v = cnt[q];
                        left += (4 * q < r.plane ? v.x : 0) + (4 * q + 1 < r.plane ? v.y : 0) +
                                (4 * q + 2 < r.plane ? v.z : 0) + (4 * q + 3 < r.plane ? v.w : 0);
                    }
                }
            }
            int total = 0;
            const int ex = blockExclusive(left, lds, &total);
            if (e < e1) T.chunkOff[e] = acc + ex;
            acc += total;
        }
        // the parts disagree with the task's histogram
        if (threadIdx.x == 0 && acc != r.nL) atomicOr(&T.ctl[kCtlError], 2);
    }

    if (threadIdx.x == 0) {
        const int nL = r.nL;
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
}

// One workgroup per chunk, one wave per part: each part's records go to their
// sides at the offsets the chunk scan gave it, in one pass.
__global__ __launch_bounds__(kBlock) void sahPartScatterKernel(SahTables T, int cur) {
    if (static_cast<int>(blockIdx.x) >= T.ctl[kCtlChunks]) return;
    const Chunk ch = T.chunks[blockIdx.x];
    const SahTask t = bigOf(T, cur)[ch.task];
    const SahResult s = T.results[ch.task];
    const AxisMap m = axisMap(t);
    const int w = threadIdx.x >> 6;
    const int pBeg = w * kPartLen;
    const int pCnt = ch.count - pBeg < kPartLen ? ch.count - pBeg : kPartLen;
    if (pCnt > 0) {
        const int e = blockIdx.x * kParts + w;
        const int partLeft = T.chunkOff[e];  // lefts of the task's earlier parts
        const int partPos = ch.start + pBeg - t.start;
        partitionWaveRange(recOf(T, t.meta & 1), recOf(T, 1 - (t.meta & 1)), ch.start + pBeg, pCnt,
                           t.start + partLeft, t.start + s.nL + (partPos - partLeft), t, s, m, &T.ctl[kCtlError]);
    }
    if (threadIdx.x == 0) atomicAdd(&T.ctl[kCtlParted], ch.count);
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
    __shared__ uint32_t h[histWords<false>()];
    __shared__ SahResult sres;
    __shared__ SahTask stk[kSubtreeStack];
    __shared__ SahTask curTask;
    __shared__ int ssp;
    __shared__ int lds[4];
    __shared__ SerialTask serBuf[kSerialBatch];
    __shared__ uint32_t waveBox[kBT / 64][12];
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
            histInit<false>(h);
            __syncthreads();
            histAddRange<kBT, false>(h, src + t.start, t.count, m);
            __syncthreads();
            if (threadIdx.x < 64) chooseSplit<false>(h, t, m, &sres);
        } else if (threadIdx.x == 0) {
            sres = medianResult(t);
        }
        __syncthreads();
        SahResult s = sres;
        if (!median) {
            // the histogram kept no centroid bounds: a split by plane takes its
            // children's centroid boxes from the records the partition moves (the same
            // min / max the bins would give)
            partitionRange<kBT, true>(src, dst, t.start, t.count, t.start, t.start + s.nL, t, s, m, lds,
                                      &T.ctl[kCtlError], &waveBox[0][0]);
            __syncthreads();
            if (s.axis >= 0) {
                uint32_t bx[12];
#pragma unroll
                for (int k = 0; k < 12; ++k) {
                    const bool isMin = k < 3 || (k >= 6 && k < 9);
                    uint32_t x = waveBox[0][k];
                    for (int v = 1; v < kBT / 64; ++v) x = isMin ? MinU::op(x, waveBox[v][k]) : MaxU::op(x, waveBox[v][k]);
                    bx[k] = x;
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    s.lLo[k] = fromOrderedBits(bx[k]);
                    s.lHi[k] = fromOrderedBits(bx[3 + k]);
                    s.rLo[k] = fromOrderedBits(bx[6 + k]);
                    s.rHi[k] = fromOrderedBits(bx[9 + k]);
                }
            }
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

// Bottom tier: one small task (kSerial+1 .. kSmallTask primitives) per workgroup of
// kSmallBT threads, grid-strided over the small list.  The task's records are read
// into LDS once; its subtree is then split entirely in LDS -- histograms read the
// records through a position -> record index, and each stable partition moves
// 16-bit indices between two LDS index buffers -- so the lower levels cost no HBM
// traffic.  The histograms keep no centroid bounds: after a split by plane the
// children's centroid boxes are reduced from their records (the same min / max the
// bins would give).  A range that reaches kSerial or fewer primitives writes its
// leaf order (scene ids) at once and goes to the serial list marked kOrderWritten
// for its balanced splits.
constexpr int kOrderWritten = 2;  // SerialTask.buf: the range's leaf order is already in place
constexpr int kSmallBT = 128;     // threads per small task
constexpr int kSmallStack = 24;   // a chain of count/8 sides from 512 to 64 is 16 deep
constexpr int kSmallSerialBatch = 16;

// A record as the small tier keeps it in LDS: centroid and box (36 B); the scene id
// is read from the record buffer when the task's leaf order is written at its end.
// With the 7-value histogram and the shorter stack and serial batch, a workgroup
// takes 25 KB of LDS, so six fit a CU instead of five.
struct PrimBox {
    float c[3];
    float lo[3];
    float hi[3];
};

template <int kBT>
__device__ __forceinline__ void histAddLds(uint32_t* h, const PrimBox* recL, const unsigned short* idx, int count,
                                           const AxisMap& m) {
    const int per = (count + kBT - 1) / kBT;
    const int beg = threadIdx.x * per;
    const int end = beg + per < count ? beg + per : count;
    BinRun run[3];
    binRunsInit(run);
    for (int i = beg; i < end; ++i) binRunAdd<false>(h, run, recL[idx[i]], m);
    binRunsFlush<false>(h, run, m);
}

template <int kBT>
__global__ __launch_bounds__(kBT) void sahSmallKernel(SahTables T, int n, uint32_t* __restrict__ order,
                                                      Node2* __restrict__ nodes, int* __restrict__ parent,
                                                      int* __restrict__ farEnd) {
    constexpr int kWaves = kBT / 64;
    __shared__ PrimBox recL[kSmallTask];
    __shared__ unsigned char finalBuf[kSmallTask];  // index buffer holding the final serial ranges' positions
    __shared__ unsigned short idxL[2][kSmallTask];
    __shared__ uint32_t h[histWords<false>()];
    __shared__ SahResult sres;
    __shared__ SahTask stk[kSmallStack];
    __shared__ SerialTask serBuf[kSmallSerialBatch];
    __shared__ int waveLefts[kWaves];
    __shared__ uint32_t waveBox[kWaves][12];
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
            const PrimRec q = src[i];
            recL[i] = {{q.c[0], q.c[1], q.c[2]}, {q.lo[0], q.lo[1], q.lo[2]}, {q.hi[0], q.hi[1], q.hi[2]}};
            idxL[0][i] = static_cast<unsigned short>(i);
            finalBuf[i] = 0xFF;
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
                histInit<false>(h);
                __syncthreads();
                histAddLds<kBT>(h, recL, idxS, t.count, m);
                __syncthreads();
                if (threadIdx.x < 64) chooseSplit<false>(h, t, m, &sres);
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
            SahResult sc = s;  // s with the children's centroid boxes
            if (!median && s.axis >= 0) {
                // left and right centroid boxes as ordered bits: min xyz, max xyz each
                uint32_t bx[12];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    bx[k] = bx[6 + k] = 0xFFFFFFFFu;
                    bx[3 + k] = bx[9 + k] = 0u;
                }
                const unsigned short* idxD = idxL[childBuf] + ls;
                for (int j = threadIdx.x; j < t.count; j += kBT) {
                    const PrimBox r = recL[idxD[j]];
                    const bool isLeft = j < nL;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const uint32_t u = orderedBits(r.c[k]);
                        const uint32_t lo = isLeft ? bx[k] : bx[6 + k], hi = isLeft ? bx[3 + k] : bx[9 + k];
                        const uint32_t nlo = u < lo ? u : lo, nhi = u > hi ? u : hi;
                        bx[k] = isLeft ? nlo : bx[k];
                        bx[6 + k] = isLeft ? bx[6 + k] : nlo;
                        bx[3 + k] = isLeft ? nhi : bx[3 + k];
                        bx[9 + k] = isLeft ? bx[9 + k] : nhi;
                    }
                }
#pragma unroll
                for (int k = 0; k < 12; ++k) {
                    const bool isMin = k < 3 || (k >= 6 && k < 9);
                    bx[k] = isMin ? reduce64<MinU>(bx[k]) : reduce64<MaxU>(bx[k]);
                }
                if (lane == 0) {
#pragma unroll
                    for (int k = 0; k < 12; ++k) waveBox[w][k] = bx[k];
                }
                __syncthreads();
#pragma unroll
                for (int k = 0; k < 12; ++k) {
                    const bool isMin = k < 3 || (k >= 6 && k < 9);
                    uint32_t x = waveBox[0][k];
                    for (int v = 1; v < kWaves; ++v) x = isMin ? MinU::op(x, waveBox[v][k]) : MaxU::op(x, waveBox[v][k]);
                    bx[k] = x;
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    sc.lLo[k] = fromOrderedBits(bx[k]);
                    sc.lHi[k] = fromOrderedBits(bx[3 + k]);
                    sc.rLo[k] = fromOrderedBits(bx[6 + k]);
                    sc.rHi[k] = fromOrderedBits(bx[9 + k]);
                }
                __syncthreads();  // waveBox is read before the next split writes it
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
                    // which index buffer holds the range's final order (the leaf order
                    // is written for the whole task at its end), one thread per position
                    if (threadIdx.x < cc[side])
                        finalBuf[cs[side] - base + threadIdx.x] = static_cast<unsigned char>(childBuf);
                    if (threadIdx.x == 0) {
                        if (serN == kSmallSerialBatch) {
                            flushSerial(T, serBuf, serN);
                            serN = 0;
                        }
                        serBuf[serN++] = {cs[side], cc[side], ci[side], kOrderWritten};
                    }
                } else if (sp < kSmallStack) {
                    if (threadIdx.x == 0)
                        stk[sp] = childTask(cs[side], cc[side], ci[side], metaOf(childBuf, level + 1), sc, side);
                    ++sp;
                } else if (threadIdx.x == 0) {
                    atomicOr(err, 1);
                }
            }
            __syncthreads();
        }
        // the task's leaf order (scene ids from the record buffer), all positions at
        // once: the reads overlap instead of one dependent read per serial range
        for (int i = threadIdx.x; i < root.count; i += kBT) {
            const int fb = finalBuf[i];
            if (fb != 0xFF) order[base + i] = static_cast<uint32_t>(src[idxL[fb][i]].id);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        flushSerial(T, serBuf, serN);
        atomicAdd(&T.ctl[kCtlBinned], binned);
        atomicAdd(&T.ctl[kCtlParted], parted);
        atomicMax(&T.ctl[kCtlMaxLevel], deepest);
    }
}

// One wave per final range of at most kSerial primitives (grid-stride over the
// serial list): its leaf order (the scene indices, read from the record buffer that
// holds the range, one lane per position) unless the small-task tier wrote it, and,
// for two or more primitives, a balanced subtree of count/2 splits.  Each internal
// node of that subtree splits after a different position of the range, so lane j
// takes the node that splits after position start + j, found by descending from the
// range's root (at most 6 steps); the nodes of a range are consecutive ids, so the
// wave's writes are contiguous.
__global__ __launch_bounds__(kBlock) void sahSerialKernel(SahTables T, int n, uint32_t* __restrict__ order,
                                                          Node2* __restrict__ nodes, int* __restrict__ parent,
                                                          int* __restrict__ farEnd) {
    static_assert(kSerial <= 64, "one lane per position of a serial range");
    const int count = T.ctl[kCtlSerial] < T.maxSerial ? T.ctl[kCtlSerial] : T.maxSerial;
    const int lane = threadIdx.x & 63;
    const int waves = static_cast<int>(gridDim.x) * (kBlock / 64);
    for (int ti = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); ti < count; ti += waves) {
        const SerialTask t = T.serial[ti];
        if (t.buf != kOrderWritten && lane < t.count)
            order[t.start + lane] = static_cast<uint32_t>(recOf(T, t.buf)[t.start + lane].id);
        if (lane >= t.count - 1) continue;  // count - 1 internal nodes, one per lane
        const int target = t.start + lane;  // the split position this lane's node owns
        int lo = t.start, hi = t.start + t.count - 1, id = t.node;
        int g = lo + (hi - lo + 1) / 2 - 1;
        while (g != target) {
            if (target < g) {
                hi = g;
                id = g;  // an internal left child is numbered by its range's last slot
            } else {
                lo = g + 1;
                id = g + 1;  // an internal right child by its range's first slot
            }
            g = lo + (hi - lo + 1) / 2 - 1;
        }
        const int l = (g == lo) ? n - 1 + lo : g;
        const int r = (g + 1 == hi) ? n - 1 + hi : g + 1;
        nodes[id].left = l;
        nodes[id].right = r;
        parent[l] = id;
        parent[r] = id;
        farEnd[id] = id == lo ? hi : lo;
    }
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
int gBuildPhaseTiming = 1;  // qr_set_build_phase_timing: 0 total only, 1 phases, 2 + SAH levels

template <typename L>
static void buildSahT(BvhState& b, const typename L::Item* dSrc, int n, hipStream_t s, BuildTimingsC* t) {
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
    b.counters.reserve(16);
    uint32_t* cb = reinterpret_cast<uint32_t*>(b.counters.p);

    SahTables T{};
    T.maxBig = n / (kChunk + 1) + 2;
    T.maxLocal = n / (kSerial + 1) + 2;
    T.maxSerial = n + 2;
    T.maxChunks = n / kChunk + T.maxBig + 2;
    // the records, leaf order and task tables, carved from the tree's arena (a dry
    // run sizes it)
    uint32_t* order = nullptr;
    auto carve = [&](unsigned char* base) {
        size_t off = 0;
        auto take = [&](auto*& ptr, size_t count) {
            off = (off + 255) & ~size_t(255);
            ptr = reinterpret_cast<std::remove_reference_t<decltype(ptr)>>(reinterpret_cast<uintptr_t>(base) + off);
            off += count * sizeof(*ptr);
        };
        SahTask* big = nullptr;
        take(T.rec[0], static_cast<size_t>(n));
        take(T.rec[1], static_cast<size_t>(n));
        take(order, static_cast<size_t>(n));
        take(big, 2 * static_cast<size_t>(T.maxBig));
        take(T.local, static_cast<size_t>(T.maxLocal));
        take(T.small, static_cast<size_t>(T.maxLocal));
        take(T.serial, static_cast<size_t>(T.maxSerial));
        take(T.chunks, static_cast<size_t>(T.maxChunks));
        take(T.first, static_cast<size_t>(T.maxBig) + 1);
        take(T.chunkBins, static_cast<size_t>(T.maxChunks) * kParts * 3 * kBins);
        take(T.chunkOff, static_cast<size_t>(T.maxChunks) * kParts);
        take(T.ctl, static_cast<size_t>(kCtlWords));
        take(T.bins, 2 * static_cast<size_t>(T.maxBig) * kHist);
        take(T.results, static_cast<size_t>(T.maxBig));
        T.big[0] = big;
        T.big[1] = reinterpret_cast<SahTask*>(reinterpret_cast<uintptr_t>(big) + T.maxBig * sizeof(SahTask));
        return off;
    };
    b.sahArena.reserve(carve(nullptr));
    carve(b.sahArena.p);
    int* hw = pinnedWords();

    EventSet<6> ev;
    ev.record(0, s);
    sahResetKernel<<<std::min(blocksFor(n - 1 > 256 ? n - 1 : 256, kBlock), 2048), kBlock, 0, s>>>(
        cb, b.arrive.p, n - 1, T, n > kChunk ? 1 : 0);
    int grid = blocksFor(n, kBlock * 4);
    if (grid > 512) grid = 512;  // two workgroups per CU stream the items; few bounds atomics
    sahPrepKernel<L><<<grid, kBlock, 0, s>>>(dSrc, n, T.rec[0], cb, b.parent.p);
    if (n <= kChunk) sahRootKernel<<<1, 64, 0, s>>>(T, cb, n);  // (else the first level's plan routes it)
    hipCheck(hipGetLastError(), "sah prep");
    // phase events only for a caller that asks for the timings with phases on
    const bool phases = t && gBuildPhaseTiming >= 1;
    if (phases) ev.record(1, s);

    // Large tasks, level by level, enqueued in batches with no host work between the
    // levels of a batch; after each batch one read of the next level's task count
    // says whether to go on.  The first batch covers the levels a balanced split
    // would need to bring every task under kChunk.
    std::vector<hipEvent_t> lev;
    // per-level phase events only when asked for (each costs a few microseconds of
    // device time per level)
    auto mark = [&](hipStream_t st) {
        if (gBuildPhaseTiming < 2) return;
        const hipEvent_t e = takeEvent();
        hipCheck(hipEventRecord(e, st), "event record");
        lev.push_back(e);
    };
    int level = 0;
    if (n > kChunk) {
        // the first batch: the levels a balanced split needs to bring every task under
        // kChunk, and two more (an empty level costs four short launches, a host
        // round trip more)
        int batch = 1;
        while ((static_cast<long long>(kChunk) << batch) < n) ++batch;
        batch += 2;
        const int chunkGrid = T.maxChunks;
        while (true) {
            for (int j = 0; j < batch; ++j, ++level) {
                const int cur = level & 1;
                const int median = level >= kMedianLevel ? 1 : 0;
                // bin (the plan's chunk list), split (plane, part offsets, node and
                // children), partition
                mark(s);
                sahPlanKernel<<<1, kBlock, 0, s>>>(T, cur, median, cb, level == 0 ? n : 0);
                sahBinKernel<<<chunkGrid, kBlock, 0, s>>>(T, cur);
                mark(s);
                sahSplitKernel<<<T.maxBig, kBlock, 0, s>>>(T, cur, median, n, b.nodes.p, b.parent.p, b.farEnd.p);
                mark(s);
                sahPartScatterKernel<<<chunkGrid, kBlock, 0, s>>>(T, cur);
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
    if (phases) ev.record(2, s);
    // local subtrees (the host knows their number from the last read), then the
    // serial tails; the bottom-up pass runs only if no kernel set the error word (a
    // device-side gate: no host round trip before it)
    if (hw[kCtlLocal] > 0)
        sahSubtreeKernel<kBlock><<<std::min(hw[kCtlLocal], T.maxLocal), kBlock, 0, s>>>(
            T, n, b.nodes.p, b.parent.p, b.farEnd.p);
    if (hw[kCtlLocal] > 0 || hw[kCtlSmall] > 0) {
        // the top tier adds small tasks: their count is known only on the device, so a
        // fixed grid strides the list
        const int smallGrid = std::min(T.maxLocal, 8192);
        sahSmallKernel<kSmallBT><<<smallGrid, kSmallBT, 0, s>>>(T, n, order, b.nodes.p, b.parent.p, b.farEnd.p);
    }
    sahSerialKernel<<<2048, kBlock, 0, s>>>(T, n, order, b.nodes.p, b.parent.p, b.farEnd.p);
    hipCheck(hipGetLastError(), "sah subtrees");
    if (phases) ev.record(4, s);

    finishTopology(b, dSrc, order, s, &T.ctl[kCtlError]);
    hipCheck(hipMemcpyAsync(hw, T.ctl, kCtlWords * sizeof(int), hipMemcpyDeviceToHost, s), "sah control");
    int depth = 0;
    hipCheck(hipMemcpyAsync(&depth, b.height.p, sizeof(int), hipMemcpyDeviceToHost, s), "depth");
    ev.record(5, s);
    ev.sync(5);
    if (hw[kCtlError] != 0) {
        b.n = 0;  // the topology is incomplete: the tree is not usable
        b.nodeCount = 0;
        failed();
    }
    b.depth = depth;
    SahBuildStats st;
    st.levels = hw[kCtlMaxLevel] + 1;
    st.binnedSlots = static_cast<unsigned int>(hw[kCtlBinned]);
    st.partitionedSlots = static_cast<unsigned int>(hw[kCtlParted]);
    st.serialSlots = static_cast<unsigned int>(n);  // every primitive ends in exactly one serial range
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
        t->mortonMs = phases ? ev.ms(0, 1) : 0.0;
        t->sortMs = 0.0;
        t->treeMs = phases ? ev.ms(1, 4) : 0.0;
        t->scatterMs = phases ? ev.ms(2, 4) : 0.0;  // local subtrees and serial tails
        t->refitMs = phases ? ev.ms(4, 5) : 0.0;
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
