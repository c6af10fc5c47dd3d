// LSD onesweep radix sort of 32-bit keys with 32-bit payloads, written for
// gfx950 wave64 (one global histogram pass + one scatter pass per 8-bit digit).
//
// Each 256-thread workgroup owns a tile of 2,048 or 4,096 keys (8 or 16 per
// thread).  Keys are ranked inside a wave with 9 ballots (8 digit bits + a
// validity bit) and a per-wave LDS digit
// counter, then across waves with an LDS prefix, then across tiles with a
// decoupled look-back over per-tile digit descriptors.  Tile ids come from an
// atomic ticket, so every predecessor a tile waits on is already resident and
// publishes its aggregate before it looks back itself: the spin always ends.
// Keys are staged through LDS in tile-sorted order before the global write so
// that consecutive lanes store consecutive addresses within a digit run.
//
// The BVH build sorts 64-bit keys (code << 32 | index) in the reference
// (src/gpu/lbvh.cu:1995-2003).  The low half is the input position, so a stable
// sort of the 30-bit code with the position as payload yields the same order;
// only ceil(30/8) = 4 digit passes are needed instead of 8.
#include "qr_internal.h"

namespace qr {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kStatAgg = 1u << 30;
constexpr uint32_t kStatIncl = 2u << 30;
constexpr uint32_t kValMask = (1u << 30) - 1u;
// Threads per 8,192-key tile: 1,024 up to kSortWideTilesMax tiles (a few per CU,
// where each tile's latency shows: 2M keys 0.119 -> 0.093 ms), 256 above (two tiles
// per CU hide it better: 48M keys 0.338 -> 0.382 ms per pass with 1,024).
constexpr int kSortHugeThreads = 1024;
constexpr int kSortWideTilesMax = 1024;
constexpr int kLookBack = 8;  // predecessors read per look-back round (32 measured slower: 0.144 vs 0.120 ms for the 2M sort)

__device__ __forceinline__ uint32_t loadAgent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void storeAgent(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Clears histogram, descriptors and tickets (a kernel, not a memset, so the
// whole build can be captured in a graph).
__global__ void sortClearKernel(uint32_t* __restrict__ hist, uint32_t* __restrict__ desc,
                                long long descWords,
                                uint32_t* __restrict__ tickets, int passes) {
    const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long j = i; j < descWords; j += stride) desc[j] = 0u;
    for (long long j = i; j < kHistCopies * kHistCopyWords; j += stride) hist[j] = 0u;
    if (i < passes) tickets[i] = 0u;
}

__global__ void sortHistogramKernel(const uint32_t* __restrict__ keys, long long n, int passes,
                                    uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[4][256];
    for (int i = threadIdx.x; i < 4 * 256; i += kBlock) (&h[0][0])[i] = 0u;
    __syncthreads();
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += stride) {
        const uint32_t k = keys[i];
        for (int p = 0; p < passes; ++p) atomicAdd(&h[p][(k >> (8 * p)) & 255u], 1u);
    }
    __syncthreads();
    flushHistogram(&h[0][0], passes * 256, hist);
}

// Exclusive scan of the values held one per thread by the first 256 threads (the
// other threads hold 0); returns the prefix and the total through *total.
template <int kThreads = kBlock>
__device__ __forceinline__ uint32_t blockScan256(uint32_t v, uint32_t* scratch,
                                                 uint32_t* total) {
    constexpr int kW = kThreads / 64;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    // inclusive wave scan
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) scratch[w] = x;
    __syncthreads();
    uint32_t waveBase = 0;
    uint32_t all = 0;
#pragma unroll
    for (int i = 0; i < kW; ++i) {
        const uint32_t s = scratch[i];
        if (i < w) waveBase += s;
        all += s;
    }
    __syncthreads();
    *total = all;
    return waveBase + x - v;
}

// kThreads = 256 or 1024: a tile of kKpt * kThreads keys; with 1024 threads four
// waves share each SIMD, so the per-row ranking chain is a quarter as long and its
// latency, and the scatter's, is hidden behind the other waves.  Digit-wise steps
// (prefixes, descriptors, look-back) run on the first 256 threads.
template <int kKpt, int kThreads = kBlock>
__global__ __launch_bounds__(kThreads) void onesweepKernel(
    const uint32_t* __restrict__ keysIn, const uint32_t* __restrict__ valsIn,
    uint32_t* __restrict__ keysOut, uint32_t* __restrict__ valsOut, long long n, int shift,
    const uint32_t* __restrict__ histPass, uint32_t* __restrict__ desc,
    uint32_t* __restrict__ ticket) {
    constexpr int kW = kThreads / 64;
    __shared__ uint32_t waveCnt[kW][256];
    __shared__ uint32_t tileStart[256];
    __shared__ uint32_t globalBase[256];
    __shared__ uint32_t scanScratch[kW];
    constexpr int kTile = kKpt * kThreads;
    __shared__ uint32_t stageKeys[kTile];
    __shared__ uint32_t stageVals[kTile];
    __shared__ int tileShared;

    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;

    const bool digitThread = t < 256;  // thread t owns digit t in the digit-wise steps
    if (t == 0) tileShared = static_cast<int>(atomicAdd(ticket, 1u));
    for (int i = t; i < kW * 256; i += kThreads) (&waveCnt[0][0])[i] = 0u;
    __syncthreads();
    const int tile = tileShared;
    const long long tileBase = static_cast<long long>(tile) * kTile;
    const long long remain = n - tileBase;
    const int validCount = remain >= kTile ? kTile : static_cast<int>(remain);

    uint32_t key[kKpt];
    uint32_t val[kKpt];
    uint32_t rank[kKpt];
    const long long waveBase = tileBase + static_cast<long long>(w) * (kTile / kW);
#pragma unroll
    for (int r = 0; r < kKpt; ++r) {
        const long long idx = waveBase + r * 64 + lane;
        if (idx < n) {
            key[r] = keysIn[idx];
            val[r] = valsIn ? valsIn[idx] : static_cast<uint32_t>(idx);
        } else {
            key[r] = 0u;
            val[r] = 0u;
        }
    }

    const uint64_t ltMask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < kKpt; ++r) {
        const long long idx = waveBase + r * 64 + lane;
        const bool valid = idx < n;
        const uint32_t d = valid ? ((key[r] >> shift) & 255u) : 256u;
        uint64_t peers = ~0ull;
#pragma unroll
        for (int b = 0; b < 9; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t below = static_cast<uint32_t>(__popcll(peers & ltMask));
        uint32_t base = 0;
        if (valid) base = waveCnt[w][d];
        rank[r] = base + below;
        if (valid && below == 0) waveCnt[w][d] = base + static_cast<uint32_t>(__popcll(peers));
    }
    __syncthreads();

    // Per-digit tile totals and per-wave exclusive prefixes (thread t = digit t).
    uint32_t tileTotal = 0;
    uint32_t* myDesc = desc + static_cast<size_t>(tile) * 256;
    if (digitThread) {
        uint32_t run = 0;
#pragma unroll
        for (int i = 0; i < kW; ++i) {
            const uint32_t c = waveCnt[i][t];
            waveCnt[i][t] = run;
            run += c;
        }
        tileTotal = run;
        if (tile == 0) {
            storeAgent(&myDesc[t], kStatIncl | tileTotal);
        } else {
            storeAgent(&myDesc[t], kStatAgg | tileTotal);
        }
    }

    uint32_t dummy;
    const uint32_t startInTile = blockScan256<kThreads>(tileTotal, scanScratch, &dummy);
    if (digitThread) tileStart[t] = startInTile;
    uint32_t digitTotal = 0;
    if (digitThread) {
#pragma unroll
        for (int c = 0; c < kHistCopies; ++c) digitTotal += histPass[c * kHistCopyWords + t];
    }
    const uint32_t bucketStart = blockScan256<kThreads>(digitTotal, scanScratch, &dummy);

    // Look-back over a window of kLookBack predecessors per round (loads in flight
    // together): aggregates are summed in order up to the first inclusive prefix; a
    // predecessor that has not published yet ends the round, which resumes there.
    // Tile 0 publishes an inclusive prefix at once, so the walk ends.
    uint32_t exclusive = 0;
    if (digitThread && tile > 0) {
        int j = tile - 1;
        while (true) {
            uint32_t v[kLookBack];
#pragma unroll
            for (int u = 0; u < kLookBack; ++u)
                v[u] = j - u >= 0 ? loadAgent(&desc[static_cast<size_t>(j - u) * 256 + t]) : 0u;
            int used = 0;
            bool done = false;
#pragma unroll
            for (int u = 0; u < kLookBack; ++u) {
                const uint32_t st = v[u] & ~kValMask;
                if (st == 0u) break;
                exclusive += v[u] & kValMask;
                ++used;
                if (st == kStatIncl) {
                    done = true;
                    break;
                }
            }
            if (done) break;
            j -= used;
            if (used < kLookBack) __builtin_amdgcn_s_sleep(1);
        }
        storeAgent(&myDesc[t], kStatIncl | (exclusive + tileTotal));
    }
    if (digitThread) globalBase[t] = bucketStart + exclusive;
    __syncthreads();

#pragma unroll
    for (int r = 0; r < kKpt; ++r) {
        const long long idx = waveBase + r * 64 + lane;
        if (idx < n) {
            const uint32_t d = (key[r] >> shift) & 255u;
            const uint32_t pos = tileStart[d] + waveCnt[w][d] + rank[r];
            stageKeys[pos] = key[r];
            stageVals[pos] = val[r];
        }
    }
    __syncthreads();

#pragma unroll
    for (int j = 0; j < kKpt; ++j) {
        const int i = j * kThreads + t;
        if (i < validCount) {
            const uint32_t k = stageKeys[i];
            const uint32_t d = (k >> shift) & 255u;
            const size_t dest = static_cast<size_t>(globalBase[d]) + (i - tileStart[d]);
            keysOut[dest] = k;
            valsOut[dest] = stageVals[i];
        }
    }
}

// ---- exclusive scan of int (reduce, scan block sums, scan + add) ------------------
constexpr int kScanBlock = 256;
constexpr int kScanPer = 16;
constexpr int kScanTile = kScanBlock * kScanPer;

__global__ void scanReduceKernel(const int* __restrict__ in, int n, int* __restrict__ sums) {
    __shared__ int s[kScanBlock / 64];
    const int base = blockIdx.x * kScanTile;
    int acc = 0;
    for (int j = 0; j < kScanPer; ++j) {
        const int i = base + j * kScanBlock + threadIdx.x;
        if (i < n) acc += in[i];
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int i = 0; i < kScanBlock / 64; ++i) tot += s[i];
        sums[blockIdx.x] = tot;
    }
}

// Single workgroup: exclusive scan of the block sums in place; writes the grand
// total to *total.
__global__ void scanSumsKernel(int* __restrict__ sums, int count, int* __restrict__ total) {
    __shared__ int s[kScanBlock / 64];
    __shared__ int carry;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    if (t == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < count; base += kScanBlock) {
        const int i = base + t;
        const int v = i < count ? sums[i] : 0;
        int x = v;
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) s[w] = x;
        __syncthreads();
        int wb = 0;
        int all = 0;
        for (int k = 0; k < kScanBlock / 64; ++k) {
            if (k < w) wb += s[k];
            all += s[k];
        }
        const int c = carry;
        if (i < count) sums[i] = c + wb + x - v;
        __syncthreads();
        if (t == 0) carry = c + all;
        __syncthreads();
    }
    if (t == 0 && total) *total = carry;
}

__global__ void scanApplyKernel(const int* __restrict__ in, int n, const int* __restrict__ sums,
                                int* __restrict__ out) {
    __shared__ int s[kScanBlock / 64];
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    const int base = blockIdx.x * kScanTile;
    int carry = sums[blockIdx.x];
    for (int j = 0; j < kScanPer; ++j) {
        const int i = base + j * kScanBlock + t;
        const int v = i < n ? in[i] : 0;
        int x = v;
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) s[w] = x;
        __syncthreads();
        int wb = 0;
        int all = 0;
        for (int k = 0; k < kScanBlock / 64; ++k) {
            if (k < w) wb += s[k];
            all += s[k];
        }
        if (i < n) out[i] = carry + wb + x - v;
        carry += all;
        __syncthreads();
    }
}

}  // namespace

static size_t sortTileKeys(size_t n) {
    return n < kSortLargeFrom ? kSortTileSmall : n < kSortHugeFrom ? kSortTileLarge : kSortTileHuge;
}

int sortTiles(size_t n) { return static_cast<int>((n + sortTileKeys(n) - 1) / sortTileKeys(n)); }

void SortWorkspace::reserve(size_t n) {
    if (n <= capacity && keysAlt.p) return;
    size_t cap = n < 4096 ? 4096 : n;
    if (keysAlt.p) cap += cap / 4;  // growing again: a quarter more, as DevBuf::reserve
    keysAlt.alloc(cap);
    valsAlt.alloc(cap);
    hist.alloc(static_cast<size_t>(kHistCopies) * kHistCopyWords);
    // enough descriptors for any n <= cap at the smaller tile
    desc.alloc(static_cast<size_t>(4) * ((cap + kSortTileSmall - 1) / kSortTileSmall) * 256);
    tickets.alloc(4);
    capacity = cap;
}

void sortClear(SortWorkspace& ws, size_t n, int passes, hipStream_t s) {
    const long long descWords = static_cast<long long>(passes) * sortTiles(n) * 256;
    long long work = descWords > 1024 ? descWords : 1024;
    int grid = blocksFor(work, kBlock);
    if (grid > 2048) grid = 2048;
    sortClearKernel<<<grid, kBlock, 0, s>>>(ws.hist.p, ws.desc.p, descWords, ws.tickets.p, passes);
    hipCheck(hipGetLastError(), "sort clear");
}

void sortHistogram(SortWorkspace& ws, const uint32_t* keys, size_t n, int passes, hipStream_t s) {
    int grid = blocksFor(static_cast<long long>(n), kBlock * 8);
    if (grid < 1) grid = 1;
    if (grid > 1024) grid = 1024;
    sortHistogramKernel<<<grid, kBlock, 0, s>>>(keys, static_cast<long long>(n), passes,
                                                ws.hist.p);
    hipCheck(hipGetLastError(), "sort histogram");
}

void sortPairs(SortWorkspace& ws, const uint32_t* keys, const uint32_t* vals, uint32_t* keysOut,
               uint32_t* valsOut, size_t n, int passes, hipStream_t s) {
    if (n == 0) return;
    if (ws.capacity < n) throw ArgError("sort workspace too small");
    const int tiles = sortTiles(n);
    // Pass p reads src, writes dst; the last pass must land in keysOut/valsOut.
    const uint32_t* srcK = keys;
    const uint32_t* srcV = vals;
    for (int p = 0; p < passes; ++p) {
        const bool last = p == passes - 1;
        // Alternate so that the final pass writes the caller's output.
        const bool toOut = ((passes - 1 - p) % 2) == 0;
        uint32_t* dstK = toOut ? keysOut : ws.keysAlt.p;
        uint32_t* dstV = toOut ? valsOut : ws.valsAlt.p;
        (void)last;
        uint32_t* desc = ws.desc.p + static_cast<size_t>(p) * tiles * 256;
        if (sortTileKeys(n) == kSortTileSmall)
            onesweepKernel<kSortTileSmall / kBlock><<<tiles, kBlock, 0, s>>>(
                srcK, srcV, dstK, dstV, static_cast<long long>(n), 8 * p, ws.hist.p + p * 256, desc,
                ws.tickets.p + p);
        else if (sortTileKeys(n) == kSortTileLarge)
            onesweepKernel<kSortTileLarge / kBlock><<<tiles, kBlock, 0, s>>>(
                srcK, srcV, dstK, dstV, static_cast<long long>(n), 8 * p, ws.hist.p + p * 256, desc,
                ws.tickets.p + p);
        else if (tiles <= kSortWideTilesMax)
            onesweepKernel<kSortTileHuge / kSortHugeThreads, kSortHugeThreads><<<tiles, kSortHugeThreads, 0, s>>>(
                srcK, srcV, dstK, dstV, static_cast<long long>(n), 8 * p, ws.hist.p + p * 256, desc,
                ws.tickets.p + p);
        else
            onesweepKernel<kSortTileHuge / kBlock><<<tiles, kBlock, 0, s>>>(
                srcK, srcV, dstK, dstV, static_cast<long long>(n), 8 * p, ws.hist.p + p * 256, desc,
                ws.tickets.p + p);
        hipCheck(hipGetLastError(), "onesweep");
        srcK = dstK;
        srcV = dstV;
    }
}

int scanScratchInts(int n) { return blocksFor(n > 0 ? n : 1, kScanTile) + 1; }

void scanExclusive(const int* in, int* out, int n, int* total_dev, int* scratch, hipStream_t s) {
    if (n <= 0) return;
    const int blocks = blocksFor(n, kScanTile);
    scanReduceKernel<<<blocks, kScanBlock, 0, s>>>(in, n, scratch);
    scanSumsKernel<<<1, kScanBlock, 0, s>>>(scratch, blocks, total_dev);
    scanApplyKernel<<<blocks, kScanBlock, 0, s>>>(in, n, scratch, out);
    hipCheck(hipGetLastError(), "scan");
}

}  // namespace qr
