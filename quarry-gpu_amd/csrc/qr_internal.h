// Internal host/device declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "qr_geom.h"

namespace qr {

// ---- device node layouts ------------------------------------------------------
// Binary node, 32 B, byte-identical to the reference's GpuBVHNode
// (src/gpu/gpu_node.h:9-35): internal nodes 0..n-2 hold child ids, leaf n-1+slot
// holds left = -(slot+1), right = 1.
struct Node2 {
    float lo[3];
    float hi[3];
    int left;
    int right;
};
static_assert(sizeof(Node2) == 32, "Node2 layout");

// 4-wide node, 128 B = one cache line (src/gpu/gpu_node.h:38-64): child >= 0 is a
// wide index, child < 0 a leaf reference -(slot+1), kWideEmpty an unused slot.
constexpr int kWideEmpty = INT_MIN;
struct WideEntry {
    float lo[3];
    float hi[3];
    int child;
    int pad;
};
struct alignas(128) Wide4 {
    WideEntry c[4];
};
static_assert(sizeof(Wide4) == 128, "Wide4 layout");

QR_HD Box nodeBox(const Node2& n) {
    return {{n.lo[0], n.lo[1], n.lo[2]}, {n.hi[0], n.hi[1], n.hi[2]}};
}
QR_HD Box entryBox(const WideEntry& e) {
    return {{e.lo[0], e.lo[1], e.lo[2]}, {e.hi[0], e.hi[1], e.hi[2]}};
}

// Result record for closest / kNN (src/gpu/gpu_bvh.h:250-253).
struct ClosestResult {
    uint32_t dist2Bits;
    int32_t prim;
};

// ---- errors -------------------------------------------------------------------
// Device failures -> DeviceError (runtime_error in the C++ API); bad arguments ->
// ArgError; invalid topology edits -> EditError (invalid_argument in the C++ API).
struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct ArgError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct EditError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct OomError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hipCheck(hipError_t e, const char* what);

inline int blocksFor(long long n, int block) { return static_cast<int>((n + block - 1) / block); }

// Plain device allocation owned by a scope (or a borrowed view of memory someone
// else owns: borrow() makes release() forget the pointer instead of freeing it).
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    bool owned = true;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), owned(o.owned) {
        o.p = nullptr;
        o.n = 0;
        o.owned = true;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            owned = o.owned;
            o.p = nullptr;
            o.n = 0;
            o.owned = true;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        owned = true;
        if (count == 0) count = 1;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            throw OomError(std::string("hipMalloc failed: ") + hipGetErrorString(e));
        }
        n = count;
    }
    // Grows to at least count elements; contents are not preserved.  A buffer that
    // grows again takes a quarter more than asked, so that a tree growing by small
    // edits reallocates rarely.
    void reserve(size_t count) {
        if (p == nullptr) {
            alloc(count);
        } else if (count > n) {
            alloc(count + count / 4);
        }
    }
    void release() {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        n = 0;
        owned = true;
    }
    // A view of count elements at q, never freed here.
    void borrow(T* q, size_t count) {
        release();
        p = q;
        n = count;
        owned = false;
    }
    T* release_ownership() {
        T* q = p;
        p = nullptr;
        n = 0;
        return q;
    }
};

// Stream the library issues work on (per host thread; null = legacy default).
hipStream_t currentStream();

// A fixed set of timing events, destroyed with the scope (so a query of the
// elapsed time can never touch a destroyed event).
// Per-host-thread pool of HIP events: a call's timing events are taken from it and
// returned, so no query creates or destroys events.
inline std::vector<hipEvent_t>& eventPool() {
    static thread_local std::vector<hipEvent_t> pool;
    return pool;
}
inline hipEvent_t takeEvent() {
    auto& pool = eventPool();
    if (!pool.empty()) {
        const hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    hipCheck(hipEventCreate(&e), "event create");
    return e;
}
inline void giveEvent(hipEvent_t e) { eventPool().push_back(e); }

template <int N>
struct EventSet {
    hipEvent_t e[N] = {};
    // events come from a per-thread free list (created once, reused by every call)
    EventSet() {
        for (auto& x : e) x = takeEvent();
    }
    ~EventSet() {
        for (auto& x : e)
            if (x) giveEvent(x);
    }
    EventSet(const EventSet&) = delete;
    EventSet& operator=(const EventSet&) = delete;
    void record(int i, hipStream_t s) { hipCheck(hipEventRecord(e[i], s), "event record"); }
    void sync(int i) { hipCheck(hipEventSynchronize(e[i]), "event sync"); }
    float ms(int a, int b) const {
        float v = 0.0f;
        hipCheck(hipEventElapsedTime(&v, e[a], e[b]), "event elapsed");
        return v;
    }
};

// ---- sort ---------------------------------------------------------------------
// Workspace for the onesweep LSD radix sort of 32-bit keys with 32-bit values.
// The sort's global digit histogram is kept in kHistCopies copies of
// [4 passes][256 digits]: a producing workgroup adds its counts into copy
// blockIdx.x % kHistCopies, which divides the same-address atomics by the copy
// count, and the sort sums the copies when it reads a digit's total.
constexpr int kHistCopies = 8;
constexpr int kHistCopyWords = 4 * 256;

__device__ __forceinline__ void flushHistogram(const uint32_t* lds, int words, uint32_t* hist) {
    uint32_t* dst = hist + (blockIdx.x % kHistCopies) * kHistCopyWords;
    for (int i = threadIdx.x; i < words; i += blockDim.x) {
        const uint32_t v = lds[i];
        if (v) atomicAdd(&dst[i], v);
    }
}

struct SortWorkspace {
    DevBuf<uint32_t> keysAlt, valsAlt;   // ping-pong buffers
    DevBuf<uint32_t> hist;               // [kHistCopies][4][256] global digit counts
    DevBuf<uint32_t> desc;               // [passes][tiles][256] look-back descriptors
    DevBuf<uint32_t> tickets;            // [passes] tile tickets
    size_t capacity = 0;
    void reserve(size_t n);
};

// Keys per onesweep tile: 2,048 (8 per thread) below 1M keys, so that a small sort
// still spreads over the CUs; 4,096 (16 per thread) from 1M keys on (fewer tiles,
// shorter look-back walks: 2,048-key tiles measured 25% slower on 2M keys).
constexpr int kSortTileSmall = 2048;
constexpr int kSortTileLarge = 4096;
constexpr size_t kSortLargeFrom = size_t(1) << 20;
// 8,192 keys (32 per thread) from 2M keys on: half the tiles again on the look-back
// walk, two workgroups per CU by LDS.
constexpr int kSortTileHuge = 8192;
constexpr size_t kSortHugeFrom = size_t(1) << 21;

int sortTiles(size_t n);

// Clears the histogram, descriptors and tickets for `passes` passes over n keys.
void sortClear(SortWorkspace& ws, size_t n, int passes, hipStream_t s);

// Sorts (keys, vals) stably by bits [0, 8*passes) of the key.  If vals == nullptr
// the pass-0 values are the element indices.  The histogram for every pass must
// already be in ws.hist (sortHistogram or a fused producer).  Results land in
// keysOut/valsOut.  keys is not modified; buffers must not alias.
void sortPairs(SortWorkspace& ws, const uint32_t* keys, const uint32_t* vals, uint32_t* keysOut,
               uint32_t* valsOut, size_t n, int passes, hipStream_t s);

// Standalone histogram of all passes (when the producer did not fuse it).
void sortHistogram(SortWorkspace& ws, const uint32_t* keys, size_t n, int passes, hipStream_t s);

// Device-side exclusive scan of n ints (reduce / scan block sums / apply).  in and
// out may alias.  scratch must hold scanScratchInts(n) ints; *total_dev (optional)
// receives the sum of all n inputs.
int scanScratchInts(int n);
void scanExclusive(const int* in, int* out, int n, int* total_dev, int* scratch, hipStream_t s);

// ---- cross-lane steps within each 32-lane half of a wave (DPP) -------------------------
// The 32-lane forms serve waves whose lanes 32..63 repeat lanes 0..31 (every step
// works on the two halves alike and a half's result can be read from lanes 0..31);
// reduce64 takes 64 distinct lanes.  All are called by the whole wave.
// kCtrl is a DPP control: row_shr:k 0x110+k (lane i reads lane i-k of its 16-lane
// row), row_shl:k 0x100+k (lane i+k), wave_shr:1 0x138 (lane i-1 across the wave),
// quad_perm [1,0,3,2] 0xB1 and [2,3,0,1] 0x4E, row_half_mirror 0x141, row_mirror
// 0x140.  A lane whose source lies outside its row keeps `id`.
template <int kCtrl>
__device__ __forceinline__ int dppI(int x, int id) {
    return __builtin_amdgcn_update_dpp(id, x, kCtrl, 0xF, 0xF, false);
}
template <int kCtrl>
__device__ __forceinline__ float dppF(float x, float id) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(id), __float_as_int(x), kCtrl, 0xF, 0xF, false));
}
__device__ __forceinline__ float readlaneF(float x, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

struct MinF {
    static constexpr float id = kInf;
    __device__ static float op(float a, float b) { return fminf(a, b); }
};
struct MaxF {
    static constexpr float id = -kInf;
    __device__ static float op(float a, float b) { return fmaxf(a, b); }
};

// Inclusive prefix over the lanes of each half: four row shifts, then rows 1 and 3
// take the total of the row below (lane 15 of the half; the halves are equal).
template <typename Op>
__device__ __forceinline__ float prefix32(float x, int lane) {
    x = Op::op(x, dppF<0x111>(x, Op::id));
    x = Op::op(x, dppF<0x112>(x, Op::id));
    x = Op::op(x, dppF<0x114>(x, Op::id));
    x = Op::op(x, dppF<0x118>(x, Op::id));
    const float low = readlaneF(x, 15);
    return (lane & 16) ? Op::op(x, low) : x;
}
__device__ __forceinline__ int prefixAdd32(int x, int lane) {
    x += dppI<0x111>(x, 0);
    x += dppI<0x112>(x, 0);
    x += dppI<0x114>(x, 0);
    x += dppI<0x118>(x, 0);
    const int low = __builtin_amdgcn_readlane(x, 15);
    return (lane & 16) ? x + low : x;
}
// Inclusive suffix (the lane and the lanes above it in its half).
template <typename Op>
__device__ __forceinline__ float suffix32(float x, int lane) {
    x = Op::op(x, dppF<0x101>(x, Op::id));
    x = Op::op(x, dppF<0x102>(x, Op::id));
    x = Op::op(x, dppF<0x104>(x, Op::id));
    x = Op::op(x, dppF<0x108>(x, Op::id));
    const float high = readlaneF(x, 16);
    return (lane & 16) ? x : Op::op(x, high);
}
// Reduction over a half, the same value returned to every lane.
template <typename Op>
__device__ __forceinline__ uint32_t reduce32(uint32_t x) {
    int v = static_cast<int>(x);
    v = static_cast<int>(Op::op(static_cast<uint32_t>(v), static_cast<uint32_t>(dppI<0xB1>(v, 0))));
    v = static_cast<int>(Op::op(static_cast<uint32_t>(v), static_cast<uint32_t>(dppI<0x4E>(v, 0))));
    v = static_cast<int>(Op::op(static_cast<uint32_t>(v), static_cast<uint32_t>(dppI<0x141>(v, 0))));
    v = static_cast<int>(Op::op(static_cast<uint32_t>(v), static_cast<uint32_t>(dppI<0x140>(v, 0))));
    return Op::op(static_cast<uint32_t>(__builtin_amdgcn_readlane(v, 0)),
                  static_cast<uint32_t>(__builtin_amdgcn_readlane(v, 16)));
}
// Reduction over a whole wave (lanes 0..63 all distinct), the same value returned to
// every lane.
template <typename Op>
__device__ __forceinline__ uint32_t reduce64(uint32_t x) {
    int v = static_cast<int>(x);
    v = static_cast<int>(Op::op(static_cast<uint32_t>(v), static_cast<uint32_t>(dppI<0xB1>(v, 0))));
    v = static_cast<int>(Op::op(static_cast<uint32_t>(v), static_cast<uint32_t>(dppI<0x4E>(v, 0))));
    v = static_cast<int>(Op::op(static_cast<uint32_t>(v), static_cast<uint32_t>(dppI<0x141>(v, 0))));
    v = static_cast<int>(Op::op(static_cast<uint32_t>(v), static_cast<uint32_t>(dppI<0x140>(v, 0))));
    const uint32_t r0 = Op::op(static_cast<uint32_t>(__builtin_amdgcn_readlane(v, 0)),
                               static_cast<uint32_t>(__builtin_amdgcn_readlane(v, 16)));
    const uint32_t r1 = Op::op(static_cast<uint32_t>(__builtin_amdgcn_readlane(v, 32)),
                               static_cast<uint32_t>(__builtin_amdgcn_readlane(v, 48)));
    return Op::op(r0, r1);
}
struct MinU {
    __device__ static uint32_t op(uint32_t a, uint32_t b) { return a < b ? a : b; }
};
struct MaxU {
    __device__ static uint32_t op(uint32_t a, uint32_t b) { return a > b ? a : b; }
};
struct AddU {
    __device__ static uint32_t op(uint32_t a, uint32_t b) { return a + b; }
};


}  // namespace qr
