// C ABI of libquarry_hip.so (declared in include/quarry/quarry_abi.h): handle
// management, host<->device transfers, timing and error mapping around the HIP
// kernels.  No kernel math lives here.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/quarry/quarry_abi.h"
#include "qr_adjoint.h"
#include "qr_bvh.h"
#include "qr_traverse.h"

using namespace qr;

struct qr_bvh {
    BvhState s;
};

struct qr_rays {
    DevBuf<Ray> rays;
    DevBuf<uint64_t> results;
    int count = 0;
    // reordering scratch, allocated on first use; bb holds the origin bounds
    // measured at upload (the reference computes them on the host as well,
    // src/gpu/raycast.cu:315-328)
    DevBuf<uint32_t> keys, keysSorted, order, bb;
    bool bbValid = false;
    SortWorkspace ws;
};

struct qr_points {
    DevBuf<V3> pts;
    DevBuf<ClosestResult> results;
    int count = 0;
    int stride = 1;
    // bb: query bounds measured at upload (src/gpu/query.cu:822-847)
    DevBuf<uint32_t> keys, keysSorted, order, bb;
    bool bbValid = false;
    SortWorkspace ws;
    // radius scratch, kept across calls: per-query counts, scan / classification
    // scratch, [0] = largest per-query count, and the captured first hits per query
    DevBuf<int> rCounts, rScratch, rMax;
    DevBuf<uint64_t> rCapture;
    // the last radius call's hit total plus an eighth: the size of the result keys
    // compacted before the total reaches the host (0: none yet)
    int rKeysGuess = 0;
};

// Result buffers handed to the caller: stream-ordered allocations from the device's
// default pool, kept warm (release threshold raised once), so a radius call does not
// stall the device at hipMalloc.  Freed on the null stream: calls are synchronous,
// nothing still runs on them when the caller destroys the result.
void warmDefaultPool() {
    static thread_local int warmed = -1;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    if (warmed == dev) return;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t keep = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    (void)hipGetLastError();
    warmed = dev;
}

template <typename T>
struct AsyncBuf {
    T* p = nullptr;
    AsyncBuf() = default;
    AsyncBuf(const AsyncBuf&) = delete;
    AsyncBuf& operator=(const AsyncBuf&) = delete;
    ~AsyncBuf() { release(); }
    void alloc(size_t count, hipStream_t s) {
        release();
        warmDefaultPool();
        const hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&p), (count > 0 ? count : 1) * sizeof(T), s);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            throw OomError(std::string("hipMallocAsync failed: ") + hipGetErrorString(e));
        }
    }
    void release() {
        if (p) (void)hipFreeAsync(p, nullptr);
        p = nullptr;
    }
};

struct qr_csr {
    AsyncBuf<int> offsets;
    AsyncBuf<uint64_t> keys;
    int count = 0;
    int total = 0;
};

// Pinned host words for the small device-to-host reads of a call (per host thread).
int* hostWords() {
    static thread_local int* w = nullptr;
    if (!w) hipCheck(hipHostMalloc(reinterpret_cast<void**>(&w), (1 + kMaxCopies) * sizeof(int), hipHostMallocDefault),
                     "pinned words");
    return w;
}
// The same words as the device addresses them (kernels store the call's figures
// there directly).
int* hostWordsDevice() {
    static thread_local int* d = nullptr;
    if (!d) hipCheck(hipHostGetDevicePointer(reinterpret_cast<void**>(&d), hostWords(), 0), "pinned words map");
    return d;
}

struct qr_batch {
    BatchState s;
};

// A prepared batched query: the query set, its scratch, the chosen kernel variant
// and the captured graph (work-queue reset -> per-scene Morton order of the queries
// -> persistent traversal) that each run replays.
struct qr_batch_plan {
    qr_batch* batch = nullptr;
    const V3* points = nullptr;
    int n = 0, k = 1, stride = 1, scenes = 0;
    int variant = kBatchLds256;
    bool morton = true;
    double tunedMs = 0.0;
    DevBuf<int> qOff, qScene, counter;
    DevBuf<ClosestResult> results;
    DevBuf<uint32_t> cb, codes, tmpA, tmpB, sceneSorted, order;
    SortWorkspace ws;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~qr_batch_plan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace qr {

void hipCheck(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        (void)hipGetLastError();
        if (e == hipErrorOutOfMemory) throw OomError(std::string("HIP ") + what + ": " + hipGetErrorString(e));
        throw DeviceError(std::string("HIP ") + what + ": " + hipGetErrorString(e));
    }
}

namespace {
thread_local hipStream_t tStream = nullptr;
}
hipStream_t currentStream() { return tStream; }

}  // namespace qr

namespace {

thread_local std::string tLastError;
thread_local qr_traversal_stats tRayStats{};
thread_local qr_traversal_stats tQueryStats{};

template <typename F>
int guard(F&& f) {
    try {
        f();
        return QR_OK;
    } catch (const ArgError& e) {
        tLastError = e.what();
        return QR_ERR_INVALID;
    } catch (const EditError& e) {
        tLastError = e.what();
        return QR_ERR_EDIT;
    } catch (const OomError& e) {
        tLastError = e.what();
        return QR_ERR_OOM;
    } catch (const std::exception& e) {
        tLastError = e.what();
        return QR_ERR_DEVICE;
    }
}

void need(bool ok, const char* msg) {
    if (!ok) throw ArgError(msg);
}

struct Timer {
    EventSet<2> ev;
    hipStream_t s;
    explicit Timer(hipStream_t st) : s(st) { ev.record(0, s); }
    double stop() {
        ev.record(1, s);
        ev.sync(1);
        return ev.ms(0, 1);
    }
};

// Traversal counters of a call: one buffer per host thread, allocated on first use
// and kept for the life of the process (never freed: no HIP call runs at exit).
struct Stats {
    unsigned long long* p = nullptr;
    Stats() {
        static thread_local unsigned long long* buf = nullptr;
        if (!buf) hipCheck(hipMalloc(reinterpret_cast<void**>(&buf), kStatWords * sizeof(unsigned long long)), "stats buffer");
        p = buf;
    }
};

void readStats(const unsigned long long* dev, qr_traversal_stats* out) {
    unsigned long long h[kStatWords];
    hipCheck(hipMemcpy(h, dev, sizeof(h), hipMemcpyDeviceToHost), "stats");
    unsigned long long sum[4] = {0, 0, 0, 0};
    for (int i = 0; i < kStatWords; ++i) sum[i % 4] += h[i];
    out->node_tests = sum[0];
    out->leaf_tests = sum[1];
    out->stack_overflows = sum[2];
    out->nodes_opened = sum[3];
}

// arm: point the kernels at the overflow word first (allocates it once per device)
TraverseArgs argsFor(const BvhState& b, bool wide, bool arm = true) {
    if (arm) armOverflowWord();
    TraverseArgs a;
    a.kind = b.kind;
    a.wide = wide;
    a.nodes = b.nodes.p;
    a.nodeCount = b.nodeCount;
    a.wideNodes = b.wide.p;
    a.wideCount = b.wideCount;
    a.items = b.items.p;
    a.primIds = b.primIds.p;
    return a;
}

void checkTree(const qr_bvh* bvh, bool wide) {
    need(bvh != nullptr, "null bvh handle");
    need(bvh->s.n >= 2, "bvh is empty");
    if (wide && !bvh->s.wideValid) throw ArgError("wide nodes not built");
}

// Morton visiting order for a buffer of elements `stride` floats apart.
template <typename Buf>
const uint32_t* reorder(Buf& buf, const float* base, int stride, int count, int boxMode,
                        hipStream_t s) {
    buf.keys.reserve(static_cast<size_t>(count));
    buf.keysSorted.reserve(static_cast<size_t>(count));
    buf.order.reserve(static_cast<size_t>(count));
    buf.bb.reserve(8);
    mortonOrder(base, stride, count, boxMode, buf.ws, buf.keys.p, buf.keysSorted.p, buf.order.p,
                buf.bb.p, s);
    return buf.order.p;
}

// Bounds of the first three floats of n elements `stride` floats apart, as the
// ordered-bit words the Morton kernels read (lo xyz, hi xyz).
void hostBounds(const float* base, int stride, int n, uint32_t out[6]) {
    float lo[3] = {kInf, kInf, kInf}, hi[3] = {-kInf, -kInf, -kInf};
    for (int i = 0; i < n; ++i) {
        const float* p = base + static_cast<size_t>(i) * stride;
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::fmin(lo[a], p[a]);
            hi[a] = std::fmax(hi[a], p[a]);
        }
    }
    for (int a = 0; a < 3; ++a) {
        out[a] = orderedBits(lo[a]);
        out[3 + a] = orderedBits(hi[a]);
    }
}

void buildInto(BvhState& b, int prim, int builder, const void* dItems, int n, qr_build_timings* t) {
    need(prim >= 0 && prim <= 2, "unknown primitive kind");
    need(builder == QR_BUILDER_LBVH || builder == QR_BUILDER_SAH, "unknown builder");
    need(n >= 2, "at least 2 primitives are required");
    BuildTimingsC bt{};
    // no timings asked for: the builders record no phase events
    BuildTimingsC* want = t ? &bt : nullptr;
    if (builder == QR_BUILDER_LBVH) {
        buildLbvhDevice(b, prim, dItems, n, currentStream(), want);
    } else {
        buildSahDevice(b, prim, dItems, n, currentStream(), want);
    }
    if (t) {
        t->morton_ms = bt.mortonMs;
        t->sort_ms = bt.sortMs;
        t->tree_ms = bt.treeMs;
        t->scatter_ms = bt.scatterMs;
        t->refit_ms = bt.refitMs;
        t->sah_bin_ms = bt.sahBinMs;
        t->sah_split_ms = bt.sahSplitMs;
        t->sah_part_ms = bt.sahPartMs;
        t->total_ms = bt.totalMs;
    }
}

}  // namespace

extern "C" {

const char* qr_last_error(void) { return tLastError.c_str(); }

int qr_describe(char* buf, int len) {
    int devices = 0;
    const hipError_t e = hipGetDeviceCount(&devices);
    if (e != hipSuccess || devices == 0) {
        (void)hipGetLastError();
        tLastError = e != hipSuccess ? hipGetErrorString(e) : "no HIP device";
        if (buf && len > 0) std::snprintf(buf, static_cast<size_t>(len), "%s", tLastError.c_str());
        return QR_ERR_NO_DEVICE;
    }
    return guard([&] {
        int dev = 0;
        hipCheck(hipGetDevice(&dev), "get device");
        hipDeviceProp_t prop{};
        hipCheck(hipGetDeviceProperties(&prop, dev), "device properties");
        if (buf && len > 0) {
            std::snprintf(buf, static_cast<size_t>(len), "%s %s (%d CU, %.1f GiB)", prop.name,
                          prop.gcnArchName, prop.multiProcessorCount,
                          static_cast<double>(prop.totalGlobalMem) / (1024.0 * 1024.0 * 1024.0));
        }
    });
}

int qr_warmup(void) {
    return guard([&] {
        DevBuf<unsigned long long> d(kStatWords);
        zeroStats(d.p, currentStream());
        hipCheck(hipStreamSynchronize(currentStream()), "warmup");
    });
}

int qr_last_sah_stats(qr_sah_stats* out) {
    return guard([&] {
        need(out != nullptr, "null argument");
        out->levels = tSahStats.levels;
        out->binned_slots = tSahStats.binnedSlots;
        out->partitioned_slots = tSahStats.partitionedSlots;
        out->serial_slots = tSahStats.serialSlots;
    });
}

int qr_set_traversal_mode(int mode) {
    return guard([&] {
        need(mode >= 0 && mode <= 6, "traversal mode must be 0 to 6");
        setTraversalMode(mode);
    });
}

int qr_set_frontier_cap(int cap) {
    return guard([&] { setFrontierCap(cap); });
}

int qr_set_build_phase_timing(int on) {
    return guard([&] {
        need(on >= 0 && on <= 2, "phase timing must be 0, 1 or 2");
        gBuildPhaseTiming = on;
    });
}

int qr_set_stream(void* s) {
    tStream = static_cast<hipStream_t>(s);
    return QR_OK;
}

int qr_dev_alloc(size_t bytes, void** out) {
    return guard([&] {
        need(out != nullptr, "null output pointer");
        *out = nullptr;
        hipCheck(hipMalloc(out, bytes > 0 ? bytes : 1), "hipMalloc");
    });
}

int qr_dev_free(void* ptr) {
    return guard([&] {
        if (ptr) hipCheck(hipFree(ptr), "hipFree");
    });
}

int qr_dev_upload(void* dst, const void* src, size_t bytes) {
    return guard([&] {
        if (bytes) hipCheck(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "H2D");
    });
}

int qr_dev_download(void* dst, const void* src, size_t bytes) {
    return guard([&] {
        if (bytes) hipCheck(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "D2H");
    });
}

namespace {
// Streaming device copy, 16 B per lane per iteration (grid-stride); the tail
// below 16 B is copied bytewise by the first lanes.
__global__ __launch_bounds__(256) void copyKernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                  size_t vec, const uint8_t* __restrict__ srcTail,
                                                  uint8_t* __restrict__ dstTail, int tail) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < vec; i += stride)
        dst[i] = src[i];
    const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < static_cast<size_t>(tail)) dstTail[t] = srcTail[t];
}
}  // namespace

int qr_dev_copy(void* dst, const void* src, size_t bytes, double* ms) {
    return guard([&] {
        need(bytes == 0 || (dst != nullptr && src != nullptr), "null pointer");
        need((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) % 16 == 0,
             "pointers must be 16-byte aligned");
        hipStream_t s = currentStream();
        EventSet<2> ev;
        ev.record(0, s);
        if (bytes) {
            const size_t vec = bytes / 16;
            const int tail = static_cast<int>(bytes % 16);
            const size_t want = (vec + 255) / 256;
            const int blocks = static_cast<int>(std::max<size_t>(1, std::min<size_t>(want, 8192)));
            copyKernel<<<blocks, 256, 0, s>>>(static_cast<const uint4*>(src), static_cast<uint4*>(dst),
                                              vec, static_cast<const uint8_t*>(src) + vec * 16,
                                              static_cast<uint8_t*>(dst) + vec * 16, tail);
            hipCheck(hipGetLastError(), "copy kernel");
        }
        ev.record(1, s);
        ev.sync(1);
        if (ms) *ms = ev.ms(0, 1);
    });
}

int qr_dev_synchronize(void) {
    return guard([&] { hipCheck(hipDeviceSynchronize(), "synchronize"); });
}

int qr_bvh_build(int prim, int builder, const void* items, int n, qr_build_timings* t,
                 qr_bvh** out) {
    return guard([&] {
        need(out != nullptr, "null output handle");
        need(items != nullptr || n == 0, "null primitive array");
        need(prim >= 0 && prim <= 2, "unknown primitive kind");
        need(n >= 2, "at least 2 primitives are required");
        const size_t bytes = static_cast<size_t>(n) * itemSize(prim);
        DevBuf<unsigned char> src(bytes);
        hipCheck(hipMemcpy(src.p, items, bytes, hipMemcpyHostToDevice), "H2D primitives");
        auto* h = new qr_bvh();
        try {
            buildInto(h->s, prim, builder, src.p, n, t);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int qr_bvh_build_device(int prim, int builder, const void* items, int n, qr_build_timings* t,
                        qr_bvh** out) {
    return guard([&] {
        need(out != nullptr, "null output handle");
        need(items != nullptr, "null primitive array");
        auto* h = new qr_bvh();
        try {
            buildInto(h->s, prim, builder, items, n, t);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int qr_bvh_destroy(qr_bvh* bvh) {
    return guard([&] { delete bvh; });
}

int qr_bvh_get_info(const qr_bvh* bvh, qr_bvh_info* out) {
    return guard([&] {
        need(bvh && out, "null argument");
        const BvhState& b = bvh->s;
        out->prim_count = b.n;
        out->node_count = b.nodeCount;
        out->depth = b.depth;
        out->wide_count = b.wideValid ? b.wideCount : 0;
        out->prim_kind = b.kind;
        out->builder = b.builder;
        out->d_nodes = b.nodes.p;
        out->d_items = b.items.p;
        out->d_prim_ids = b.primIds.p;
        out->d_wide_nodes = b.wideValid ? b.wide.p : nullptr;
    });
}

int qr_bvh_update(qr_bvh* bvh, const void* items, int n, const int* deleted, int nDeleted,
                  const int* moved, int nMoved, const int* inserted, int nInserted, int appended,
                  qr_topo_timings* t, double* ms) {
    return guard([&] {
        // Edit validation follows src/gpu/lbvh.cu:2555-2585: every violation is an
        // invalid edit (std::invalid_argument in the reference API).
        if (bvh == nullptr || bvh->s.n < 2) throw EditError("updateBVHGPU: BVH not built");
        BvhState& b = bvh->s;
        const int nOld = b.n;
        if (appended < 0) throw EditError("updateBVHGPU: appended must be >= 0");
        if (nDeleted < 0 || nMoved < 0 || nInserted < 0) throw EditError("updateBVHGPU: negative list size");
        for (int i = 0; i < nDeleted; ++i) {
            const int id = deleted[i];
            if (id < 0 || id >= nOld) throw EditError("updateBVHGPU: deleted out of range");
            if (i > 0 && id <= deleted[i - 1])
                throw EditError("updateBVHGPU: deleted must be ascending and unique");
        }
        const long long nNew = static_cast<long long>(nOld) - nDeleted + nInserted + appended;
        if (nNew < 2) throw EditError("updateBVHGPU: at least 2 primitives required");
        if (static_cast<long long>(n) != nNew)
            throw EditError("updateBVHGPU: scene size does not match the edit");
        const long long tailBeg = nNew - appended;
        for (int i = 0; i < nInserted; ++i) {
            const int id = inserted[i];
            if (id < 0 || id >= tailBeg) throw EditError("updateBVHGPU: inserted index out of range");
            if (i > 0 && id <= inserted[i - 1])
                throw EditError("updateBVHGPU: inserted must be ascending and unique");
        }
        for (int i = 0; i < nMoved; ++i)
            if (moved[i] < 0 || moved[i] >= nNew) throw EditError("updateBVHGPU: moved index out of range");
        need(items != nullptr, "null primitive array");

        qr_topo_timings tt{};
        tt.deleted = nDeleted;
        tt.inserted = nInserted;
        tt.appended = appended;
        tt.new_depth = b.depth;
        if (nDeleted == 0 && nInserted == 0 && appended == 0) {
            // Shape unchanged: the dirty refit path (only the moved primitives travel).
            qr_refit_timings rt{};
            double rms = 0.0;
            const int st = qr_bvh_refit(bvh, items, n, moved, nMoved, &rt, &rms);
            if (st != QR_OK) throw DeviceError(tLastError);
            tt.refit_ms = rms;
            tt.moved = rt.updated;
            tt.total_ms = rms;
        } else {
            // New shape.  Only the new and moved primitives travel; the new scene is
            // composed on the device from the old tree's items, then the topology is
            // rebuilt with the tree's own builder and the 4-wide level re-collapsed
            // when it existed.  Queries on the result equal those on a fresh build of
            // the new scene bit for bit.
            hipStream_t s = currentStream();
            const bool hadWide = b.wideValid;
            const size_t isz = itemSize(b.kind);
            const auto* src = static_cast<const unsigned char*>(items);
            const int nStaged = nInserted + appended;
            std::vector<unsigned char> staged(static_cast<size_t>(nStaged + nMoved) * isz + 1);
            for (int i = 0; i < nInserted; ++i)
                std::memcpy(&staged[static_cast<size_t>(i) * isz], src + static_cast<size_t>(inserted[i]) * isz, isz);
            for (int i = 0; i < appended; ++i)
                std::memcpy(&staged[static_cast<size_t>(nInserted + i) * isz],
                            src + static_cast<size_t>(tailBeg + i) * isz, isz);
            for (int i = 0; i < nMoved; ++i)
                std::memcpy(&staged[static_cast<size_t>(nStaged + i) * isz], src + static_cast<size_t>(moved[i]) * isz,
                            isz);
            std::vector<int> lists;
            lists.insert(lists.end(), deleted, deleted + nDeleted);
            lists.insert(lists.end(), inserted, inserted + nInserted);
            lists.insert(lists.end(), moved, moved + nMoved);
            lists.push_back(0);
            EventSet<5> ev;
            ev.record(0, s);
            // the staged items and index lists share the tree's edit buffer; the new
            // scene is composed into its stage buffer and the tree rebuilt in place,
            // so repeated edits of one tree reuse its device memory
            const size_t listOff = (staged.size() + 255) & ~size_t(255);
            b.editBuf.reserve(listOff + lists.size() * sizeof(int));
            unsigned char* dStaged = b.editBuf.p;
            int* dLists = reinterpret_cast<int*>(b.editBuf.p + listOff);
            hipCheck(hipMemcpyAsync(dStaged, staged.data(), staged.size(), hipMemcpyHostToDevice, s), "H2D edit");
            hipCheck(hipMemcpyAsync(dLists, lists.data(), lists.size() * sizeof(int), hipMemcpyHostToDevice, s),
                     "H2D edit lists");
            ev.record(1, s);
            b.stage.reserve(static_cast<size_t>(n) * isz);
            composeEdit(b, dLists, nDeleted, dLists + nDeleted, nInserted, static_cast<int>(tailBeg), n, dStaged,
                        dLists + nDeleted + nInserted, nMoved, dStaged + static_cast<size_t>(nStaged) * isz, b.stage.p,
                        s);
            ev.record(2, s);
            try {
                buildInto(b, b.kind, b.builder, b.stage.p, n, nullptr);
            } catch (...) {
                b.n = 0;  // the old topology is partly overwritten: the tree is no longer usable
                b.nodeCount = 0;
                throw;
            }
            ev.record(3, s);
            if (hadWide) buildWide(b, s);
            ev.record(4, s);
            ev.sync(4);
            tt.upload_ms = ev.ms(0, 1);
            tt.plan_ms = ev.ms(1, 2);
            tt.surgery_ms = ev.ms(2, 3);
            tt.wide_ms = hadWide ? ev.ms(3, 4) : 0.0;
            tt.total_ms = ev.ms(0, 4);
            tt.moved = nMoved;
            tt.surgery_roots = 1;
            tt.rebuilt_leaves = n;
            tt.new_depth = b.depth;
            tt.full_rebuild = 1;
        }
        if (t) *t = tt;
        if (ms) *ms = tt.total_ms;
    });
}

int qr_rays_device_ptrs(const qr_rays* r, void** rays_dev, uint64_t** results_dev, int* count) {
    return guard([&] {
        need(r != nullptr, "null ray buffer");
        if (rays_dev) *rays_dev = r->rays.p;
        if (results_dev) *results_dev = r->results.p;
        if (count) *count = r->count;
    });
}

int qr_points_device_ptrs(const qr_points* p, float** points_dev, void** results_dev) {
    return guard([&] {
        need(p != nullptr, "null point buffer");
        if (points_dev) *points_dev = reinterpret_cast<float*>(p->pts.p);
        if (results_dev) *results_dev = p->results.p;
    });
}

int qr_radius_device_ptrs(const qr_csr* r, int** offsets_dev, uint64_t** keys_dev) {
    return guard([&] {
        need(r != nullptr, "null radius result");
        if (offsets_dev) *offsets_dev = r->offsets.p;
        if (keys_dev) *keys_dev = r->keys.p;
    });
}

int qr_bvh_build_wide(qr_bvh* bvh, double* ms) {
    return guard([&] {
        need(bvh != nullptr, "null bvh handle");
        need(bvh->s.n >= 2, "bvh is empty");
        hipStream_t s = currentStream();
        // The depth table is part of the collapse cost (the reference keeps it from
        // the build); measure both.  Allocation stays outside the timed region, as
        // in the build.
        reserveWide(bvh->s);
        Timer tm(s);
        buildWide(bvh->s, s);
        const double t = tm.stop();
        if (ms) *ms = t;
    });
}

int qr_bvh_refit(qr_bvh* bvh, const void* items, int n, const int* dirty, int nDirty,
                 qr_refit_timings* t, double* ms) {
    return guard([&] {
        need(bvh != nullptr, "null bvh handle");
        BvhState& b = bvh->s;
        need(n == b.n, "refit: primitive count differs from the tree");
        need(items != nullptr, "null primitive array");
        hipStream_t s = currentStream();
        const size_t isz = itemSize(b.kind);
        qr_refit_timings tt{};
        EventSet<5> ev;
        if (dirty == nullptr) {
            b.stage.reserve(static_cast<size_t>(n) * isz);
            ev.record(0, s);
            hipCheck(hipMemcpyAsync(b.stage.p, items, static_cast<size_t>(n) * isz,
                                    hipMemcpyHostToDevice, s), "H2D refit");
            ev.record(1, s);
            refitFull(b, b.stage.p, s);
            ev.record(2, s);
            ev.record(3, s);
            tt.updated = n;
        } else {
            std::vector<int> ids(dirty, dirty + nDirty);
            std::sort(ids.begin(), ids.end());
            ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
            for (int id : ids) need(id >= 0 && id < n, "refit: dirty index out of range");
            const int m = static_cast<int>(ids.size());
            std::vector<unsigned char> staged(static_cast<size_t>(m) * isz);
            const auto* src = static_cast<const unsigned char*>(items);
            for (int i = 0; i < m; ++i)
                std::memcpy(&staged[static_cast<size_t>(i) * isz], src + static_cast<size_t>(ids[i]) * isz, isz);
            b.stage.reserve(staged.size() + 1);
            b.dirtyIds.reserve(static_cast<size_t>(m) + 1);
            b.dirtyLeaves.reserve(static_cast<size_t>(m) + 1);
            ev.record(0, s);
            if (m > 0) {
                hipCheck(hipMemcpyAsync(b.stage.p, staged.data(), staged.size(), hipMemcpyHostToDevice, s), "H2D dirty");
                hipCheck(hipMemcpyAsync(b.dirtyIds.p, ids.data(), static_cast<size_t>(m) * sizeof(int),
                                        hipMemcpyHostToDevice, s), "H2D dirty ids");
            }
            ev.record(1, s);
            refitDirty(b, b.stage.p, b.dirtyIds.p, m, b.dirtyLeaves.p, s, ev.e[2]);
            if (m <= 0) ev.record(2, s);
            ev.record(3, s);
            tt.updated = m;
        }
        reemitWide(b, s);
        ev.record(4, s);
        ev.sync(4);
        float v[4];
        for (int i = 0; i < 4; ++i) v[i] = ev.ms(i, i + 1);
        const float total = ev.ms(0, 4);
        tt.upload_ms = v[0];
        if (dirty == nullptr) {
            tt.gather_ms = 0.0;
            tt.merge_ms = v[1];
            tt.mark_ms = 0.0;
        } else {
            tt.mark_ms = v[1];
            tt.merge_ms = v[2];
        }
        tt.tree_ms = tt.mark_ms + tt.merge_ms;
        tt.wide_ms = v[3];
        tt.total_ms = total;
        if (t) *t = tt;
        if (ms) *ms = total;
    });
}

int qr_bvh_refit_device(qr_bvh* bvh, const void* items, int n, qr_refit_timings* t, double* ms) {
    return guard([&] {
        need(bvh != nullptr, "null bvh handle");
        BvhState& b = bvh->s;
        need(n == b.n, "refit: primitive count differs from the tree");
        hipStream_t s = currentStream();
        Timer tm(s);
        refitFull(b, items, s);
        reemitWide(b, s);
        const double total = tm.stop();
        if (t) {
            *t = qr_refit_timings{};
            t->merge_ms = total;
            t->tree_ms = total;
            t->total_ms = total;
            t->updated = n;
        }
        if (ms) *ms = total;
    });
}

int qr_bvh_validate(qr_bvh* bvh, const void* items, int n, char* why, int len) {
    return guard([&] {
        need(bvh != nullptr, "null bvh handle");
        BvhState& b = bvh->s;
        std::string msg;
        if (n != b.n) {
            msg = "prim count mismatch";
        } else {
            const size_t isz = itemSize(b.kind);
            DevBuf<unsigned char> src(static_cast<size_t>(n) * isz);
            hipCheck(hipMemcpy(src.p, items, static_cast<size_t>(n) * isz, hipMemcpyHostToDevice), "H2D validate");
            int errors[6] = {0, 0, 0, 0, 0, 0};
            validateTree(b, src.p, errors, currentStream());
            // Root box must equal the scene bounds (union of primitive boxes).
            Node2 root{};
            hipCheck(hipMemcpy(&root, b.nodes.p, sizeof(Node2), hipMemcpyDeviceToHost), "root D2H");
            Box sb = emptyBox();
            const auto* raw = static_cast<const unsigned char*>(items);
            for (int i = 0; i < n; ++i) {
                const void* it = raw + static_cast<size_t>(i) * isz;
                Box pb;
                if (b.kind == kPrimTri) pb = triBounds(*static_cast<const Tri*>(it));
                else if (b.kind == kPrimPoint) pb = PointLeaf::bounds(*static_cast<const V3*>(it));
                else pb = gaussBounds(*static_cast<const Gaussian*>(it));
                growBox(sb, pb);
            }
            const bool badRoot = root.lo[0] != sb.lo.x || root.lo[1] != sb.lo.y || root.lo[2] != sb.lo.z ||
                                 root.hi[0] != sb.hi.x || root.hi[1] != sb.hi.y || root.hi[2] != sb.hi.z;
            const char* names[6] = {"bad_children_or_box", "bad_leaf_box", "bad_refcount",
                                    "bad_leaf_id", "bad_prim_map", "bad_depth"};
            char buf[96];
            for (int i = 0; i < 6; ++i) {
                if (errors[i]) {
                    std::snprintf(buf, sizeof(buf), "%s%s=%d", msg.empty() ? "" : " ", names[i], errors[i]);
                    msg += buf;
                }
            }
            if (badRoot) msg += msg.empty() ? "root_box_mismatch=1" : " root_box_mismatch=1";
        }
        if (why && len > 0) std::snprintf(why, static_cast<size_t>(len), "%s", msg.c_str());
    });
}

int qr_bvh_tree_hash(const qr_bvh* bvh, uint64_t* out) {
    return guard([&] {
        need(bvh && out, "null argument");
        const BvhState& b = bvh->s;
        std::vector<unsigned char> bytes(static_cast<size_t>(b.nodeCount) * sizeof(Node2));
        if (!bytes.empty())
            hipCheck(hipMemcpy(bytes.data(), b.nodes.p, bytes.size(), hipMemcpyDeviceToHost), "tree D2H");
        uint64_t h = 0xcbf29ce484222325ull;
        for (unsigned char c : bytes) {
            h ^= static_cast<uint64_t>(c);
            h *= 0x100000001b3ull;
        }
        *out = h;
    });
}

int qr_bvh_read(const qr_bvh* bvh, void* nodes, int* primIds) {
    return guard([&] {
        need(bvh != nullptr, "null bvh handle");
        const BvhState& b = bvh->s;
        if (nodes)
            hipCheck(hipMemcpy(nodes, b.nodes.p, static_cast<size_t>(b.nodeCount) * sizeof(Node2),
                               hipMemcpyDeviceToHost), "nodes D2H");
        if (primIds)
            hipCheck(hipMemcpy(primIds, b.primIds.p, static_cast<size_t>(b.n) * sizeof(int),
                               hipMemcpyDeviceToHost), "prim ids D2H");
    });
}

// ---- rays ---------------------------------------------------------------------------

int qr_rays_upload(const void* rays, int n, qr_rays** out) {
    return guard([&] {
        need(out != nullptr, "null output handle");
        need(n >= 0, "negative ray count");
        need(rays != nullptr || n == 0, "null ray array");
        auto* r = new qr_rays();
        try {
            r->count = n;
            r->rays.alloc(static_cast<size_t>(n));
            r->results.alloc(static_cast<size_t>(n));
            if (n > 0) {
                hipCheck(hipMemcpy(r->rays.p, rays, static_cast<size_t>(n) * sizeof(Ray), hipMemcpyHostToDevice), "H2D rays");
                uint32_t bb[6];
                hostBounds(static_cast<const float*>(rays), 8, n, bb);
                r->bb.reserve(8);
                hipCheck(hipMemcpy(r->bb.p, bb, sizeof(bb), hipMemcpyHostToDevice), "H2D ray bounds");
                r->bbValid = true;
            }
        } catch (...) {
            delete r;
            throw;
        }
        *out = r;
    });
}

int qr_rays_destroy(qr_rays* r) {
    return guard([&] { delete r; });
}

namespace {
double raycastImpl(const qr_bvh* bvh, const Ray* rays, int n, uint64_t* out, qr_rays* scratch,
                   int collectStats, int order, int wide, int* hitPrim = nullptr) {
    checkTree(bvh, wide != 0);
    need(bvh->s.kind == kPrimTri, "raycast needs a triangle BVH");
    need(order >= 0 && order <= 2, "unknown ray order");
    tRayStats = qr_traversal_stats{};
    if (n <= 0) return 0.0;
    hipStream_t s = currentStream();
    Stats st;
    if (collectStats) zeroStats(st.p, s);
    EventSet<3> ev;
    ev.record(0, s);
    TraverseArgs a = argsFor(bvh->s, wide != 0);
    a.stats = collectStats ? st.p : nullptr;
    if (order != QR_ORDER_INPUT) {
        const float* base = reinterpret_cast<const float*>(rays) + (order == QR_ORDER_MORTON ? 0 : 3);
        const int mode = order == QR_ORDER_MORTON_DIRECTION ? kBoxUnit
                                                            : (scratch->bbValid ? kBoxGiven : kBoxMeasure);
        a.order = reorder(*scratch, base, 8, n, mode, s);
    }
    ev.record(1, s);
    if (hitPrim) {
        launchRaysSave(a, rays, n, out, hitPrim, s);
    } else {
        launchRays(a, rays, n, out, s);
    }
    ev.record(2, s);
    ev.sync(2);
    checkOverflow();
    const double total = ev.ms(0, 2);
    if (collectStats) readStats(st.p, &tRayStats);
    tRayStats.sort_ms = order != QR_ORDER_INPUT ? ev.ms(0, 1) : 0.0;
    tRayStats.kernel_ms = ev.ms(1, 2);
    return total;
}
}  // namespace

int qr_raycast(const qr_bvh* bvh, qr_rays* r, int collectStats, int order, int wide, double* ms) {
    return guard([&] {
        need(r != nullptr, "null ray buffer");
        const double t = raycastImpl(bvh, r->rays.p, r->count, r->results.p, r, collectStats, order, wide);
        if (ms) *ms = t;
    });
}

int qr_raycast_device(const qr_bvh* bvh, const void* rays, int n, uint64_t* results, int order,
                      int wide, double* ms) {
    return guard([&] {
        need(n >= 0, "negative ray count");
        qr_rays scratch;
        const double t = raycastImpl(bvh, static_cast<const Ray*>(rays), n, results, &scratch, 0, order, wide);
        if (ms) *ms = t;
    });
}

int qr_raycast_save_device(const qr_bvh* bvh, const void* rays, int n, uint64_t* results, int* hitPrim,
                           int order, int wide, double* ms) {
    return guard([&] {
        need(n >= 0, "negative ray count");
        need(n == 0 || (rays && results && hitPrim), "null device pointer");
        qr_rays scratch;
        const double t =
            raycastImpl(bvh, static_cast<const Ray*>(rays), n, results, &scratch, 0, order, wide, hitPrim);
        if (ms) *ms = t;
    });
}

int qr_closest_save_aux(qr_bvh* bvh, const float* points, int n, const void* results, int* aux) {
    return guard([&] {
        checkTree(bvh, false);
        need(n >= 0, "negative point count");
        need(n == 0 || (points && results && aux), "null device pointer");
        const int bad = saveClosestAux(bvh->s, points, n, results, aux, currentStream());
        if (bad != 0) throw DeviceError("closest save: saved results disagree with their primitives");
    });
}

int qr_adjoint_closest(qr_bvh* bvh, const float* points, int n, const void* results, const int* aux,
                       const float* grad_dist2, float* grad_points, float* grad_items) {
    return guard([&] {
        checkTree(bvh, false);
        need(n >= 0, "negative point count");
        need(n == 0 || (points && results && grad_dist2), "null device pointer");
        need(n == 0 || bvh->s.kind == kPrimPoint || aux != nullptr, "saved region / root required");
        adjointClosest(bvh->s, points, n, results, aux, grad_dist2, grad_points, grad_items, currentStream());
    });
}

int qr_adjoint_ray(qr_bvh* bvh, const void* rays, int n, const uint64_t* results, const int* hit_prim,
                   const float* grad_t, float* grad_origins, float* grad_dirs, float* grad_tris) {
    return guard([&] {
        checkTree(bvh, false);
        need(n >= 0, "negative ray count");
        need(n == 0 || (rays && results && hit_prim && grad_t), "null device pointer");
        adjointRay(bvh->s, rays, n, results, hit_prim, grad_t, grad_origins, grad_dirs, grad_tris,
                   currentStream());
    });
}

namespace {
void enqueuePlan(qr_batch_plan& p, int variant, bool morton, hipStream_t s) {
    const BatchState& b = p.batch->s;
    launchBatchReset(p.counter.p, s);
    if (morton)
        batchMortonOrder(kPrimPoint, p.points, p.n, p.qScene.p, p.scenes, p.ws, p.cb.p, p.codes.p, p.tmpA.p,
                         p.tmpB.p, p.sceneSorted.p, p.order.p, s);
    BatchQueryArgs a;
    a.kind = b.kind;
    a.nodes = b.nodes.p;
    a.wide = b.wide.p;
    a.wideRoot = b.wideRoot.p;
    a.items = b.items.p;
    a.primIds = b.primIds.p;
    a.primOff = b.primOff.p;
    a.points = p.points;
    a.qScene = p.qScene.p;
    a.order = morton ? p.order.p : nullptr;
    a.count = p.n;
    a.k = p.k < 1 ? 1 : p.k;
    a.outStride = p.stride;
    a.out = p.results.p;
    a.counter = p.counter.p;
    launchBatchNearest(a, variant, s);
}

double timePlan(qr_batch_plan& p, int variant, bool morton) {
    EventSet<2> ev;
    double best = 1e30;
    for (int r = 0; r < 2; ++r) {
        ev.record(0, p.stream);
        enqueuePlan(p, variant, morton, p.stream);
        ev.record(1, p.stream);
        ev.sync(1);
        const double ms = ev.ms(0, 1);
        best = ms < best ? ms : best;
    }
    return best;
}
}  // namespace

int qr_batch_build_device(int prim, const void* items, const int* offsets, int scenes, qr_build_timings* t,
                          qr_batch** out) {
    return guard([&] {
        need(out != nullptr, "null output handle");
        need(prim >= 0 && prim <= 2, "unknown primitive kind");
        need(scenes >= 1 && offsets != nullptr && items != nullptr, "empty batch");
        need(scenes < 65536, "at most 65535 scenes per batch");
        need(offsets[0] == 0, "offsets must start at 0");
        for (int i = 0; i < scenes; ++i)
            need(offsets[i + 1] - offsets[i] >= 2, "every scene needs at least 2 primitives");
        auto* h = new qr_batch();
        try {
            BatchState& b = h->s;
            b.scenes = scenes;
            b.total = offsets[scenes];
            b.primOffH.assign(offsets, offsets + scenes + 1);
            b.primOff.alloc(static_cast<size_t>(scenes) + 1);
            hipCheck(hipMemcpy(b.primOff.p, offsets, sizeof(int) * (static_cast<size_t>(scenes) + 1),
                               hipMemcpyHostToDevice), "batch offsets");
            BuildTimingsC bt{};
            buildBatchDevice(b, prim, items, currentStream(), t ? &bt : nullptr);
            if (t) {
                *t = qr_build_timings{};
                t->morton_ms = bt.mortonMs;
                t->tree_ms = bt.treeMs;
                t->refit_ms = bt.refitMs;
                t->total_ms = bt.totalMs;
            }
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int qr_batch_build(int prim, const void* items, const int* offsets, int scenes, qr_build_timings* t,
                   qr_batch** out) {
    return guard([&] {
        need(prim >= 0 && prim <= 2, "unknown primitive kind");
        need(scenes >= 1 && offsets != nullptr && items != nullptr, "empty batch");
        const size_t bytes = static_cast<size_t>(offsets[scenes]) * itemSize(prim);
        DevBuf<unsigned char> src(bytes);
        hipCheck(hipMemcpy(src.p, items, bytes, hipMemcpyHostToDevice), "H2D batch items");
        const int st = qr_batch_build_device(prim, src.p, offsets, scenes, t, out);
        if (st != QR_OK) throw ArgError(tLastError);
    });
}

int qr_batch_destroy(qr_batch* b) {
    return guard([&] { delete b; });
}

int qr_batch_info(const qr_batch* b, int* scenes, int* total, int* maxDepth) {
    return guard([&] {
        need(b != nullptr, "null batch");
        if (scenes) *scenes = b->s.scenes;
        if (total) *total = b->s.total;
        if (maxDepth) *maxDepth = b->s.maxDepth;
    });
}

int qr_batch_read_scene(const qr_batch* b, int scene, void* nodes, int* primIds) {
    return guard([&] {
        need(b != nullptr, "null batch");
        need(scene >= 0 && scene < b->s.scenes, "scene out of range");
        const int ps = b->s.primOffH[static_cast<size_t>(scene)];
        const int ns = b->s.primOffH[static_cast<size_t>(scene) + 1] - ps;
        if (nodes)
            hipCheck(hipMemcpy(nodes, b->s.nodes.p + (2 * ps - scene), sizeof(Node2) * static_cast<size_t>(2 * ns - 1),
                               hipMemcpyDeviceToHost), "batch nodes D2H");
        if (primIds)
            hipCheck(hipMemcpy(primIds, b->s.primIds.p + ps, sizeof(int) * static_cast<size_t>(ns),
                               hipMemcpyDeviceToHost), "batch ids D2H");
    });
}

int qr_batch_tree_hash(const qr_batch* b, int scene, uint64_t* out) {
    return guard([&] {
        need(b != nullptr && out != nullptr, "null argument");
        need(scene >= 0 && scene < b->s.scenes, "scene out of range");
        const int ps = b->s.primOffH[static_cast<size_t>(scene)];
        const int ns = b->s.primOffH[static_cast<size_t>(scene) + 1] - ps;
        std::vector<unsigned char> bytes(sizeof(Node2) * static_cast<size_t>(2 * ns - 1));
        hipCheck(hipMemcpy(bytes.data(), b->s.nodes.p + (2 * ps - scene), bytes.size(), hipMemcpyDeviceToHost),
                 "batch nodes D2H");
        uint64_t h = 0xcbf29ce484222325ull;
        for (unsigned char c : bytes) {
            h ^= static_cast<uint64_t>(c);
            h *= 0x100000001b3ull;
        }
        *out = h;
    });
}

int qr_batch_plan_create(qr_batch* b, const float* points, const int* qoffsets, int k, int tune,
                         qr_batch_plan** out) {
    return guard([&] {
        need(b != nullptr && out != nullptr && qoffsets != nullptr, "null argument");
        need(k >= 0 && k <= 32, "k must be in [0, 32] (0: closest)");
        const int S = b->s.scenes;
        need(qoffsets[0] == 0, "query offsets must start at 0");
        for (int i = 0; i < S; ++i) need(qoffsets[i + 1] >= qoffsets[i], "query offsets must not decrease");
        const int n = qoffsets[S];
        need(n == 0 || points != nullptr, "null points");
        armOverflowWord();
        auto* p = new qr_batch_plan();
        try {
            p->batch = b;
            p->points = reinterpret_cast<const V3*>(points);
            p->n = n;
            p->k = k;
            p->stride = k < 1 ? 1 : k;
            p->scenes = S;
            p->qOff.alloc(static_cast<size_t>(S) + 1);
            hipCheck(hipMemcpy(p->qOff.p, qoffsets, sizeof(int) * (static_cast<size_t>(S) + 1), hipMemcpyHostToDevice),
                     "query offsets");
            p->qScene.alloc(static_cast<size_t>(n));
            p->counter.alloc(1);
            p->results.alloc(static_cast<size_t>(n) * p->stride);
            p->cb.alloc(static_cast<size_t>(S) * 6);
            for (DevBuf<uint32_t>* d : {&p->codes, &p->tmpA, &p->tmpB, &p->sceneSorted, &p->order})
                d->alloc(static_cast<size_t>(n));
            p->ws.reserve(static_cast<size_t>(n));
            hipCheck(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking), "plan stream");
            batchSceneIds(p->qOff.p, S, p->qScene.p, p->stream);
            // warm every candidate once (occupancy, code objects), then choose
            int bestV = kBatchLds256;
            bool bestM = true;
            double best = 1e30;
            if (tune >= 2) {
                need(tune < 2 + 2 * kBatchVariants, "unknown forced batch configuration");
                bestV = (tune - 2) / 2;
                bestM = ((tune - 2) % 2) != 0;
                if (batchVariantWide(bestV)) buildBatchWide(b->s, p->stream);
                best = timePlan(*p, bestV, bestM);
            } else if (tune == 1) {
                buildBatchWide(b->s, p->stream);  // the 4-wide variants are candidates too
                for (int v = 0; v < kBatchVariants; ++v)
                    for (int m = 0; m < 2; ++m) {
                        const double ms = timePlan(*p, v, m != 0);
                        if (ms < best) {
                            best = ms;
                            bestV = v;
                            bestM = m != 0;
                        }
                    }
            } else {
                best = timePlan(*p, bestV, bestM);
            }
            p->variant = bestV;
            p->morton = bestM;
            p->tunedMs = best;
            hipCheck(hipStreamSynchronize(p->stream), "plan sync");
            hipCheck(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal), "begin capture");
            enqueuePlan(*p, p->variant, p->morton, p->stream);
            hipCheck(hipStreamEndCapture(p->stream, &p->graph), "end capture");
            hipCheck(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0), "graph instantiate");
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
    });
}

int qr_batch_plan_run(qr_batch_plan* p, double* ms) {
    return guard([&] {
        need(p != nullptr && p->exec != nullptr, "null plan");
        EventSet<2> ev;
        ev.record(0, p->stream);
        hipCheck(hipGraphLaunch(p->exec, p->stream), "graph launch");
        ev.record(1, p->stream);
        ev.sync(1);
        checkOverflow();
        if (ms) *ms = ev.ms(0, 1);
    });
}

int qr_batch_plan_run_eager(qr_batch_plan* p, double* ms) {
    return guard([&] {
        need(p != nullptr, "null plan");
        EventSet<2> ev;
        ev.record(0, p->stream);
        enqueuePlan(*p, p->variant, p->morton, p->stream);
        ev.record(1, p->stream);
        ev.sync(1);
        checkOverflow();
        if (ms) *ms = ev.ms(0, 1);
    });
}

int qr_batch_plan_read(const qr_batch_plan* p, void* host) {
    return guard([&] {
        need(p != nullptr && host != nullptr, "null argument");
        if (p->n > 0)
            hipCheck(hipMemcpy(host, p->results.p, sizeof(ClosestResult) * static_cast<size_t>(p->n) * p->stride,
                               hipMemcpyDeviceToHost), "plan results D2H");
    });
}

int qr_batch_plan_config(const qr_batch_plan* p, int* variant, int* morton, double* tunedMs) {
    return guard([&] {
        need(p != nullptr, "null plan");
        if (variant) *variant = p->variant;
        if (morton) *morton = p->morton ? 1 : 0;
        if (tunedMs) *tunedMs = p->tunedMs;
    });
}

int qr_batch_plan_destroy(qr_batch_plan* p) {
    return guard([&] { delete p; });
}

int qr_rays_summarize(const qr_rays* r, uint64_t* perRay, qr_ray_stats* out) {
    return guard([&] {
        need(r && out, "null argument");
        std::vector<uint64_t> host(static_cast<size_t>(r->count));
        if (r->count > 0)
            hipCheck(hipMemcpy(host.data(), r->results.p, host.size() * sizeof(uint64_t), hipMemcpyDeviceToHost), "D2H results");
        qr_ray_stats st{0, 0, 0xcbf29ce484222325ull, 0.0};
        for (uint64_t v : host) {
            ++st.rays;
            if (v >> 32) {
                ++st.hits;
                st.sum_t += static_cast<double>(bitsFloat(static_cast<uint32_t>(v)));
            }
            st.hash ^= v;
            st.hash *= 0x100000001b3ull;
        }
        if (perRay && !host.empty()) std::memcpy(perRay, host.data(), host.size() * sizeof(uint64_t));
        *out = st;
    });
}

int qr_last_traversal_stats(qr_traversal_stats* out) {
    if (!out) return QR_ERR_INVALID;
    *out = tRayStats;
    return QR_OK;
}

// ---- point queries ------------------------------------------------------------------------

int qr_points_upload(const float* pts, int n, int rpp, qr_points** out) {
    return guard([&] {
        need(out != nullptr, "null output handle");
        need(n >= 0, "negative point count");
        need(pts != nullptr || n == 0, "null point array");
        auto* p = new qr_points();
        try {
            p->count = n;
            p->stride = rpp > 0 ? rpp : 1;
            p->pts.alloc(static_cast<size_t>(n));
            p->results.alloc(static_cast<size_t>(n) * p->stride);
            if (n > 0) {
                hipCheck(hipMemcpy(p->pts.p, pts, static_cast<size_t>(n) * sizeof(V3), hipMemcpyHostToDevice), "H2D points");
                uint32_t bb[6];
                hostBounds(pts, 3, n, bb);
                p->bb.reserve(8);
                hipCheck(hipMemcpy(p->bb.p, bb, sizeof(bb), hipMemcpyHostToDevice), "H2D point bounds");
                p->bbValid = true;
            }
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
    });
}

int qr_points_destroy(qr_points* p) {
    return guard([&] { delete p; });
}

int qr_points_info(const qr_points* p, int* count, int* stride) {
    if (!p) return QR_ERR_INVALID;
    if (count) *count = p->count;
    if (stride) *stride = p->stride;
    return QR_OK;
}

namespace {

// Shared body of closest and kNN (k = 0 selects closest).
double pointQuery(const qr_bvh* bvh, const V3* pts, int n, int k, int outStride, ClosestResult* out,
                  qr_points* work, int collectStats, int order, int wide) {
    checkTree(bvh, wide != 0);
    need(order == QR_ORDER_INPUT || order == QR_ORDER_MORTON, "query order must be input or morton");
    tQueryStats = qr_traversal_stats{};
    if (n <= 0) return 0.0;
    hipStream_t s = currentStream();
    Stats st;
    if (collectStats) zeroStats(st.p, s);
    EventSet<3> ev;
    ev.record(0, s);
    TraverseArgs a = argsFor(bvh->s, wide != 0);
    a.stats = collectStats ? st.p : nullptr;
    if (order == QR_ORDER_MORTON)
        a.order = reorder(*work, reinterpret_cast<const float*>(pts), 3, n,
                          work->bbValid ? kBoxGiven : kBoxMeasure, s);
    ev.record(1, s);
    if (k == 0) {
        launchClosest(a, pts, n, out, s);
    } else {
        launchKnn(a, pts, n, k, outStride, out, s);
    }
    ev.record(2, s);
    ev.sync(2);
    checkOverflow();
    const double total = ev.ms(0, 2);
    if (collectStats) readStats(st.p, &tQueryStats);
    tQueryStats.sort_ms = order == QR_ORDER_MORTON ? ev.ms(0, 1) : 0.0;
    tQueryStats.kernel_ms = ev.ms(1, 2);
    return total;
}

}  // namespace

int qr_closest(const qr_bvh* bvh, qr_points* p, int collectStats, int order, int wide, double* ms) {
    return guard([&] {
        need(p != nullptr, "null point buffer");
        const double t = pointQuery(bvh, p->pts.p, p->count, 0, p->stride, p->results.p, p, collectStats, order, wide);
        if (ms) *ms = t;
    });
}

int qr_knn(const qr_bvh* bvh, qr_points* p, int k, int collectStats, int order, int wide, double* ms) {
    return guard([&] {
        need(p != nullptr, "null point buffer");
        need(k >= 1 && k <= 32, "kNearestGPU: k must be in [1, 32]");
        need(p->stride >= k, "kNearestGPU: result buffer was uploaded with a smaller stride");
        const double t = pointQuery(bvh, p->pts.p, p->count, k, p->stride, p->results.p, p, collectStats, order, wide);
        if (ms) *ms = t;
    });
}

int qr_closest_device(const qr_bvh* bvh, const float* pts, int n, void* results, int order, int wide,
                      double* ms) {
    return guard([&] {
        qr_points work;
        const double t = pointQuery(bvh, reinterpret_cast<const V3*>(pts), n, 0, 1,
                                    static_cast<ClosestResult*>(results), &work, 0, order, wide);
        if (ms) *ms = t;
    });
}

int qr_knn_device(const qr_bvh* bvh, const float* pts, int n, int k, void* results, int order,
                  int wide, double* ms) {
    return guard([&] {
        need(k >= 1 && k <= 32, "kNearestGPU: k must be in [1, 32]");
        qr_points work;
        const double t = pointQuery(bvh, reinterpret_cast<const V3*>(pts), n, k, k,
                                    static_cast<ClosestResult*>(results), &work, 0, order, wide);
        if (ms) *ms = t;
    });
}

// ---- kNN without allocation ---------------------------------------------------------
// The scratch of qr_knn_device's Morton order (codes, sorted codes, visiting order,
// query bounds, and the radix sort's ping-pong buffers, digit histograms, look-back
// descriptors and tile tickets) lives in one fixed device array, so the call makes no
// device allocation.  Each GPU holds its own copy until the process ends.  Every call
// on a GPU shares that GPU's copy: calls that overlap on one GPU race.
namespace {

constexpr size_t kNaQueries = QR_NOALLOC_MAX_QUERIES;
constexpr size_t kNaKeys = 0;
constexpr size_t kNaSorted = kNaKeys + kNaQueries;
constexpr size_t kNaOrder = kNaSorted + kNaQueries;
constexpr size_t kNaBounds = kNaOrder + kNaQueries;          // 6 words used, 64 kept
constexpr size_t kNaAltKeys = kNaBounds + 64;
constexpr size_t kNaAltVals = kNaAltKeys + kNaQueries;
constexpr size_t kNaHist = kNaAltVals + kNaQueries;
// descriptors for 4 passes at the smallest sort tile, as SortWorkspace::reserve sizes them
constexpr size_t kNaDesc = kNaHist + static_cast<size_t>(kHistCopies) * kHistCopyWords;
constexpr size_t kNaTickets = kNaDesc + 4 * ((kNaQueries + kSortTileSmall - 1) / kSortTileSmall) * 256;
constexpr size_t kNaWords = kNaTickets + 64;
static_assert(kNaQueries % 64 == 0, "every part stays 256-byte aligned");

__device__ __align__(256) uint32_t scratch[kNaWords];

constexpr int kNaDevices = 64;
std::mutex gScratchMutex;
uint32_t* gScratchDev[kNaDevices] = {};

// The current device's address of `scratch`, looked up once per device with that
// device current (the array's host-side name is not a device address).
uint32_t* scratchOnDevice() {
    int dev = 0;
    hipCheck(hipGetDevice(&dev), "get device");
    need(dev >= 0 && dev < kNaDevices, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (!gScratchDev[dev]) {
        void* ptr = nullptr;
        hipCheck(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "scratch address");
        gScratchDev[dev] = static_cast<uint32_t*>(ptr);
    }
    return gScratchDev[dev];
}

}  // namespace

int qr_knn_device_noalloc(const qr_bvh* bvh, const float* pts, int n, int k, void* results, int order,
                          int wide) {
    return guard([&] {
        need(k >= 1 && k <= 32, "kNearestGPU: k must be in [1, 32]");
        need(n >= 0 && static_cast<size_t>(n) <= kNaQueries,
             "qr_knn_device_noalloc: more queries than its fixed scratch holds (QR_NOALLOC_MAX_QUERIES)");
        checkTree(bvh, wide != 0);
        need(order == QR_ORDER_INPUT || order == QR_ORDER_MORTON, "query order must be input or morton");
        tQueryStats = qr_traversal_stats{};
        if (n == 0) return;
        hipStream_t s = currentStream();
        TraverseArgs a = argsFor(bvh->s, wide != 0, false);
        a.tune = false;  // timing trials would take events; shapes give the same bits
        if (order == QR_ORDER_MORTON) {
            uint32_t* w = scratchOnDevice();
            SortWorkspace ws;  // views into the fixed array, never freed
            ws.keysAlt.borrow(w + kNaAltKeys, kNaQueries);
            ws.valsAlt.borrow(w + kNaAltVals, kNaQueries);
            ws.hist.borrow(w + kNaHist, static_cast<size_t>(kHistCopies) * kHistCopyWords);
            ws.desc.borrow(w + kNaDesc, kNaTickets - kNaDesc);
            ws.tickets.borrow(w + kNaTickets, 4);
            ws.capacity = kNaQueries;
            mortonOrder(pts, 3, n, kBoxMeasure, ws, w + kNaKeys, w + kNaSorted, w + kNaOrder, w + kNaBounds, s);
            a.order = w + kNaOrder;
        }
        launchKnn(a, reinterpret_cast<const V3*>(pts), n, k, k, static_cast<ClosestResult*>(results), s);
        hipCheck(hipStreamSynchronize(s), "knn");
        checkOverflow();  // reported once another call has armed the device's word
    });
}

int qr_radius(const qr_bvh* bvh, qr_points* p, float radius, int collectStats, int order, int wide,
              qr_csr** out, double* ms) {
    return guard([&] {
        need(p != nullptr && out != nullptr, "null argument");
        checkTree(bvh, wide != 0);
        need(order == QR_ORDER_INPUT || order == QR_ORDER_MORTON, "query order must be input or morton");
        tQueryStats = qr_traversal_stats{};
        auto* res = new qr_csr();
        try {
            const int n = p->count;
            res->count = n;
            hipStream_t s = currentStream();
            res->offsets.alloc(static_cast<size_t>(n) + 1, s);
            if (n <= 0 || radius < 0.0f) {
                // Empty CSR: every segment is empty (the reference returns no hits for r < 0).
                std::vector<int> zeros(static_cast<size_t>(n) + 1, 0);
                hipCheck(hipMemcpyAsync(res->offsets.p, zeros.data(), zeros.size() * sizeof(int), hipMemcpyHostToDevice, s),
                         "H2D offsets");
                hipCheck(hipStreamSynchronize(s), "H2D offsets");
                *out = res;
                if (ms) *ms = 0.0;
                return;
            }
            Stats st;
            if (collectStats) zeroStats(st.p, s);
            const float r2 = radius * radius;
            EventSet<3> ev;
            EventSet<4> kev;
            ev.record(0, s);
            TraverseArgs a = argsFor(bvh->s, wide != 0);
            a.stats = collectStats ? st.p : nullptr;
            if (order == QR_ORDER_MORTON)
                a.order = reorder(*p, reinterpret_cast<const float*>(p->pts.p), 3, n,
                                  p->bbValid ? kBoxGiven : kBoxMeasure, s);
            ev.record(1, s);
            p->rCounts.reserve(static_cast<size_t>(n));
            p->rScratch.reserve(static_cast<size_t>(std::max(scanScratchInts(n), 4 + 2 * n)));
            p->rMax.reserve(kMaxCopies);
            const int cap = radiusCaptureCap(n);
            if (cap > 0) p->rCapture.reserve(static_cast<size_t>(n) * cap);
            int* counts = p->rCounts.p;
            int* scratch = p->rScratch.p;
            zeroInts(p->rMax.p, kMaxCopies, s);
            kev.record(0, s);
            const bool captured = launchRadiusCapture(a, p->pts.p, n, r2, counts, p->rCapture.p, cap, p->rMax.p, s);
            kev.record(1, s);
            scanExclusive(counts, res->offsets.p, n, res->offsets.p + n, scratch, s);
            int* hw = hostWords();
            hw[0] = hw[1] = -1;
            radiusSummary(res->offsets.p + n, p->rMax.p, hostWordsDevice(), s);
            // Compaction before the total is known on the host, into keys sized by the
            // last call on these queries: the common case then needs no host round
            // trip in the middle of the call.  A short guess skips the segments past it
            // and the compaction is redone below at the exact size.
            const int guess = captured ? p->rKeysGuess : 0;
            if (guess > 0) {
                res->keys.alloc(static_cast<size_t>(guess), s);
                radiusCompact(counts, res->offsets.p, p->rCapture.p, n, cap, res->keys.p, guess, s);
            }
            hipCheck(hipStreamSynchronize(s), "radius total sync");
            const int total = hw[0];
            const int maxSeg = hw[1];
            if (total < 0 || maxSeg < 0) throw DeviceError("radius: summary words not written");
            res->total = total;
            const bool compacted = guess > 0 && total <= guess;
            if (!compacted) res->keys.alloc(static_cast<size_t>(total > 0 ? total : 1), s);
            p->rKeysGuess = total + total / 8 + 64;
            kev.record(2, s);
            if (total > 0) {
                a.stats = nullptr;
                if (captured && !compacted)
                    radiusCompact(counts, res->offsets.p, p->rCapture.p, n, cap, res->keys.p, total, s);
                // a second walk only for the queries the capture could not hold
                if (!captured || maxSeg > cap)
                    launchRadiusFill(a, p->pts.p, n, r2, counts, captured ? cap : 0, res->offsets.p, res->keys.p, s);
            }
            kev.record(3, s);
            // segments of at most 64 captured hits come sorted from the one-wave capture
            const bool sorted = captured && maxSeg <= 64 && maxSeg <= cap && radiusCapturePresorts(a, n);
            if (total > 0 && !sorted) segmentSort(res->offsets.p, res->keys.p, n, maxSeg, scratch, s);
            ev.record(2, s);
            ev.sync(2);
            checkOverflow();
            const double total_ms = ev.ms(0, 2);
            if (collectStats) readStats(st.p, &tQueryStats);
            tQueryStats.sort_ms = order == QR_ORDER_MORTON ? ev.ms(0, 1) : 0.0;
            tQueryStats.kernel_ms = kev.ms(0, 1) + kev.ms(2, 3);
            if (ms) *ms = total_ms;
        } catch (...) {
            delete res;
            throw;
        }
        *out = res;
    });
}

int qr_points_read(const qr_points* p, void* host) {
    return guard([&] {
        need(p != nullptr && host != nullptr, "null argument");
        const size_t bytes = static_cast<size_t>(p->count) * p->stride * sizeof(ClosestResult);
        if (bytes) hipCheck(hipMemcpy(host, p->results.p, bytes, hipMemcpyDeviceToHost), "D2H results");
    });
}

int qr_radius_info(const qr_csr* r, int* count, int* total) {
    if (!r) return QR_ERR_INVALID;
    if (count) *count = r->count;
    if (total) *total = r->total;
    return QR_OK;
}

int qr_radius_read(const qr_csr* r, int* offsets, uint64_t* keys) {
    return guard([&] {
        need(r != nullptr, "null radius buffer");
        if (offsets)
            hipCheck(hipMemcpy(offsets, r->offsets.p, (static_cast<size_t>(r->count) + 1) * sizeof(int), hipMemcpyDeviceToHost), "D2H offsets");
        if (keys && r->total > 0)
            hipCheck(hipMemcpy(keys, r->keys.p, static_cast<size_t>(r->total) * sizeof(uint64_t), hipMemcpyDeviceToHost), "D2H keys");
    });
}

int qr_radius_destroy(qr_csr* r) {
    return guard([&] { delete r; });
}

int qr_last_query_stats(qr_traversal_stats* out) {
    if (!out) return QR_ERR_INVALID;
    *out = tQueryStats;
    return QR_OK;
}

}  // extern "C"
