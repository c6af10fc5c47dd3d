// gpu/gpu_bvh.h - host C++ API of the quarry GPU backend, MI355X (gfx950) build.
//
// Drop-in for src/gpu/gpu_bvh.h of ycl-berserk/quarry-gpu: the same types,
// function names, argument meanings, defaults and exception types, so a program
// written against the reference (its examples/ and bench/gpu_*_bench.cpp) compiles
// unchanged against this header and links against libquarry_hip.so instead of the
// CUDA library.  Every function is an inline wrapper over the C ABI declared in
// <quarry/quarry_abi.h>; no HIP type appears here.
//
// The scene types (Scene, PointCloud, GaussianCloud, Ray, Vec3, AABB, ClosestHit,
// RayStats) come from the CPU core library (core/*.h), which this backend does not
// replace.
//
// Error behaviour (reference src/gpu/lbvh.cu:28-32, query.cu:969-974,
// lbvh.cu:2555-2585): device failures and invalid query parameters throw
// std::runtime_error; invalid topology edits throw std::invalid_argument.
//
// Handles are value structs holding an opaque library handle next to the public
// fields of the reference (primCount, nodeCount, depth, wideCount, device
// pointers).  As in the reference, a handle must not be copied or shared between
// threads while in use, and every upload/build needs its matching destroy.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "core/fingerprint.h"
#include "core/gaussian.h"
#include "core/geometry.h"
#include "core/point_cloud.h"
#include "core/scene.h"
#include "quarry/quarry_abi.h"

namespace quarry_gpu_detail {

inline void check(int status, const char* what) {
    if (status == QR_OK) return;
    std::string msg = std::string(what) + ": " + qr_last_error();
    if (status == QR_ERR_EDIT) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

}  // namespace quarry_gpu_detail

// GpuBVH (src/gpu/gpu_bvh.h:26-78).  Internal nodes are 0..n-2, leaf n-1+slot.
struct GpuBVH {
    void* d_nodes = nullptr;    // 32 B nodes, nodeCount of them
    void* d_tris = nullptr;     // primitives in leaf order (Tri, Vec3 or Gaussian)
    void* d_primIds = nullptr;  // int per leaf slot: scene index
    int primCount = 0;
    int nodeCount = 0;
    int depth = 0;
    void* d_wideNodes = nullptr;  // 128 B 4-wide nodes after buildWideNodesGPU
    int wideCount = 0;
    bool sah = false;
    qr_bvh* handle = nullptr;   // library state (refit caches, wide arrays, scratch)
};

// GpuBuildTimings (src/gpu/gpu_bvh.h:79-89)
struct GpuBuildTimings {
    double mortonMs = 0.0;
    double sortMs = 0.0;
    double treeMs = 0.0;
    double scatterMs = 0.0;
    double refitMs = 0.0;
    double sahBinMs = 0.0;
    double sahSplitMs = 0.0;
    double sahPartMs = 0.0;
    double totalMs = 0.0;
};

// GpuRayBuffer (src/gpu/gpu_bvh.h:91-96): rays and results stay in device memory.
struct GpuRayBuffer {
    void* d_rays = nullptr;
    void* d_results = nullptr;  // uint64_t per ray: (hit << 32) | t bits
    int count = 0;
    void* d_sortScratch = nullptr;  // unused here: reordering scratch lives in the handle
    qr_rays* handle = nullptr;
};

// GpuRefitTimings (src/gpu/gpu_bvh.h:129-138)
struct GpuRefitTimings {
    double uploadMs = 0.0;
    double gatherMs = 0.0;
    double treeMs = 0.0;
    double markMs = 0.0;
    double mergeMs = 0.0;
    double wideMs = 0.0;
    double totalMs = 0.0;
    int updated = 0;
};

// GpuTopoEdit / GpuTopoUpdateTimings (src/gpu/gpu_bvh.h:182-206)
struct GpuTopoEdit {
    std::vector<int> deleted;
    std::vector<int> moved;
    std::vector<int> inserted;
    int appended = 0;
};

// The fields keep the reference's names; an edit that changes the primitive count is
// served by a full rebuild of the device-composed scene here (fullRebuild = true,
// surgeryRoots = 1, rebuiltLeaves = n, planMs = composition, surgeryMs = rebuild,
// relayoutMs = mergeMs = 0); a pure-move edit is a dirty refit (refitMs).
struct GpuTopoUpdateTimings {
    double uploadMs = 0.0;
    double planMs = 0.0;
    double surgeryMs = 0.0;
    double relayoutMs = 0.0;
    double mergeMs = 0.0;
    double wideMs = 0.0;
    double totalMs = 0.0;
    int deleted = 0;
    int moved = 0;
    int inserted = 0;
    int appended = 0;
    int surgeryRoots = 0;
    int rebuiltLeaves = 0;
    int newDepth = 0;
    bool fullRebuild = false;
    double refitMs = 0.0;
};

// RayOrder (src/gpu/gpu_bvh.h:224)
enum class RayOrder { Input, Morton, MortonDirection };

// GpuTraversalStats (src/gpu/gpu_bvh.h:240-245)
struct GpuTraversalStats {
    uint64_t nodeTests = 0;
    uint64_t triTests = 0;
    uint64_t stackOverflows = 0;
    double sortMs = 0.0;
};

// GpuClosestResult (src/gpu/gpu_bvh.h:250-253)
struct GpuClosestResult {
    uint32_t dist2Bits = 0;
    int32_t triangle = -1;
};
static_assert(sizeof(GpuClosestResult) == 8, "GpuClosestResult is 8 bytes");

// GpuPointBuffer (src/gpu/gpu_bvh.h:255-262)
struct GpuPointBuffer {
    void* d_points = nullptr;
    void* d_results = nullptr;
    int count = 0;
    int stride = 1;
    AABB bounds;
    void* d_sortScratch = nullptr;  // unused here: reordering scratch lives in the handle
    qr_points* handle = nullptr;
};

// GpuQueryStats (src/gpu/gpu_bvh.h:264-269)
struct GpuQueryStats {
    uint64_t nodeTests = 0;
    uint64_t triTests = 0;
    uint64_t stackOverflows = 0;
    double sortMs = 0.0;
};

// QueryOrder (src/gpu/gpu_bvh.h:274)
enum class QueryOrder { Input, Morton };

// GpuRadiusBuffer (src/gpu/gpu_bvh.h:296-301)
struct GpuRadiusBuffer {
    void* d_offsets = nullptr;
    void* d_keys = nullptr;
    int count = 0;
    int total = 0;
    qr_csr* handle = nullptr;
};

namespace quarry_gpu_detail {

inline void sync(GpuBVH& b) {
    qr_bvh_info info{};
    check(qr_bvh_get_info(b.handle, &info), "bvh info");
    b.d_nodes = info.d_nodes;
    b.d_tris = info.d_items;
    b.d_primIds = info.d_prim_ids;
    b.primCount = info.prim_count;
    b.nodeCount = info.node_count;
    b.depth = info.depth;
    b.d_wideNodes = info.d_wide_nodes;
    b.wideCount = info.wide_count;
    b.sah = info.builder == QR_BUILDER_SAH;
}

inline void copyTimings(const qr_build_timings& t, GpuBuildTimings* out) {
    if (!out) return;
    out->mortonMs = t.morton_ms;
    out->sortMs = t.sort_ms;
    out->treeMs = t.tree_ms;
    out->scatterMs = t.scatter_ms;
    out->refitMs = t.refit_ms;
    out->sahBinMs = t.sah_bin_ms;
    out->sahSplitMs = t.sah_split_ms;
    out->sahPartMs = t.sah_part_ms;
    out->totalMs = t.total_ms;
}

inline GpuBVH build(int prim, int builder, const void* items, size_t n, GpuBuildTimings* timings,
                    const char* what) {
    if (n < 2) throw std::runtime_error(std::string(what) + ": at least 2 primitives required");
    qr_build_timings t{};
    GpuBVH b;
    check(qr_bvh_build(prim, builder, items, static_cast<int>(n), &t, &b.handle), what);
    sync(b);
    copyTimings(t, timings);
    return b;
}

inline void requireKind(const GpuBVH& b, int kind, const char* what) {
    if (b.handle == nullptr) throw std::runtime_error(std::string(what) + ": BVH not built");
    qr_bvh_info info{};
    check(qr_bvh_get_info(b.handle, &info), what);
    if (info.prim_kind != kind)
        throw std::runtime_error(std::string(what) + ": the tree holds another primitive kind");
}

inline double refit(GpuBVH& b, int kind, const void* items, size_t n, const std::vector<int>* dirty,
                    GpuRefitTimings* timings, const char* what) {
    requireKind(b, kind, what);
    qr_refit_timings t{};
    double ms = 0.0;
    check(qr_bvh_refit(b.handle, items, static_cast<int>(n), dirty ? dirty->data() : nullptr,
                       dirty ? static_cast<int>(dirty->size()) : 0, &t, &ms),
          what);
    sync(b);
    if (timings) {
        timings->uploadMs = t.upload_ms;
        timings->gatherMs = t.gather_ms;
        timings->treeMs = t.tree_ms;
        timings->markMs = t.mark_ms;
        timings->mergeMs = t.merge_ms;
        timings->wideMs = t.wide_ms;
        timings->totalMs = t.total_ms;
        timings->updated = t.updated;
    }
    return ms;
}

inline double update(GpuBVH& b, int kind, const void* items, size_t n, const GpuTopoEdit& e,
                     GpuTopoUpdateTimings* timings) {
    if (b.handle == nullptr) throw std::invalid_argument("updateBVHGPU: BVH not built");
    requireKind(b, kind, "updateBVHGPU");
    qr_topo_timings t{};
    double ms = 0.0;
    check(qr_bvh_update(b.handle, items, static_cast<int>(n), e.deleted.data(),
                        static_cast<int>(e.deleted.size()), e.moved.data(),
                        static_cast<int>(e.moved.size()), e.inserted.data(),
                        static_cast<int>(e.inserted.size()), e.appended, &t, &ms),
          "updateBVHGPU");
    sync(b);
    if (timings) {
        timings->uploadMs = t.upload_ms;
        timings->planMs = t.plan_ms;
        timings->surgeryMs = t.surgery_ms;
        timings->relayoutMs = t.relayout_ms;
        timings->mergeMs = t.merge_ms;
        timings->wideMs = t.wide_ms;
        timings->totalMs = t.total_ms;
        timings->deleted = t.deleted;
        timings->moved = t.moved;
        timings->inserted = t.inserted;
        timings->appended = t.appended;
        timings->surgeryRoots = t.surgery_roots;
        timings->rebuiltLeaves = t.rebuilt_leaves;
        timings->newDepth = t.new_depth;
        timings->fullRebuild = t.full_rebuild != 0;
        timings->refitMs = t.refit_ms;
    }
    return ms;
}

inline std::string validate(const GpuBVH& b, int kind, const void* items, size_t n) {
    requireKind(b, kind, "validateBVHGPU");
    char why[512] = {0};
    check(qr_bvh_validate(b.handle, items, static_cast<int>(n), why, sizeof(why)), "validateBVHGPU");
    return why;
}

inline double closest(const GpuBVH& b, int kind, GpuPointBuffer& buf, bool stats, QueryOrder order,
                      bool wide, const char* what) {
    if (buf.count <= 0) return 0.0;
    requireKind(b, kind, what);
    double ms = 0.0;
    check(qr_closest(b.handle, buf.handle, stats ? 1 : 0, static_cast<int>(order), wide ? 1 : 0, &ms),
          what);
    return ms;
}

inline double knn(const GpuBVH& b, int kind, GpuPointBuffer& buf, int k, bool stats, QueryOrder order,
                  bool wide, const char* what) {
    if (buf.count <= 0) return 0.0;
    requireKind(b, kind, what);
    double ms = 0.0;
    check(qr_knn(b.handle, buf.handle, k, stats ? 1 : 0, static_cast<int>(order), wide ? 1 : 0, &ms),
          what);
    return ms;
}

inline void releaseRadius(GpuRadiusBuffer& out) {
    if (out.handle) qr_radius_destroy(out.handle);
    out = GpuRadiusBuffer{};
}

inline double radius(const GpuBVH& b, int kind, GpuPointBuffer& buf, float r, GpuRadiusBuffer& out,
                     bool stats, QueryOrder order, bool wide, const char* what) {
    releaseRadius(out);  // out is reallocated on every call (src/gpu/gpu_bvh.h:305-310)
    if (buf.count <= 0) return 0.0;
    requireKind(b, kind, what);
    double ms = 0.0;
    check(qr_radius(b.handle, buf.handle, r, stats ? 1 : 0, static_cast<int>(order), wide ? 1 : 0,
                    &out.handle, &ms),
          what);
    int count = 0, total = 0, *offs = nullptr;
    uint64_t* keys = nullptr;
    check(qr_radius_info(out.handle, &count, &total), what);
    check(qr_radius_device_ptrs(out.handle, &offs, &keys), what);
    out.count = count;
    out.total = total;
    out.d_offsets = offs;
    out.d_keys = keys;
    return ms;
}

}  // namespace quarry_gpu_detail

// ---- device --------------------------------------------------------------------------

// gpuDescribe (src/gpu/gpu_bvh.h:100): device name, or false with the reason.
inline bool gpuDescribe(std::string* why) {
    char buf[256] = {0};
    const int st = qr_describe(buf, sizeof(buf));
    if (why) *why = st == QR_OK ? std::string(buf) : std::string(qr_last_error());
    return st == QR_OK;
}

// warmupBuildGPU (src/gpu/gpu_bvh.h:122)
inline void warmupBuildGPU() { quarry_gpu_detail::check(qr_warmup(), "warmupBuildGPU"); }

// ---- build (src/gpu/gpu_bvh.h:103-118) --------------------------------------------------

inline GpuBVH buildBVHGPU(const Scene& scene, GpuBuildTimings* timings = nullptr) {
    return quarry_gpu_detail::build(QR_PRIM_TRIANGLE, QR_BUILDER_LBVH, scene.tris.data(),
                                    scene.tris.size(), timings, "buildBVHGPU");
}
inline GpuBVH buildBVHSahGPU(const Scene& scene, GpuBuildTimings* timings = nullptr) {
    return quarry_gpu_detail::build(QR_PRIM_TRIANGLE, QR_BUILDER_SAH, scene.tris.data(),
                                    scene.tris.size(), timings, "buildBVHSahGPU");
}
inline GpuBVH buildPointBVHGPU(const PointCloud& cloud, GpuBuildTimings* timings = nullptr) {
    return quarry_gpu_detail::build(QR_PRIM_POINT, QR_BUILDER_LBVH, cloud.pts.data(), cloud.pts.size(),
                                    timings, "buildPointBVHGPU");
}
inline GpuBVH buildPointBVHSahGPU(const PointCloud& cloud, GpuBuildTimings* timings = nullptr) {
    return quarry_gpu_detail::build(QR_PRIM_POINT, QR_BUILDER_SAH, cloud.pts.data(), cloud.pts.size(),
                                    timings, "buildPointBVHSahGPU");
}
inline GpuBVH buildGaussianBVHGPU(const GaussianCloud& cloud, GpuBuildTimings* timings = nullptr) {
    return quarry_gpu_detail::build(QR_PRIM_GAUSSIAN, QR_BUILDER_LBVH, cloud.gaussians.data(),
                                    cloud.gaussians.size(), timings, "buildGaussianBVHGPU");
}
inline GpuBVH buildGaussianBVHSahGPU(const GaussianCloud& cloud, GpuBuildTimings* timings = nullptr) {
    return quarry_gpu_detail::build(QR_PRIM_GAUSSIAN, QR_BUILDER_SAH, cloud.gaussians.data(),
                                    cloud.gaussians.size(), timings, "buildGaussianBVHSahGPU");
}

// ---- refit (src/gpu/gpu_bvh.h:141-155) ---------------------------------------------------

inline double refitBVHGPU(GpuBVH& bvh, const Scene& scene, GpuRefitTimings* timings = nullptr) {
    return quarry_gpu_detail::refit(bvh, QR_PRIM_TRIANGLE, scene.tris.data(), scene.tris.size(), nullptr,
                                    timings, "refitBVHGPU");
}
inline double refitBVHGPU(GpuBVH& bvh, const Scene& scene, const std::vector<int>& dirty,
                          GpuRefitTimings* timings = nullptr) {
    return quarry_gpu_detail::refit(bvh, QR_PRIM_TRIANGLE, scene.tris.data(), scene.tris.size(), &dirty,
                                    timings, "refitBVHGPU");
}
inline double refitPointBVHGPU(GpuBVH& bvh, const PointCloud& cloud, GpuRefitTimings* timings = nullptr) {
    return quarry_gpu_detail::refit(bvh, QR_PRIM_POINT, cloud.pts.data(), cloud.pts.size(), nullptr,
                                    timings, "refitPointBVHGPU");
}
inline double refitPointBVHGPU(GpuBVH& bvh, const PointCloud& cloud, const std::vector<int>& dirty,
                               GpuRefitTimings* timings = nullptr) {
    return quarry_gpu_detail::refit(bvh, QR_PRIM_POINT, cloud.pts.data(), cloud.pts.size(), &dirty,
                                    timings, "refitPointBVHGPU");
}
inline double refitGaussianBVHGPU(GpuBVH& bvh, const GaussianCloud& cloud,
                                  GpuRefitTimings* timings = nullptr) {
    return quarry_gpu_detail::refit(bvh, QR_PRIM_GAUSSIAN, cloud.gaussians.data(), cloud.gaussians.size(),
                                    nullptr, timings, "refitGaussianBVHGPU");
}
inline double refitGaussianBVHGPU(GpuBVH& bvh, const GaussianCloud& cloud, const std::vector<int>& dirty,
                                  GpuRefitTimings* timings = nullptr) {
    return quarry_gpu_detail::refit(bvh, QR_PRIM_GAUSSIAN, cloud.gaussians.data(), cloud.gaussians.size(),
                                    &dirty, timings, "refitGaussianBVHGPU");
}

// ---- topology (src/gpu/gpu_bvh.h:208-213) -------------------------------------------------

inline double updateBVHGPU(GpuBVH& bvh, const Scene& scene, const GpuTopoEdit& edit,
                           GpuTopoUpdateTimings* timings = nullptr) {
    return quarry_gpu_detail::update(bvh, QR_PRIM_TRIANGLE, scene.tris.data(), scene.tris.size(), edit,
                                     timings);
}
inline double updatePointBVHGPU(GpuBVH& bvh, const PointCloud& cloud, const GpuTopoEdit& edit,
                                GpuTopoUpdateTimings* timings = nullptr) {
    return quarry_gpu_detail::update(bvh, QR_PRIM_POINT, cloud.pts.data(), cloud.pts.size(), edit,
                                     timings);
}
inline double updateGaussianBVHGPU(GpuBVH& bvh, const GaussianCloud& cloud, const GpuTopoEdit& edit,
                                   GpuTopoUpdateTimings* timings = nullptr) {
    return quarry_gpu_detail::update(bvh, QR_PRIM_GAUSSIAN, cloud.gaussians.data(), cloud.gaussians.size(),
                                     edit, timings);
}

// destroyBVHGPU (src/gpu/gpu_bvh.h:215)
inline void destroyBVHGPU(GpuBVH& bvh) {
    if (bvh.handle) qr_bvh_destroy(bvh.handle);
    bvh = GpuBVH{};
}

// ---- rays (src/gpu/gpu_bvh.h:217-246) ----------------------------------------------------

inline GpuRayBuffer uploadRaysGPU(const std::vector<Ray>& rays) {
    static_assert(sizeof(Ray) == 32, "Ray is 32 bytes");
    GpuRayBuffer buf;
    quarry_gpu_detail::check(qr_rays_upload(rays.data(), static_cast<int>(rays.size()), &buf.handle),
                             "uploadRaysGPU");
    uint64_t* res = nullptr;
    quarry_gpu_detail::check(qr_rays_device_ptrs(buf.handle, &buf.d_rays, &res, &buf.count),
                             "uploadRaysGPU");
    buf.d_results = res;
    return buf;
}

inline void destroyRayBufferGPU(GpuRayBuffer& buf) {
    if (buf.handle) qr_rays_destroy(buf.handle);
    buf = GpuRayBuffer{};
}

// buildWideNodesGPU (src/gpu/gpu_bvh.h:229)
inline double buildWideNodesGPU(GpuBVH& bvh) {
    if (bvh.handle == nullptr) throw std::runtime_error("buildWideNodesGPU: BVH not built");
    double ms = 0.0;
    quarry_gpu_detail::check(qr_bvh_build_wide(bvh.handle, &ms), "buildWideNodesGPU");
    quarry_gpu_detail::sync(bvh);
    return ms;
}

// raycastGPU (src/gpu/gpu_bvh.h:234-235): kernel milliseconds.
inline double raycastGPU(const GpuBVH& bvh, GpuRayBuffer& buf, bool collectStats = false,
                         RayOrder order = RayOrder::Input, bool wide = false) {
    if (buf.count <= 0) return 0.0;
    quarry_gpu_detail::requireKind(bvh, QR_PRIM_TRIANGLE, "raycastGPU");
    double ms = 0.0;
    quarry_gpu_detail::check(qr_raycast(bvh.handle, buf.handle, collectStats ? 1 : 0,
                                        static_cast<int>(order), wide ? 1 : 0, &ms),
                             "raycastGPU");
    return ms;
}

// summarizeGPU (src/gpu/gpu_bvh.h:238): the RayStats fold over the per-ray results.
inline RayStats summarizeGPU(const GpuRayBuffer& buf, std::vector<uint64_t>* perRay = nullptr) {
    RayStats stats;
    if (buf.count <= 0) return stats;
    std::vector<uint64_t> host(static_cast<size_t>(buf.count));
    qr_ray_stats s{};
    quarry_gpu_detail::check(qr_rays_summarize(buf.handle, host.data(), &s), "summarizeGPU");
    stats.rays = s.rays;
    stats.hits = s.hits;
    stats.hash = s.hash;
    stats.sumT = s.sum_t;
    if (perRay) *perRay = std::move(host);
    return stats;
}

// lastTraversalStatsGPU (src/gpu/gpu_bvh.h:246)
inline GpuTraversalStats lastTraversalStatsGPU() {
    qr_traversal_stats s{};
    quarry_gpu_detail::check(qr_last_traversal_stats(&s), "lastTraversalStatsGPU");
    GpuTraversalStats out;
    out.nodeTests = s.node_tests;
    out.triTests = s.leaf_tests;
    out.stackOverflows = s.stack_overflows;
    out.sortMs = s.sort_ms;
    return out;
}

// ---- point queries (src/gpu/gpu_bvh.h:276-342) -------------------------------------------

inline GpuPointBuffer uploadPointsGPU(const std::vector<Vec3>& points, int resultsPerPoint = 1) {
    static_assert(sizeof(Vec3) == 12, "Vec3 is 12 bytes");
    if (resultsPerPoint < 1) throw std::runtime_error("uploadPointsGPU: resultsPerPoint must be >= 1");
    GpuPointBuffer buf;
    quarry_gpu_detail::check(qr_points_upload(reinterpret_cast<const float*>(points.data()),
                                              static_cast<int>(points.size()), resultsPerPoint,
                                              &buf.handle),
                             "uploadPointsGPU");
    float* pts = nullptr;
    quarry_gpu_detail::check(qr_points_device_ptrs(buf.handle, &pts, &buf.d_results), "uploadPointsGPU");
    buf.d_points = pts;
    buf.count = static_cast<int>(points.size());
    buf.stride = resultsPerPoint;
    // query bounds, as the reference records them at upload (src/gpu/query.cu:826-836)
    if (!points.empty()) {
        buf.bounds.mn = points[0];
        buf.bounds.mx = points[0];
        for (const Vec3& p : points) buf.bounds.extend(p);
    }
    return buf;
}

inline void destroyPointBufferGPU(GpuPointBuffer& buf) {
    if (buf.handle) qr_points_destroy(buf.handle);
    buf = GpuPointBuffer{};
}

inline double closestPointGPU(const GpuBVH& bvh, GpuPointBuffer& buf, bool collectStats = false,
                              QueryOrder order = QueryOrder::Input, bool wide = false) {
    return quarry_gpu_detail::closest(bvh, QR_PRIM_TRIANGLE, buf, collectStats, order, wide,
                                      "closestPointGPU");
}
inline double kNearestGPU(const GpuBVH& bvh, GpuPointBuffer& buf, int k, bool collectStats = false,
                          QueryOrder order = QueryOrder::Input, bool wide = false) {
    return quarry_gpu_detail::knn(bvh, QR_PRIM_TRIANGLE, buf, k, collectStats, order, wide, "kNearestGPU");
}

inline void destroyRadiusBufferGPU(GpuRadiusBuffer& buf) { quarry_gpu_detail::releaseRadius(buf); }

inline double radiusQueryGPU(const GpuBVH& bvh, GpuPointBuffer& buf, float radius, GpuRadiusBuffer& out,
                             bool collectStats = false, QueryOrder order = QueryOrder::Input,
                             bool wide = false) {
    return quarry_gpu_detail::radius(bvh, QR_PRIM_TRIANGLE, buf, radius, out, collectStats, order, wide,
                                     "radiusQueryGPU");
}

// readRadiusResultsGPU (src/gpu/gpu_bvh.h:314): per query, hits ascending by
// (dist2, index) with found, dist2 and triangle set.
inline std::vector<std::vector<ClosestHit>> readRadiusResultsGPU(const GpuRadiusBuffer& buf) {
    std::vector<std::vector<ClosestHit>> out;
    if (buf.count <= 0 || buf.handle == nullptr) return out;
    std::vector<int> offsets(static_cast<size_t>(buf.count) + 1);
    std::vector<uint64_t> keys(static_cast<size_t>(buf.total > 0 ? buf.total : 1));
    quarry_gpu_detail::check(qr_radius_read(buf.handle, offsets.data(), keys.data()),
                             "readRadiusResultsGPU");
    out.resize(static_cast<size_t>(buf.count));
    for (int i = 0; i < buf.count; ++i) {
        std::vector<ClosestHit>& list = out[static_cast<size_t>(i)];
        list.reserve(static_cast<size_t>(offsets[i + 1] - offsets[i]));
        for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
            ClosestHit h;
            h.found = true;
            const uint64_t key = keys[static_cast<size_t>(j)];
            const uint32_t bits = static_cast<uint32_t>(key >> 32);
            std::memcpy(&h.dist2, &bits, sizeof(h.dist2));
            h.triangle = static_cast<int>(static_cast<uint32_t>(key));
            list.push_back(h);
        }
    }
    return out;
}

inline double closestVertexGPU(const GpuBVH& bvh, GpuPointBuffer& buf, bool collectStats = false,
                               QueryOrder order = QueryOrder::Input, bool wide = false) {
    return quarry_gpu_detail::closest(bvh, QR_PRIM_POINT, buf, collectStats, order, wide,
                                      "closestVertexGPU");
}
inline double kNearestVertexGPU(const GpuBVH& bvh, GpuPointBuffer& buf, int k, bool collectStats = false,
                                QueryOrder order = QueryOrder::Input, bool wide = false) {
    return quarry_gpu_detail::knn(bvh, QR_PRIM_POINT, buf, k, collectStats, order, wide,
                                  "kNearestVertexGPU");
}
inline double radiusQueryVertexGPU(const GpuBVH& bvh, GpuPointBuffer& buf, float radius,
                                   GpuRadiusBuffer& out, bool collectStats = false,
                                   QueryOrder order = QueryOrder::Input, bool wide = false) {
    return quarry_gpu_detail::radius(bvh, QR_PRIM_POINT, buf, radius, out, collectStats, order, wide,
                                     "radiusQueryVertexGPU");
}

inline double closestGaussianGPU(const GpuBVH& bvh, GpuPointBuffer& buf, bool collectStats = false,
                                 QueryOrder order = QueryOrder::Input, bool wide = false) {
    return quarry_gpu_detail::closest(bvh, QR_PRIM_GAUSSIAN, buf, collectStats, order, wide,
                                      "closestGaussianGPU");
}
inline double kNearestGaussianGPU(const GpuBVH& bvh, GpuPointBuffer& buf, int k,
                                  bool collectStats = false, QueryOrder order = QueryOrder::Input,
                                  bool wide = false) {
    return quarry_gpu_detail::knn(bvh, QR_PRIM_GAUSSIAN, buf, k, collectStats, order, wide,
                                  "kNearestGaussianGPU");
}
inline double radiusQueryGaussianGPU(const GpuBVH& bvh, GpuPointBuffer& buf, float radius,
                                     GpuRadiusBuffer& out, bool collectStats = false,
                                     QueryOrder order = QueryOrder::Input, bool wide = false) {
    return quarry_gpu_detail::radius(bvh, QR_PRIM_GAUSSIAN, buf, radius, out, collectStats, order, wide,
                                     "radiusQueryGaussianGPU");
}

// readQueryResultsGPU (src/gpu/gpu_bvh.h:339): count * stride records, query order.
inline std::vector<GpuClosestResult> readQueryResultsGPU(const GpuPointBuffer& buf) {
    std::vector<GpuClosestResult> host(static_cast<size_t>(buf.count) * static_cast<size_t>(buf.stride));
    if (buf.count <= 0 || buf.handle == nullptr) return host;
    quarry_gpu_detail::check(qr_points_read(buf.handle, host.data()), "readQueryResultsGPU");
    return host;
}

// lastQueryStatsGPU (src/gpu/gpu_bvh.h:342)
inline GpuQueryStats lastQueryStatsGPU() {
    qr_traversal_stats s{};
    quarry_gpu_detail::check(qr_last_query_stats(&s), "lastQueryStatsGPU");
    GpuQueryStats out;
    out.nodeTests = s.node_tests;
    out.triTests = s.leaf_tests;
    out.stackOverflows = s.stack_overflows;
    out.sortMs = s.sort_ms;
    return out;
}

// warmupQueryGPU (src/gpu/gpu_bvh.h:343)
inline void warmupQueryGPU() { quarry_gpu_detail::check(qr_warmup(), "warmupQueryGPU"); }

// ---- checks (src/gpu/gpu_bvh.h:347-352) --------------------------------------------------

inline std::string validateBVHGPU(const GpuBVH& bvh, const Scene& scene) {
    return quarry_gpu_detail::validate(bvh, QR_PRIM_TRIANGLE, scene.tris.data(), scene.tris.size());
}
inline std::string validatePointBVHGPU(const GpuBVH& bvh, const PointCloud& cloud) {
    return quarry_gpu_detail::validate(bvh, QR_PRIM_POINT, cloud.pts.data(), cloud.pts.size());
}
inline std::string validateGaussianBVHGPU(const GpuBVH& bvh, const GaussianCloud& cloud) {
    return quarry_gpu_detail::validate(bvh, QR_PRIM_GAUSSIAN, cloud.gaussians.data(),
                                       cloud.gaussians.size());
}

// treeHashGPU (src/gpu/gpu_bvh.h:351-352): FNV-1a over the node array bytes.
inline uint64_t treeHashGPU(const GpuBVH& bvh) {
    if (bvh.handle == nullptr) return 0xcbf29ce484222325ull;
    uint64_t h = 0;
    quarry_gpu_detail::check(qr_bvh_tree_hash(bvh.handle, &h), "treeHashGPU");
    return h;
}
