// C entry points over the reference's own CPU sources, compiled from
// /root/reference by `make ref` into _ref/libquarry_ref.so.  Test
// infrastructure only: it validates the oracle restatement and serves as the
// "reference" CPU baseline.  It adds no arithmetic of its own.
#include <cstring>
#include <thread>
#include <vector>

#include "core/fingerprint.h"
#include "core/gaussian.h"
#include "core/point_cloud.h"
#include "core/scene.h"
#include "cpu/brute.h"
#include "cpu/bvh.h"
#include "cpu/gauss_bvh.h"
#include "cpu/point_kd.h"
#include "cpu/selftest.h"

namespace {

template <typename F>
void forRange(int n, int threads, F&& f) {
    if (threads <= 1) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> pool;
    const int chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const int b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back([&, b, e] { for (int i = b; i < e; ++i) f(i); });
    }
    for (auto& th : pool) th.join();
}

void putHit(uint32_t* rec, const ClosestHit& h) {
    if (h.found) {
        std::memcpy(&rec[0], &h.dist2, 4);
        rec[1] = static_cast<uint32_t>(h.triangle);
    } else {
        rec[0] = 0;
        rec[1] = 0xffffffffu;
    }
}

uint64_t packRay(const RayHit& h) {
    uint32_t bits = 0;
    if (h.hit) std::memcpy(&bits, &h.t, 4);
    return h.hit ? ((1ull << 32) | bits) : 0ull;
}

}  // namespace

extern "C" {

int ref_selftest() { return runSelfTest() ? 1 : 0; }

// ---- reference generators (only for replaying the reference's published hashes) --
// kind: 0 terrain(grid), 1 city(buildings), 2 random(count), 3 mixed(grid, count)
int ref_scene(int kind, int grid, int count, unsigned long long seed, float* out, int cap,
              float* bounds6) {
    SceneSpec spec;
    spec.kind = kind == 0 ? "terrain" : kind == 1 ? "city" : kind == 2 ? "random" : "mixed";
    spec.grid = grid;
    spec.randomTris = count;
    spec.buildings = count;
    spec.seed = seed;
    const Scene s = buildScene(spec);
    const int n = static_cast<int>(s.tris.size());
    if (out && n <= cap) std::memcpy(out, s.tris.data(), s.tris.size() * sizeof(Tri));
    if (bounds6) {
        const float b[6] = {s.bounds.mn.x, s.bounds.mn.y, s.bounds.mn.z,
                            s.bounds.mx.x, s.bounds.mx.y, s.bounds.mx.z};
        std::memcpy(bounds6, b, sizeof(b));
    }
    return n;
}

void ref_rays(const float* bounds6, int mode, int count, unsigned long long seed, float* out) {
    AABB b;
    b.mn = {bounds6[0], bounds6[1], bounds6[2]};
    b.mx = {bounds6[3], bounds6[4], bounds6[5]};
    const auto rays = makeRays(b, mode == 0 ? RayMode::Down : RayMode::Random, count, seed);
    std::memcpy(out, rays.data(), rays.size() * sizeof(Ray));
}

void ref_query_points(const float* bounds6, int count, unsigned long long seed, float* out) {
    AABB b;
    b.mn = {bounds6[0], bounds6[1], bounds6[2]};
    b.mx = {bounds6[3], bounds6[4], bounds6[5]};
    const auto pts = makeQueryPoints(b, count, seed);
    std::memcpy(out, pts.data(), pts.size() * sizeof(Vec3));
}

// ---- triangle BVH (src/cpu/bvh.cpp) ------------------------------------------------
void* ref_bvh_new(const float* tris, int n) {
    Scene s;
    s.tris.assign(reinterpret_cast<const Tri*>(tris), reinterpret_cast<const Tri*>(tris) + n);
    s.updateBounds();
    return new BVH(buildBVH(s));
}
void ref_bvh_free(void* h) { delete static_cast<BVH*>(h); }

void ref_bvh_ray(void* h, const float* rays, int n, uint64_t* out, int threads) {
    const BVH& b = *static_cast<BVH*>(h);
    const Ray* r = reinterpret_cast<const Ray*>(rays);
    forRange(n, threads, [&](int i) { out[i] = packRay(raycastBVH(b, r[i])); });
}

void ref_bvh_knn(void* h, const float* pts, int n, int k, uint32_t* out, int threads) {
    const BVH& b = *static_cast<BVH*>(h);
    const Vec3* p = reinterpret_cast<const Vec3*>(pts);
    const int kk = k == 0 ? 1 : k;
    forRange(n, threads, [&](int i) {
        if (k == 0) {
            putHit(out + 2 * static_cast<size_t>(i), closestPointBVH(b, p[i]));
            return;
        }
        const auto list = kNearestBVH(b, p[i], k);
        for (int j = 0; j < kk; ++j) {
            uint32_t* rec = out + 2 * (static_cast<size_t>(i) * kk + j);
            if (j < static_cast<int>(list.size())) putHit(rec, list[j]);
            else { rec[0] = 0; rec[1] = 0xffffffffu; }
        }
    });
}

void ref_bvh_radius(void* h, const float* pts, int n, float r, int* counts, const int* offsets,
                    uint64_t* keys, int threads) {
    const BVH& b = *static_cast<BVH*>(h);
    const Vec3* p = reinterpret_cast<const Vec3*>(pts);
    forRange(n, threads, [&](int i) {
        const auto list = radiusBVH(b, p[i], r);
        counts[i] = static_cast<int>(list.size());
        if (!keys) return;
        for (size_t j = 0; j < list.size(); ++j) {
            uint32_t bits;
            std::memcpy(&bits, &list[j].dist2, 4);
            keys[offsets[i] + j] = (static_cast<uint64_t>(bits) << 32) | static_cast<uint32_t>(list[j].triangle);
        }
    });
}

// ---- point kd-tree (src/cpu/point_kd.cpp) -----------------------------------------------
void* ref_kd_new(const float* pts, int n) {
    PointCloud c;
    c.pts.assign(reinterpret_cast<const Vec3*>(pts), reinterpret_cast<const Vec3*>(pts) + n);
    c.updateBounds();
    return new PointKdTree(buildPointKdTree(c));
}
void ref_kd_free(void* h) { delete static_cast<PointKdTree*>(h); }
void ref_kd_knn(void* h, const float* pts, int n, int k, uint32_t* out, int threads) {
    const PointKdTree& t = *static_cast<PointKdTree*>(h);
    const Vec3* p = reinterpret_cast<const Vec3*>(pts);
    const int kk = k == 0 ? 1 : k;
    forRange(n, threads, [&](int i) {
        if (k == 0) {
            putHit(out + 2 * static_cast<size_t>(i), kdClosest(t, p[i]));
            return;
        }
        const auto list = kdKNearest(t, p[i], k);
        for (int j = 0; j < kk; ++j) {
            uint32_t* rec = out + 2 * (static_cast<size_t>(i) * kk + j);
            if (j < static_cast<int>(list.size())) putHit(rec, list[j]);
            else { rec[0] = 0; rec[1] = 0xffffffffu; }
        }
    });
}
void ref_kd_radius(void* h, const float* pts, int n, float r, int* counts, const int* offsets,
                   uint64_t* keys, int threads) {
    const PointKdTree& t = *static_cast<PointKdTree*>(h);
    const Vec3* p = reinterpret_cast<const Vec3*>(pts);
    forRange(n, threads, [&](int i) {
        const auto list = kdRadius(t, p[i], r);
        counts[i] = static_cast<int>(list.size());
        if (!keys) return;
        for (size_t j = 0; j < list.size(); ++j) {
            uint32_t bits;
            std::memcpy(&bits, &list[j].dist2, 4);
            keys[offsets[i] + j] = (static_cast<uint64_t>(bits) << 32) | static_cast<uint32_t>(list[j].triangle);
        }
    });
}

// ---- ellipsoid BVH (src/cpu/gauss_bvh.cpp) -------------------------------------------------
void* ref_gauss_new(const float* g, int n) {
    GaussianCloud c;
    c.gaussians.assign(reinterpret_cast<const Gaussian*>(g), reinterpret_cast<const Gaussian*>(g) + n);
    c.updateBounds();
    return new GaussBVH(buildGaussBVH(c));
}
void ref_gauss_free(void* h) { delete static_cast<GaussBVH*>(h); }
void ref_gauss_knn(void* h, const float* pts, int n, int k, uint32_t* out, int threads) {
    const GaussBVH& b = *static_cast<GaussBVH*>(h);
    const Vec3* p = reinterpret_cast<const Vec3*>(pts);
    const int kk = k == 0 ? 1 : k;
    forRange(n, threads, [&](int i) {
        if (k == 0) {
            putHit(out + 2 * static_cast<size_t>(i), closestGaussianBVH(b, p[i]));
            return;
        }
        const auto list = kNearestGaussianBVH(b, p[i], k);
        for (int j = 0; j < kk; ++j) {
            uint32_t* rec = out + 2 * (static_cast<size_t>(i) * kk + j);
            if (j < static_cast<int>(list.size())) putHit(rec, list[j]);
            else { rec[0] = 0; rec[1] = 0xffffffffu; }
        }
    });
}
void ref_gauss_radius(void* h, const float* pts, int n, float r, int* counts, const int* offsets,
                      uint64_t* keys, int threads) {
    const GaussBVH& b = *static_cast<GaussBVH*>(h);
    const Vec3* p = reinterpret_cast<const Vec3*>(pts);
    forRange(n, threads, [&](int i) {
        const auto list = radiusGaussianBVH(b, p[i], r);
        counts[i] = static_cast<int>(list.size());
        if (!keys) return;
        for (size_t j = 0; j < list.size(); ++j) {
            uint32_t bits;
            std::memcpy(&bits, &list[j].dist2, 4);
            keys[offsets[i] + j] = (static_cast<uint64_t>(bits) << 32) | static_cast<uint32_t>(list[j].triangle);
        }
    });
}

// ---- leaf math ---------------------------------------------------------------------------
float ref_gauss_dist2(const float* p, const float* g) {
    return gaussianDist2(*reinterpret_cast<const Vec3*>(p), *reinterpret_cast<const Gaussian*>(g));
}
float ref_tri_closest(const float* p, const float* tri) {
    return closestPointOnTriangle(*reinterpret_cast<const Vec3*>(p), *reinterpret_cast<const Tri*>(tri)).dist2;
}
int ref_tri_intersect(const float* ray, const float* tri, float* t) {
    return intersectTriangle(*reinterpret_cast<const Ray*>(ray), *reinterpret_cast<const Tri*>(tri), *t) ? 1 : 0;
}

}  // extern "C"
