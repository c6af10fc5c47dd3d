// FP32 leaf and box arithmetic for the MI355X (gfx950) kernels.
//
// Every function here reproduces the reference's single-precision operation
// order so that hit flags, t bits and squared-distance bits are identical to
// the reference CPU path.  The rules that make that possible:
//   * the whole library is compiled with -ffp-contract=off (no implicit FMA);
//     explicit fused multiply-adds are written as fmaf, which lowers to
//     v_fma_f32 (one correctly rounded operation);
//   * float division and sqrtf stay IEEE correctly rounded (clang's HIP default);
//   * min/max of coordinates use compare-select (a < b ? a : b), never the
//     hardware min/max, because the choice between +0 and -0 reaches node bytes;
//   * dot products are summed ((x*x + y*y) + z*z), left to right.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define QR_HD __host__ __device__ __forceinline__

namespace qr {

constexpr float kInf = __builtin_huge_valf();
constexpr float kFltMax = 3.402823466e+38f;

struct V3 {
    float x, y, z;
};

struct Box {
    V3 lo, hi;
};

// Reference item layouts (src/core/geometry.h:27-101, src/core/gaussian.h:12-20).
struct Tri {
    V3 a, b, c;
};  // 36 B

struct Quat {
    float x, y, z, w;
};

struct Gaussian {
    V3 mean;
    Quat rot;
    V3 scale;
};  // 40 B

struct Ray {
    V3 origin;
    V3 dir;
    float tMin;
    float tMax;
};  // 32 B

static_assert(sizeof(Tri) == 36, "Tri layout");
static_assert(sizeof(Gaussian) == 40, "Gaussian layout");
static_assert(sizeof(Ray) == 32, "Ray layout");

QR_HD V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
QR_HD V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
QR_HD V3 scale(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
QR_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
QR_HD V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
QR_HD float comp(V3 v, int axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }

QR_HD float selMin(float a, float b) { return a < b ? a : b; }
QR_HD float selMax(float a, float b) { return a > b ? a : b; }

QR_HD Box emptyBox() { return {{kInf, kInf, kInf}, {-kInf, -kInf, -kInf}}; }

// AABB::extend(Vec3) (geometry.h:57-64): the incoming point is the left operand.
QR_HD void grow(Box& b, V3 p) {
    b.lo.x = p.x < b.lo.x ? p.x : b.lo.x;
    b.lo.y = p.y < b.lo.y ? p.y : b.lo.y;
    b.lo.z = p.z < b.lo.z ? p.z : b.lo.z;
    b.hi.x = p.x > b.hi.x ? p.x : b.hi.x;
    b.hi.y = p.y > b.hi.y ? p.y : b.hi.y;
    b.hi.z = p.z > b.hi.z ? p.z : b.hi.z;
}

QR_HD bool boxValid(const Box& b) {
    return b.lo.x <= b.hi.x && b.lo.y <= b.hi.y && b.lo.z <= b.hi.z;
}

// AABB::extend(AABB) (geometry.h:66-70).
QR_HD void growBox(Box& b, const Box& o) {
    if (!boxValid(o)) return;
    grow(b, o.lo);
    grow(b, o.hi);
}

// Parent box of two children as the reference's refit kernel forms it
// (src/gpu/lbvh.cu:506-512): left child is the first operand of every select.
QR_HD Box mergeChildren(const Box& l, const Box& r) {
    return {{l.lo.x < r.lo.x ? l.lo.x : r.lo.x, l.lo.y < r.lo.y ? l.lo.y : r.lo.y,
             l.lo.z < r.lo.z ? l.lo.z : r.lo.z},
            {l.hi.x > r.hi.x ? l.hi.x : r.hi.x, l.hi.y > r.hi.y ? l.hi.y : r.hi.y,
             l.hi.z > r.hi.z ? l.hi.z : r.hi.z}};
}

// Tri::bounds (geometry.h:94-100).
QR_HD Box triBounds(const Tri& t) {
    Box b = emptyBox();
    grow(b, t.a);
    grow(b, t.b);
    grow(b, t.c);
    return b;
}

// TriOps::centroid (src/gpu/prim_ops.h:16-19): sum first, then divide.
QR_HD V3 triCentroid(const Tri& t) {
    return {(t.a.x + t.b.x + t.c.x) / 3.0f, (t.a.y + t.b.y + t.c.y) / 3.0f,
            (t.a.z + t.b.z + t.c.z) / 3.0f};
}

// Slab test with padded pruning threshold (geometry.h:282-316).
QR_HD bool rayBox(const Ray& ray, const Box& box, float tMax, float& tEnter) {
    float t0 = ray.tMin;
    float t1 = tMax * (1.0f + 1e-6f) + 1e-6f;
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        const float o = comp(ray.origin, axis);
        const float d = comp(ray.dir, axis);
        const float lo = comp(box.lo, axis);
        const float hi = comp(box.hi, axis);
        if (fabsf(d) < 1e-20f) {
            if (o < lo || o > hi) return false;
            continue;
        }
        const float inv = 1.0f / d;
        float ta = (lo - o) * inv;
        float tb = (hi - o) * inv;
        if (ta > tb) {
            const float tmp = ta;
            ta = tb;
            tb = tmp;
        }
        t0 = ta > t0 ? ta : t0;
        t1 = tb < t1 ? tb : t1;
        if (t0 > t1 * (1.0f + 1e-6f) + 1e-6f) return false;
    }
    tEnter = t0;
    return true;
}

// Same slab test with the per-axis reciprocal hoisted out of the traversal loop.
// (lo - o) * inv with inv = 1/d is exactly what rayBox computes per call, so the
// result bits are identical; the |d| < 1e-20 axes are flagged by inv == 0 sentinel
// handling below (a separate mask keeps the skip rule exact).
struct RayPre {
    V3 o;
    V3 inv;
    int flat;  // bit a set: |dir[a]| < 1e-20, the axis is tested by containment
    float tMin;
};

QR_HD RayPre rayPrepare(const Ray& ray) {
    RayPre p;
    p.o = ray.origin;
    p.tMin = ray.tMin;
    p.flat = 0;
    const float d[3] = {ray.dir.x, ray.dir.y, ray.dir.z};
    float inv[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (fabsf(d[a]) < 1e-20f) {
            p.flat |= 1 << a;
            inv[a] = 0.0f;
        } else {
            inv[a] = 1.0f / d[a];
        }
    }
    p.inv = {inv[0], inv[1], inv[2]};
    return p;
}

QR_HD bool rayBoxPre(const RayPre& r, const Box& box, float tMax, float& tEnter) {
    float t0 = r.tMin;
    float t1 = tMax * (1.0f + 1e-6f) + 1e-6f;
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        const float o = comp(r.o, axis);
        const float lo = comp(box.lo, axis);
        const float hi = comp(box.hi, axis);
        if (r.flat & (1 << axis)) {
            if (o < lo || o > hi) return false;
            continue;
        }
        const float inv = comp(r.inv, axis);
        float ta = (lo - o) * inv;
        float tb = (hi - o) * inv;
        if (ta > tb) {
            const float tmp = ta;
            ta = tb;
            tb = tmp;
        }
        t0 = ta > t0 ? ta : t0;
        t1 = tb < t1 ? tb : t1;
        if (t0 > t1 * (1.0f + 1e-6f) + 1e-6f) return false;
    }
    tEnter = t0;
    return true;
}

// hitParamInsideBox (geometry.h:112-143).
QR_HD bool tInsideBox(const Ray& ray, const Box& box, float t) {
    constexpr float kPad = 1e-6f;
    float t0 = ray.tMin;
    float t1 = kInf;
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        const float o = comp(ray.origin, axis);
        const float d = comp(ray.dir, axis);
        const float lo = comp(box.lo, axis);
        const float hi = comp(box.hi, axis);
        if (fabsf(d) < 1e-20f) {
            if (o < lo || o > hi) return false;
            continue;
        }
        const float inv = 1.0f / d;
        float ta = (lo - o) * inv;
        float tb = (hi - o) * inv;
        if (ta > tb) {
            const float tmp = ta;
            ta = tb;
            tb = tmp;
        }
        t0 = ta > t0 ? ta : t0;
        t1 = tb < t1 ? tb : t1;
        if (t0 > t1) return false;
    }
    return t >= t0 * (1.0f - kPad) - kPad && t <= t1 * (1.0f + kPad) + kPad;
}

// Moller-Trumbore with the triangle's own slab guard (geometry.h:146-176).
QR_HD bool rayTri(const Ray& ray, const Tri& tri, float& tOut) {
    constexpr float kEps = 1e-7f;
    const V3 e1 = sub(tri.b, tri.a);
    const V3 e2 = sub(tri.c, tri.a);
    const V3 p = cross(ray.dir, e2);
    const float det = dot(e1, p);
    if (fabsf(det) < kEps) return false;
    const float inv = 1.0f / det;
    const V3 tv = sub(ray.origin, tri.a);
    const float u = dot(tv, p) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    const V3 q = cross(tv, e1);
    const float v = dot(ray.dir, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    const float t = dot(e2, q) * inv;
    if (t < ray.tMin || t > ray.tMax) return false;
    if (!tInsideBox(ray, triBounds(tri), t)) return false;
    tOut = t;
    return true;
}

// Which of Ericson's exits produced the closest point (geometry.h:214-258).
// The differentiable mode stores it so the adjoint never re-derives the branch.
enum TriRegion : int {
    kRegionA = 0,
    kRegionB = 1,
    kRegionAB = 2,
    kRegionC = 3,
    kRegionAC = 4,
    kRegionBC = 5,
    kRegionFace = 6,
    kRegionDegenA = 7,
    kRegionDegenB = 8,
    kRegionDegenC = 9,
};

struct TriClosest {
    float dist2;
    float b1, b2;  // point = a + b1*(b-a) + b2*(c-a)
    int region;
};

QR_HD float distTo(V3 p, V3 q) {
    const V3 d = sub(p, q);
    return dot(d, d);
}

// closestPointOnTriangle (geometry.h:198-259).
QR_HD TriClosest closestOnTri(V3 p, const Tri& tri) {
    const V3 ab = sub(tri.b, tri.a);
    const V3 ac = sub(tri.c, tri.a);
    const V3 ap = sub(p, tri.a);

    const float d1 = dot(ab, ap);
    const float d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) return {distTo(p, tri.a), 0.0f, 0.0f, kRegionA};

    const V3 bp = sub(p, tri.b);
    const float d3 = dot(ab, bp);
    const float d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) return {distTo(p, tri.b), 1.0f, 0.0f, kRegionB};

    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        const float t = d1 / (d1 - d3);
        return {distTo(p, add(tri.a, scale(ab, t))), t, 0.0f, kRegionAB};
    }

    const V3 cp = sub(p, tri.c);
    const float d5 = dot(ab, cp);
    const float d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) return {distTo(p, tri.c), 0.0f, 1.0f, kRegionC};

    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        const float t = d2 / (d2 - d6);
        return {distTo(p, add(tri.a, scale(ac, t))), 0.0f, t, kRegionAC};
    }

    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        const float t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return {distTo(p, add(tri.b, scale(sub(tri.c, tri.b), t))), 1.0f - t, t, kRegionBC};
    }

    const float denom = va + vb + vc;
    if (!(denom > 1e-20f)) {
        TriClosest best{distTo(p, tri.a), 0.0f, 0.0f, kRegionDegenA};
        const float db = distTo(p, tri.b);
        const float dc = distTo(p, tri.c);
        if (db < best.dist2) best = {db, 1.0f, 0.0f, kRegionDegenB};
        if (dc < best.dist2) best = {dc, 0.0f, 1.0f, kRegionDegenC};
        return best;
    }
    const float b1 = vb / denom;
    const float b2 = vc / denom;
    const V3 q = add(add(tri.a, scale(ab, b1)), scale(ac, b2));
    return {distTo(p, q), b1, b2, kRegionFace};
}

// distanceSquaredToAABB (geometry.h:262-277).
QR_HD float boxDist2(V3 p, const Box& box) {
    float d2 = 0.0f;
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        const float v = comp(p, axis);
        const float lo = comp(box.lo, axis);
        const float hi = comp(box.hi, axis);
        if (v < lo) {
            const float d = lo - v;
            d2 += d * d;
        } else if (v > hi) {
            const float d = v - hi;
            d2 += d * d;
        }
    }
    return d2;
}

// ---- Gaussians as solid ellipsoids (src/core/gaussian.h:23-131) ----------------

// gaussianAxes: normalise the quaternion, expand to the three rotated axes.
QR_HD void gaussAxes(const Quat& q, V3& u, V3& v, V3& w) {
    const float lenSq = fmaf(q.w, q.w, fmaf(q.z, q.z, fmaf(q.y, q.y, q.x * q.x)));
    float inv = 1.0f;
    if (lenSq > 0.0f) inv = 1.0f / sqrtf(lenSq);
    const float qx = q.x * inv, qy = q.y * inv, qz = q.z * inv, qw = q.w * inv;
    const float yyzz = fmaf(qy, qy, qz * qz);
    const float xxzz = fmaf(qx, qx, qz * qz);
    const float xxyy = fmaf(qx, qx, qy * qy);
    u = {fmaf(-2.0f, yyzz, 1.0f), 2.0f * fmaf(qx, qy, qz * qw), 2.0f * fmaf(qx, qz, -(qy * qw))};
    v = {2.0f * fmaf(qx, qy, -(qz * qw)), fmaf(-2.0f, xxzz, 1.0f), 2.0f * fmaf(qy, qz, qx * qw)};
    w = {2.0f * fmaf(qx, qz, qy * qw), 2.0f * fmaf(qy, qz, -(qx * qw)), fmaf(-2.0f, xxyy, 1.0f)};
}

// gaussianBounds: padded AABB of the rotated ellipsoid.
QR_HD Box gaussBounds(const Gaussian& g) {
    V3 u, v, w;
    gaussAxes(g.rot, u, v, w);
    const float a2 = g.scale.x * g.scale.x;
    const float b2 = g.scale.y * g.scale.y;
    const float c2 = g.scale.z * g.scale.z;
    Box box{g.mean, g.mean};
    float lo[3] = {g.mean.x, g.mean.y, g.mean.z};
    float hi[3] = {g.mean.x, g.mean.y, g.mean.z};
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        const float ui = comp(u, axis), vi = comp(v, axis), wi = comp(w, axis);
        const float half = sqrtf(fmaf(c2, wi * wi, fmaf(b2, vi * vi, a2 * ui * ui)));
        const float pad = half + 1e-6f * (1.0f + fabsf(comp(g.mean, axis)) + half);
        lo[axis] -= pad;
        hi[axis] += pad;
    }
    box.lo = {lo[0], lo[1], lo[2]};
    box.hi = {hi[0], hi[1], hi[2]};
    return box;
}

constexpr int kGaussNewtonIters = 14;

// Everything the differentiable mode needs from one ellipsoid distance
// evaluation; dist2 alone is the plain path's result.
struct GaussClosest {
    float dist2;
    float t;       // Eberly root (0 when the point is inside)
    int inside;    // 1: point inside or on the surface
};

// gaussianDist2: inside test, then 14 fixed bracketed Newton/bisection steps.
QR_HD GaussClosest gaussClosest(V3 p, const Gaussian& g) {
    V3 u, v, w;
    gaussAxes(g.rot, u, v, w);
    const V3 d = sub(p, g.mean);
    const float y0 = dot(d, u), y1 = dot(d, v), y2 = dot(d, w);
    const float a = g.scale.x, b = g.scale.y, c = g.scale.z;
    const float aa = a * a, bb = b * b, cc = c * c;
    const float gy0 = a * y0, gy1 = b * y1, gy2 = c * y2;
    {
        const float r0 = y0 / a, r1 = y1 / b, r2 = y2 / c;
        if (fmaf(r2, r2, fmaf(r1, r1, r0 * r0)) <= 1.0f) return {0.0f, 0.0f, 1};
    }
    const float lenSq = fmaf(y2, y2, fmaf(y1, y1, y0 * y0));
    const float len = sqrtf(lenSq);
    const float amax = a > b ? (a > c ? a : c) : (b > c ? b : c);
    float lo = 0.0f;
    float hi = 2.0f * amax * len;
    float t = 0.0f;
    if (len > amax) t = amax * (len - amax);
    for (int iter = 0; iter < kGaussNewtonIters; ++iter) {
        const float ir0 = 1.0f / (t + aa), ir1 = 1.0f / (t + bb), ir2 = 1.0f / (t + cc);
        const float s0 = gy0 * ir0, s1 = gy1 * ir1, s2 = gy2 * ir2;
        const float f = fmaf(s2, s2, fmaf(s1, s1, s0 * s0)) - 1.0f;
        if (f > 0.0f) {
            lo = t;
        } else {
            hi = t;
        }
        const float fp = -2.0f * (s0 * s0 * ir0 + s1 * s1 * ir1 + s2 * s2 * ir2);
        float next = 0.5f * (lo + hi);
        if (fp < 0.0f) {
            const float newton = t - f / fp;
            if (newton > lo && newton < hi) next = newton;
        }
        t = next;
    }
    const float ir0 = 1.0f / (t + aa), ir1 = 1.0f / (t + bb), ir2 = 1.0f / (t + cc);
    const float e0 = y0 - gy0 * ir0 * a;
    const float e1 = y1 - gy1 * ir1 * b;
    const float e2 = y2 - gy2 * ir2 * c;
    return {fmaf(e2, e2, fmaf(e1, e1, e0 * e0)), t, 0};
}

QR_HD float gaussDist2(V3 p, const Gaussian& g) { return gaussClosest(p, g).dist2; }

// ---- leaf traits: the three primitive kinds share every kernel ------------------
// (src/gpu/prim_ops.h:10-53)

struct TriLeaf {
    using Item = Tri;
    static constexpr int kKind = 0;
    QR_HD static Box bounds(const Tri& t) { return triBounds(t); }
    QR_HD static V3 centroid(const Tri& t) { return triCentroid(t); }
    QR_HD static float dist2(V3 q, const Tri& t) { return closestOnTri(q, t).dist2; }
};

struct PointLeaf {
    using Item = V3;
    static constexpr int kKind = 1;
    QR_HD static Box bounds(const V3& p) {
        Box b = emptyBox();
        grow(b, p);
        return b;
    }
    QR_HD static V3 centroid(const V3& p) { return p; }
    QR_HD static float dist2(V3 q, const V3& p) { return distTo(q, p); }
};

struct GaussLeaf {
    using Item = Gaussian;
    static constexpr int kKind = 2;
    QR_HD static Box bounds(const Gaussian& g) { return gaussBounds(g); }
    QR_HD static V3 centroid(const Gaussian& g) { return g.mean; }
    QR_HD static float dist2(V3 q, const Gaussian& g) { return gaussDist2(q, g); }
};

// Monotone float <-> u32 map for atomic min/max on floats (src/gpu/lbvh.cu:213-220).
QR_HD uint32_t orderedBits(float f) {
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
QR_HD float fromOrderedBits(uint32_t u) {
    return __builtin_bit_cast(float, (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

QR_HD uint32_t floatBits(float f) { return __builtin_bit_cast(uint32_t, f); }
QR_HD float bitsFloat(uint32_t u) { return __builtin_bit_cast(float, u); }

}  // namespace qr
