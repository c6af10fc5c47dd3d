// Differentiable mode: forward-with-save companions and adjoint kernels.
//
// The forward results are the plain kernels' results, bit for bit.  What the
// backward pass needs beyond them is saved alongside:
//   * closest point on a triangle: the closest-feature region (which of
//     Ericson's exits, src/core/geometry.h:214-258, produced the point);
//   * closest point on a Gaussian: the Eberly root t of the final iterate
//     (src/core/gaussian.h:103-124);
//   * ray hits: the hit triangle's scene index (the reference ray kernel discards
//     it, src/gpu/raycast.cu:108-111; the save variant of the ray kernel keeps it).
// The adjoint kernels read the saved primitive (through the tree's leaf-order
// items and the scene -> slot map) and never traverse the tree again.
//
// Gradients (g = upstream dL/d(dist2) or dL/dt per query; p = closest point):
//   triangle closest point, with w the barycentric weights of p implied by the
//   saved region (one-hot at a vertex, the edge parameter on an edge, the face
//   barycentrics inside; p is the constrained minimiser, so by the envelope
//   theorem only the explicit dependence remains):
//     dL/dq = 2 g (q - p),  dL/dv_i = -2 g w_i (q - p);
//   point: dL/dq = 2 g (q - p), dL/dp = -2 g (q - p);
//   Gaussian (outside; everything is zero inside), local y = R^T (q - m),
//   p_k = s_k^2 y_k / (t + s_k^2), e = y - p:
//     dL/dq = 2 g R e,  dL/dm = -dL/dq,  dL/ds_k = -2 g t p_k^2 / s_k^3,
//     dL/dR[:,k] = 2 g e_k (q - m), chained through the normalised quaternion;
//   ray hit t with n = (b - a) x (c - a), w the hit's barycentrics:
//     dt/do = -n / (n.d),  dt/dd = -t n / (n.d),  dt/dv_i = w_i n / (n.d).
// Scene-order gradient arrays are accumulated with atomics (several queries may
// share a primitive).
#include "qr_adjoint.h"

namespace qr {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ void atomicAdd3(float* dst, V3 v) {
    atomicAdd(dst + 0, v.x);
    atomicAdd(dst + 1, v.y);
    atomicAdd(dst + 2, v.z);
}

// Barycentric weights (of a, b, c) of the closest point for a known region,
// using the forward's own expressions for that branch.
__device__ void regionWeights(V3 p, const Tri& tri, int region, float& wa, float& wb, float& wc) {
    const V3 ab = sub(tri.b, tri.a);
    const V3 ac = sub(tri.c, tri.a);
    const V3 ap = sub(p, tri.a);
    const float d1 = dot(ab, ap);
    const float d2 = dot(ac, ap);
    const V3 bp = sub(p, tri.b);
    const float d3 = dot(ab, bp);
    const float d4 = dot(ac, bp);
    const V3 cp = sub(p, tri.c);
    const float d5 = dot(ab, cp);
    const float d6 = dot(ac, cp);
    wa = wb = wc = 0.0f;
    switch (region) {
        case kRegionA:
        case kRegionDegenA: wa = 1.0f; break;
        case kRegionB:
        case kRegionDegenB: wb = 1.0f; break;
        case kRegionC:
        case kRegionDegenC: wc = 1.0f; break;
        case kRegionAB: {
            const float t = d1 / (d1 - d3);
            wa = 1.0f - t;
            wb = t;
            break;
        }
        case kRegionAC: {
            const float t = d2 / (d2 - d6);
            wa = 1.0f - t;
            wc = t;
            break;
        }
        case kRegionBC: {
            const float t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
            wb = 1.0f - t;
            wc = t;
            break;
        }
        default: {
            const float vc = d1 * d4 - d3 * d2;
            const float vb = d5 * d2 - d1 * d6;
            const float va = d3 * d6 - d5 * d4;
            const float denom = va + vb + vc;
            const float b1 = vb / denom, b2 = vc / denom;
            wa = 1.0f - b1 - b2;
            wb = b1;
            wc = b2;
            break;
        }
    }
}

// ---- forward-with-save companion: region / Eberly root of each saved closest hit ----
template <typename L>
__global__ void saveAuxKernel(const typename L::Item* __restrict__ items, const int* __restrict__ primSlot,
                              const V3* __restrict__ pts, const ClosestResult* __restrict__ res, int n,
                              int* __restrict__ aux, int* __restrict__ mismatch) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const ClosestResult r = res[i];
    int a = -1;
    if (r.prim >= 0) {
        const typename L::Item it = items[primSlot[r.prim]];
        uint32_t bits;
        if constexpr (L::kKind == 0) {
            const TriClosest c = closestOnTri(pts[i], it);
            a = c.region;
            bits = floatBits(c.dist2);
        } else if constexpr (L::kKind == 2) {
            const GaussClosest c = gaussClosest(pts[i], it);
            a = c.inside ? -2 : static_cast<int>(floatBits(c.t));
            bits = floatBits(c.dist2);
        } else {
            a = 0;
            bits = floatBits(L::dist2(pts[i], it));
        }
        if (bits != r.dist2Bits) atomicAdd(mismatch, 1);
    }
    aux[i] = a;
}

// ---- adjoints of closest-point squared distance ----------------------------------------
__global__ void adjTriKernel(const Tri* __restrict__ items, const int* __restrict__ primSlot,
                             const V3* __restrict__ pts, const ClosestResult* __restrict__ res,
                             const int* __restrict__ aux, const float* __restrict__ gout, int n,
                             float* __restrict__ gpts, float* __restrict__ gitems) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const ClosestResult r = res[i];
    const float g = gout[i];
    V3 gq{0.0f, 0.0f, 0.0f};
    if (r.prim >= 0 && g != 0.0f) {
        const Tri tri = items[primSlot[r.prim]];
        const V3 q = pts[i];
        float wa, wb, wc;
        regionWeights(q, tri, aux[i], wa, wb, wc);
        const V3 p = add(add(scale(tri.a, wa), scale(tri.b, wb)), scale(tri.c, wc));
        const V3 d = scale(sub(q, p), 2.0f * g);
        gq = d;
        if (gitems) {
            float* dst = gitems + static_cast<size_t>(r.prim) * 9;
            atomicAdd3(dst + 0, scale(d, -wa));
            atomicAdd3(dst + 3, scale(d, -wb));
            atomicAdd3(dst + 6, scale(d, -wc));
        }
    }
    if (gpts) {
        gpts[3 * i + 0] = gq.x;
        gpts[3 * i + 1] = gq.y;
        gpts[3 * i + 2] = gq.z;
    }
}

__global__ void adjPointKernel(const V3* __restrict__ items, const int* __restrict__ primSlot,
                               const V3* __restrict__ pts, const ClosestResult* __restrict__ res,
                               const float* __restrict__ gout, int n, float* __restrict__ gpts,
                               float* __restrict__ gitems) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const ClosestResult r = res[i];
    const float g = gout[i];
    V3 gq{0.0f, 0.0f, 0.0f};
    if (r.prim >= 0 && g != 0.0f) {
        const V3 p = items[primSlot[r.prim]];
        gq = scale(sub(pts[i], p), 2.0f * g);
        if (gitems) atomicAdd3(gitems + static_cast<size_t>(r.prim) * 3, scale(gq, -1.0f));
    }
    if (gpts) {
        gpts[3 * i + 0] = gq.x;
        gpts[3 * i + 1] = gq.y;
        gpts[3 * i + 2] = gq.z;
    }
}

// d(axes)/d(normalised quaternion): the rotation columns u, v, w of gaussAxes.
__device__ void quatGrad(const Quat& q, V3 gu, V3 gv, V3 gw, float out[4]) {
    const float lenSq = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
    const float inv = lenSq > 0.0f ? 1.0f / sqrtf(lenSq) : 1.0f;
    const float x = q.x * inv, y = q.y * inv, z = q.z * inv, w = q.w * inv;
    // u = (1-2(yy+zz), 2(xy+zw), 2(xz-yw)); v = (2(xy-zw), 1-2(xx+zz), 2(yz+xw));
    // w = (2(xz+yw), 2(yz-xw), 1-2(xx+yy))
    const float gx = gu.y * 2 * y + gu.z * 2 * z + gv.x * 2 * y + gv.y * -4 * x + gv.z * 2 * w + gw.x * 2 * z +
                     gw.y * -2 * w + gw.z * -4 * x;
    const float gy = gu.x * -4 * y + gu.y * 2 * x + gu.z * -2 * w + gv.x * 2 * x + gv.z * 2 * z + gw.x * 2 * w +
                     gw.y * 2 * z + gw.z * -4 * y;
    const float gz = gu.x * -4 * z + gu.y * 2 * w + gu.z * 2 * x + gv.x * -2 * w + gv.y * -4 * z + gv.z * 2 * y +
                     gw.x * 2 * x + gw.y * 2 * y;
    const float gw_ = gu.y * 2 * z + gu.z * -2 * y + gv.x * -2 * z + gv.z * 2 * x + gw.x * 2 * y + gw.y * -2 * x;
    // through the normalisation n = q / |q|: dL/dq = (gn - n (n . gn)) / |q|
    const float dotng = x * gx + y * gy + z * gz + w * gw_;
    out[0] = (gx - x * dotng) * inv;
    out[1] = (gy - y * dotng) * inv;
    out[2] = (gz - z * dotng) * inv;
    out[3] = (gw_ - w * dotng) * inv;
}

__global__ void adjGaussKernel(const Gaussian* __restrict__ items, const int* __restrict__ primSlot,
                               const V3* __restrict__ pts, const ClosestResult* __restrict__ res,
                               const int* __restrict__ aux, const float* __restrict__ gout, int n,
                               float* __restrict__ gpts, float* __restrict__ gitems) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const ClosestResult r = res[i];
    const float g = gout[i];
    V3 gq{0.0f, 0.0f, 0.0f};
    if (r.prim >= 0 && g != 0.0f && aux[i] != -2) {
        const Gaussian gs = items[primSlot[r.prim]];
        const float t = bitsFloat(static_cast<uint32_t>(aux[i]));
        V3 u, v, w;
        gaussAxes(gs.rot, u, v, w);
        const V3 d = sub(pts[i], gs.mean);
        const float y[3] = {dot(d, u), dot(d, v), dot(d, w)};
        const float s[3] = {gs.scale.x, gs.scale.y, gs.scale.z};
        float e[3], pk[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float ss = s[k] * s[k];
            pk[k] = ss * y[k] / (t + ss);
            e[k] = y[k] - pk[k];
        }
        const float g2 = 2.0f * g;
        gq = scale(add(add(scale(u, e[0]), scale(v, e[1])), scale(w, e[2])), g2);
        if (gitems) {
            float* dst = gitems + static_cast<size_t>(r.prim) * 10;
            atomicAdd3(dst + 0, scale(gq, -1.0f));
            float gqt[4];
            quatGrad(gs.rot, scale(d, g2 * e[0]), scale(d, g2 * e[1]), scale(d, g2 * e[2]), gqt);
#pragma unroll
            for (int k = 0; k < 4; ++k) atomicAdd(dst + 3 + k, gqt[k]);
#pragma unroll
            for (int k = 0; k < 3; ++k)
                atomicAdd(dst + 7 + k, -g2 * t * pk[k] * pk[k] / (s[k] * s[k] * s[k]));
        }
    }
    if (gpts) {
        gpts[3 * i + 0] = gq.x;
        gpts[3 * i + 1] = gq.y;
        gpts[3 * i + 2] = gq.z;
    }
}

// ---- adjoint of ray hit t ---------------------------------------------------------------
__global__ void adjRayKernel(const Tri* __restrict__ items, const int* __restrict__ primSlot,
                             const Ray* __restrict__ rays, const uint64_t* __restrict__ res,
                             const int* __restrict__ hitPrim, const float* __restrict__ gout, int n,
                             float* __restrict__ gorig, float* __restrict__ gdir, float* __restrict__ gitems) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint64_t rv = res[i];
    const int prim = hitPrim[i];
    const float g = gout[i];
    V3 go{0.0f, 0.0f, 0.0f}, gd{0.0f, 0.0f, 0.0f};
    if ((rv >> 32) != 0 && prim >= 0 && g != 0.0f) {
        const Tri tri = items[primSlot[prim]];
        const Ray ray = rays[i];
        const float t = bitsFloat(static_cast<uint32_t>(rv));
        const V3 e1 = sub(tri.b, tri.a);
        const V3 e2 = sub(tri.c, tri.a);
        const V3 nrm = cross(e1, e2);
        const float nd = dot(nrm, ray.dir);
        // barycentrics of the hit (Moller-Trumbore u, v)
        const V3 pv = cross(ray.dir, e2);
        const float det = dot(e1, pv);
        const float inv = 1.0f / det;
        const V3 tv = sub(ray.origin, tri.a);
        const float bu = dot(tv, pv) * inv;
        const V3 qv = cross(tv, e1);
        const float bv = dot(ray.dir, qv) * inv;
        const V3 k = scale(nrm, g / nd);
        go = scale(k, -1.0f);
        gd = scale(k, -t);
        if (gitems) {
            float* dst = gitems + static_cast<size_t>(prim) * 9;
            atomicAdd3(dst + 0, scale(k, 1.0f - bu - bv));
            atomicAdd3(dst + 3, scale(k, bu));
            atomicAdd3(dst + 6, scale(k, bv));
        }
    }
    if (gorig) {
        gorig[3 * i + 0] = go.x;
        gorig[3 * i + 1] = go.y;
        gorig[3 * i + 2] = go.z;
    }
    if (gdir) {
        gdir[3 * i + 0] = gd.x;
        gdir[3 * i + 1] = gd.y;
        gdir[3 * i + 2] = gd.z;
    }
}

}  // namespace

int saveClosestAux(BvhState& b, const float* pts, int n, const void* results, int* aux, hipStream_t s) {
    if (n <= 0) return 0;
    ensurePrimSlot(b, s);
    DevBuf<int> mism(1);
    hipCheck(hipMemsetAsync(mism.p, 0, sizeof(int), s), "aux clear");
    const int grid = blocksFor(n, kBlock);
    const auto* P = reinterpret_cast<const V3*>(pts);
    const auto* R = static_cast<const ClosestResult*>(results);
    switch (b.kind) {
        case kPrimTri:
            saveAuxKernel<TriLeaf><<<grid, kBlock, 0, s>>>(reinterpret_cast<const Tri*>(b.items.p), b.primSlot.p, P,
                                                          R, n, aux, mism.p);
            break;
        case kPrimPoint:
            saveAuxKernel<PointLeaf><<<grid, kBlock, 0, s>>>(reinterpret_cast<const V3*>(b.items.p), b.primSlot.p,
                                                            P, R, n, aux, mism.p);
            break;
        default:
            saveAuxKernel<GaussLeaf><<<grid, kBlock, 0, s>>>(reinterpret_cast<const Gaussian*>(b.items.p),
                                                            b.primSlot.p, P, R, n, aux, mism.p);
            break;
    }
    hipCheck(hipGetLastError(), "save aux");
    int h = 0;
    hipCheck(hipMemcpyAsync(&h, mism.p, sizeof(int), hipMemcpyDeviceToHost, s), "aux check");
    hipCheck(hipStreamSynchronize(s), "aux sync");
    return h;
}

void adjointClosest(BvhState& b, const float* pts, int n, const void* results, const int* aux,
                    const float* gout, float* gpts, float* gitems, hipStream_t s) {
    if (n <= 0) return;
    ensurePrimSlot(b, s);
    const int grid = blocksFor(n, kBlock);
    const auto* P = reinterpret_cast<const V3*>(pts);
    const auto* R = static_cast<const ClosestResult*>(results);
    switch (b.kind) {
        case kPrimTri:
            adjTriKernel<<<grid, kBlock, 0, s>>>(reinterpret_cast<const Tri*>(b.items.p), b.primSlot.p, P, R, aux,
                                                 gout, n, gpts, gitems);
            break;
        case kPrimPoint:
            adjPointKernel<<<grid, kBlock, 0, s>>>(reinterpret_cast<const V3*>(b.items.p), b.primSlot.p, P, R, gout,
                                                   n, gpts, gitems);
            break;
        default:
            adjGaussKernel<<<grid, kBlock, 0, s>>>(reinterpret_cast<const Gaussian*>(b.items.p), b.primSlot.p, P,
                                                   R, aux, gout, n, gpts, gitems);
            break;
    }
    hipCheck(hipGetLastError(), "adjoint closest");
}

void adjointRay(BvhState& b, const void* rays, int n, const uint64_t* results, const int* hitPrim,
                const float* gout, float* gorig, float* gdir, float* gitems, hipStream_t s) {
    if (n <= 0) return;
    if (b.kind != kPrimTri) throw ArgError("ray adjoint: the tree does not hold triangles");
    ensurePrimSlot(b, s);
    adjRayKernel<<<blocksFor(n, kBlock), kBlock, 0, s>>>(reinterpret_cast<const Tri*>(b.items.p), b.primSlot.p,
                                                         static_cast<const Ray*>(rays), results, hitPrim, gout, n,
                                                         gorig, gdir, gitems);
    hipCheck(hipGetLastError(), "adjoint ray");
}

}  // namespace qr
