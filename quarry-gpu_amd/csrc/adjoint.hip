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

// The Gaussian adjoint runs in double precision: the quaternion gradient is a sum
// of terms much larger than the result (it is projected onto the tangent of the
// unit sphere), which single precision would leave with ~1e-4 relative error.  The
// saved root t comes from the FP32 forward.
struct D3 {
    double x, y, z;
};
__device__ __forceinline__ double dotd(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

__global__ void adjGaussKernel(const Gaussian* __restrict__ items, const int* __restrict__ primSlot,
                               const V3* __restrict__ pts, const ClosestResult* __restrict__ res,
                               const int* __restrict__ aux, const float* __restrict__ gout, int n,
                               float* __restrict__ gpts, float* __restrict__ gitems) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const ClosestResult r = res[i];
    const float g = gout[i];
    double gq[3] = {0.0, 0.0, 0.0};
    if (r.prim >= 0 && g != 0.0f && aux[i] != -2) {
        const Gaussian gs = items[primSlot[r.prim]];
        const double t32 = static_cast<double>(bitsFloat(static_cast<uint32_t>(aux[i])));
        const double qx0 = gs.rot.x, qy0 = gs.rot.y, qz0 = gs.rot.z, qw0 = gs.rot.w;
        const double len2 = qx0 * qx0 + qy0 * qy0 + qz0 * qz0 + qw0 * qw0;
        const double inv = len2 > 0.0 ? 1.0 / sqrt(len2) : 1.0;
        const double x = qx0 * inv, y = qy0 * inv, z = qz0 * inv, w = qw0 * inv;
        const D3 u{1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w)};
        const D3 v{2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w)};
        const D3 ww{2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y)};
        const V3 q = pts[i];
        const D3 d{static_cast<double>(q.x) - gs.mean.x, static_cast<double>(q.y) - gs.mean.y,
                   static_cast<double>(q.z) - gs.mean.z};
        const double yy[3] = {dotd(d, u), dotd(d, v), dotd(d, ww)};
        const double sc[3] = {gs.scale.x, gs.scale.y, gs.scale.z};
        // The forward pass solves the multiplier in FP32; the gradient wants the root
        // of f(t) = sum_k (s_k y_k / (t + s_k^2))^2 - 1 itself.  f is convex and
        // decreasing for t > -min s_k^2, so Newton from the FP32 root converges
        // quadratically (three steps reach FP64 precision from ~1e-6 relative).
        double t = t32;
        for (int it = 0; it < 3; ++it) {
            double f = -1.0, df = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double ss = sc[k] * sc[k];
                const double a = sc[k] * yy[k] / (t + ss);
                f += a * a;
                df -= 2.0 * a * a / (t + ss);
            }
            if (!(df < 0.0)) break;
            const double tn = t - f / df;
            if (!(tn >= 0.0)) break;
            t = tn;
        }
        double e[3], pk[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double ss = sc[k] * sc[k];
            pk[k] = ss * yy[k] / (t + ss);
            e[k] = yy[k] - pk[k];
        }
        const double g2 = 2.0 * static_cast<double>(g);
        gq[0] = g2 * (e[0] * u.x + e[1] * v.x + e[2] * ww.x);
        gq[1] = g2 * (e[0] * u.y + e[1] * v.y + e[2] * ww.y);
        gq[2] = g2 * (e[0] * u.z + e[1] * v.z + e[2] * ww.z);
        if (gitems) {
            float* dst = gitems + static_cast<size_t>(r.prim) * 10;
            atomicAdd(dst + 0, static_cast<float>(-gq[0]));
            atomicAdd(dst + 1, static_cast<float>(-gq[1]));
            atomicAdd(dst + 2, static_cast<float>(-gq[2]));
            // dL/d(axis k) = g2 e_k d; chained through the rotation columns (see the
            // matrix above) and the normalisation n = q / |q|
            const D3 gu{g2 * e[0] * d.x, g2 * e[0] * d.y, g2 * e[0] * d.z};
            const D3 gv{g2 * e[1] * d.x, g2 * e[1] * d.y, g2 * e[1] * d.z};
            const D3 gw{g2 * e[2] * d.x, g2 * e[2] * d.y, g2 * e[2] * d.z};
            const double gx = gu.y * 2 * y + gu.z * 2 * z + gv.x * 2 * y + gv.y * -4 * x + gv.z * 2 * w +
                              gw.x * 2 * z + gw.y * -2 * w + gw.z * -4 * x;
            const double gy = gu.x * -4 * y + gu.y * 2 * x + gu.z * -2 * w + gv.x * 2 * x + gv.z * 2 * z +
                              gw.x * 2 * w + gw.y * 2 * z + gw.z * -4 * y;
            const double gz = gu.x * -4 * z + gu.y * 2 * w + gu.z * 2 * x + gv.x * -2 * w + gv.y * -4 * z +
                              gv.z * 2 * y + gw.x * 2 * x + gw.y * 2 * y;
            const double gwq = gu.y * 2 * z + gu.z * -2 * y + gv.x * -2 * z + gv.z * 2 * x + gw.x * 2 * y +
                               gw.y * -2 * x;
            const double dn = x * gx + y * gy + z * gz + w * gwq;
            atomicAdd(dst + 3, static_cast<float>((gx - x * dn) * inv));
            atomicAdd(dst + 4, static_cast<float>((gy - y * dn) * inv));
            atomicAdd(dst + 5, static_cast<float>((gz - z * dn) * inv));
            atomicAdd(dst + 6, static_cast<float>((gwq - w * dn) * inv));
#pragma unroll
            for (int k = 0; k < 3; ++k)
                atomicAdd(dst + 7 + k, static_cast<float>(-g2 * t * pk[k] * pk[k] / (sc[k] * sc[k] * sc[k])));
        }
    }
    if (gpts) {
        gpts[3 * i + 0] = static_cast<float>(gq[0]);
        gpts[3 * i + 1] = static_cast<float>(gq[1]);
        gpts[3 * i + 2] = static_cast<float>(gq[2]);
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
