// Differentiable mode: forward-with-save companions and adjoint launchers.
#pragma once

#include "qr_bvh.h"

namespace qr {

// For each closest-point result (scene index in ClosestResult::prim), writes the
// saved auxiliary word: triangle -> closest-feature region (TriRegion), Gaussian ->
// bits of the Eberly root t (-2 when the query is inside), point -> 0; -1 for a
// query without a result.  Returns how many saved results disagree with a
// re-evaluation of their primitive (0 unless something is badly wrong).
int saveClosestAux(BvhState& b, const float* pts, int n, const void* results, int* aux, hipStream_t s);

// dL/dq (gpts, n x 3, overwritten; may be null) and scene-order dL/d(primitive)
// (gitems, prim_count x item floats, accumulated; may be null) from dL/d(dist2).
void adjointClosest(BvhState& b, const float* pts, int n, const void* results, const int* aux,
                    const float* gout, float* gpts, float* gitems, hipStream_t s);

// dL/d(origin), dL/d(dir) (n x 3, overwritten; may be null) and scene-order
// dL/d(triangle) (accumulated; may be null) from dL/dt of hit rays.
void adjointRay(BvhState& b, const void* rays, int n, const uint64_t* results, const int* hitPrim,
                const float* gout, float* gorig, float* gdir, float* gitems, hipStream_t s);

}  // namespace qr
