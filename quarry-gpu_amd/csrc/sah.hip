// Binned SAH builder (placeholder until the level pipeline lands).
#include "qr_bvh.h"

namespace qr {

void buildSahDevice(BvhState&, int, const void*, int, hipStream_t, BuildTimingsC*) {
    throw ArgError("builder 'sah' is not available in this build");
}

}  // namespace qr
