// beamlet.hip -- beamlet dose coefficients (dosecalc-beamlet path).  Entry points are filled in
// by the beamlet pipeline; until then they report CCCS_EINVAL.
#include <hip/hip_runtime.h>

#include "context.hpp"

using namespace cccs;

extern "C" int cccs_beamlet_dose(cccs_ctx*, const cccs_beam*, int, const float* const*, float, int,
                                 cccs_sparse_sink, void*) {
    return set_error(CCCS_EINVAL, "beamlet path not built yet");
}

extern "C" int cccs_beamlet_dense(cccs_ctx*, const cccs_beam*, const float*, int, float*, int, int*) {
    return set_error(CCCS_EINVAL, "beamlet path not built yet");
}
