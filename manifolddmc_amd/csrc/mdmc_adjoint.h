// mdmc_adjoint.h -- launchers of the adjoint passes (mdmc_adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mdmc_internal.h"

namespace mdmc {

// per dual vertex: 8 corner contributions (contrib[8V]) from dL/dpos (dpos[3V])
void launchAdjointVertex(const GridDev& g, float iso, const int32_t* adjMap, const int32_t* adjCell, int64_t V, const float* dpos, float* contrib,
                         hipStream_t s);
// per point of the processed grid g: dL/dvalue (out[P]); accumulate adds to out
void launchAdjointGather(const GridDev& g, const int32_t* adjMap, const float* contrib, float* out, int accumulate, hipStream_t s);
// processed-grid gradient -> input-grid gradient through padding / refinement
void launchPrepareAdjoint(const int inDims[3], const int outDims[3], int r, int pad, const float* dproc, float* din, int accumulate, hipStream_t s);

}  // namespace mdmc
