#pragma once
// Pieces shared by the FISP EPG kernels (primal and tangent).

#include <cfloat>

#include "compas/device/common.h"

namespace compas {
namespace kernels {

struct FispArgs {
    dev::TissueView tissue;
    const float4* table;  // [nslices][nTR]: (deg2rad*|RF|, cos phi, sin phi, 0), then one zero entry
    int nTR;
    int nslices;
    float TR, TE, TW, TI;
    int N;  // max_state (runtime; <= SPL * LPV)
    int usf;
    int repetitions;
    bool inversion;
    bool wait_spoiling;
    cfloat* echos;  // [nTR * usf][nvoxels]
};

__device__ inline float relax(float dt, float T) {
    return expf(-dt * (1.0f / T));
}

// exp(i 2 pi dt B0), or exactly 1 when B0 == 0 (fisp_kernels.cuh:110-117)
__device__ inline cfloat off_resonance(float dt, float B0) {
    if (B0 == 0.0f) return {1.0f, 0.0f};
    float s, c;
    sincosf(dev::TWO_PI_F * dt * B0, &s, &c);
    return {c, s};
}

}  // namespace kernels
}  // namespace compas
