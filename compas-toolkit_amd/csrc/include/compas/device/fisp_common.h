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

// Largest flip angle (deg2rad |RF|) of the launch's RF table, the same in
// every lane of the wave.  Every lane of the wave must be active.
__device__ inline float fisp_table_max_angle(const FispArgs& a) {
    const int total = a.nslices * a.nTR;
    float m = 0.0f;
    for (int i = int(__lane_id()); i < total; i += 64) m = fmaxf(m, fabsf(a.table[i].x));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    return m;
}

// Forms of the half flip angle's sincos in the TR loop
enum FispHalfAngle : int {
    HALF_Q0 = 0,    // every |alpha/2| has reduction index 0: sincos_poly_q0
    HALF_POLY = 1,  // every |alpha/2| <= 64: sincos_poly
    HALF_WIDE = 2,  // sincos_fast (the library sincosf beyond 64)
};

template<int FORM>
__device__ __forceinline__ void half_angle_sincos(float x, float* s, float* c) {
    if constexpr (FORM == HALF_WIDE) {
        dev::sincos_fast(x, s, c);
    } else if constexpr (FORM == HALF_POLY) {
        dev::sincos_poly(x, s, c);
    } else {
        dev::sincos_poly_q0(x, s, c);
    }
}

// F+_0 at TE is taken from the updated state: over TR - TE the transverse
// state only decays and precesses (F+ times d_TRmTE), so F+_0(TE) =
// F+_0(TR) / d_TRmTE, one complex multiply per TR instead of a second 2x3
// operator.  A wave with a lane whose d_TRmTE is too small to invert (T2 of a
// few microseconds) applies the operator D(TE) R to state 0 instead: this
// tells whether the wave must (wave-uniform).
__device__ inline bool fisp_direct_echo(const FispArgs& a, float T2, float B0) {
    const cfloat d_TRmTE = off_resonance(a.TR - a.TE, B0) * relax(a.TR - a.TE, T2);
    const float dd = d_TRmTE.re * d_TRmTE.re + d_TRmTE.im * d_TRmTE.im;
    return __any(!(dd > 1e-30f));
}

// How a wave runs the recursion, decided once per wave (wave-uniform).  The
// half flip angles alpha/2 = t.x B1 / 2 of the wave are bounded by h = 0.5
// fl(max_angle |B1|), since rounding is monotone (|fl(t.x B1)| <=
// fl(max_angle |B1|)).  wide: h may exceed 64, where sincos_fast leaves its
// polynomial; q0: fl(h 2/pi) <= 1/2, every reduction index is 0; simple: one
// slice profile, one repetition, no undersampling, no direct echo and at
// most 2^28 voxels (each TR stores its echo once, at a 32-bit byte offset
// from the TR's row).  max_angle: the wave's fisp_table_max_angle.
struct FispWaveForm {
    bool wide, simple, q0, direct_echo;
};

__device__ inline FispWaveForm fisp_wave_form(const FispArgs& a, int voxel, float max_angle) {
    FispWaveForm f;
    const float h = 0.5f * (max_angle * fabsf(a.tissue.B1(voxel)));
    f.direct_echo = fisp_direct_echo(a, a.tissue.T2(voxel), a.tissue.B0(voxel));
    f.wide = __any(!(h <= 64.0f));
    f.q0 = !__any(!(h * 0.63661977236758134f <= 0.5f));
    f.simple = !f.direct_echo && a.nslices == 1 && a.repetitions == 1 && a.usf == 1 && a.tissue.nvoxels <= (1 << 28);
    return f;
}

}  // namespace kernels
}  // namespace compas
