#pragma once
// Device helpers shared by the HIP kernels (gfx950, wave64).

#include <hip/hip_runtime.h>

#include "compas/compas.h"
#include "compas/core/complex.h"

namespace compas {
namespace dev {

constexpr double TWO_PI = 6.283185307179586476925286766559;
constexpr float TWO_PI_F = 6.2831853071795864769f;
constexpr float DEG2RAD_F = 3.14159265358979323846f / 180.0f;

// Structure-of-arrays tissue map [NUM_FIELDS][nvoxels]; the defaults for absent
// B1 / B0 / z follow TissueParametersView::get (tissue_view.cuh:53-71).
struct TissueView {
    const float* data;
    int nvoxels;
    bool has_z, has_b0, has_b1;

    __device__ float field(int f, int v) const {
        return data[size_t(f) * size_t(nvoxels) + size_t(v)];
    }
    __device__ float T1(int v) const { return field(TissueParameterField::T1, v); }
    __device__ float T2(int v) const { return field(TissueParameterField::T2, v); }
    __device__ float B1(int v) const { return has_b1 ? field(TissueParameterField::B1, v) : 1.0f; }
    __device__ float B0(int v) const { return has_b0 ? field(TissueParameterField::B0, v) : 0.0f; }
    __device__ cfloat rho(int v) const {
        return {field(TissueParameterField::RHO_X, v), field(TissueParameterField::RHO_Y, v)};
    }
    __device__ float x(int v) const { return field(TissueParameterField::X, v); }
    __device__ float y(int v) const { return field(TissueParameterField::Y, v); }
    __device__ float z(int v) const { return has_z ? field(TissueParameterField::Z, v) : 0.0f; }
};

inline TissueView tissue_view(const TissueParameters& p) {
    return {p.data.data(), p.nvoxels, p.has_z, p.has_b0, p.has_b1};
}

// Reduce an FP64 phase to [-pi, pi) and return it as float.  The argument is
// formed in FP64 from the float32 inputs, so large k.r products (hundreds of
// radians) keep full precision before the float sincos.
__device__ inline float reduce_phase(double phi) {
    double k = rint(phi * (1.0 / TWO_PI));
    return float(fma(-k, TWO_PI, phi));
}

// sin and cos of |x| <= 64: Cody-Waite reduction by pi/2 with a
// two-constant split, then the Taylor series of sin and cos on [-pi/4, pi/4]
// evaluated by Horner's rule (truncation below 2e-9), quadrant by the
// reduction index.  No branch, so a caller that has bounded its arguments
// keeps its loop body in one block.
__device__ inline void sincos_poly(float x, float* s, float* c) {
    const float kf = rintf(x * 0.63661977236758134f);
    float r = fmaf(-kf, 1.5707963705062866f, x);  // float(pi/2)
    r = fmaf(-kf, -4.3711388286737929e-08f, r);  // pi/2 - float(pi/2)
    const float r2 = r * r;
    float ps = fmaf(r2, 2.7557319e-6f, -1.9841270e-4f);  // 1/9!, -1/7!
    ps = fmaf(r2, ps, 8.3333333e-3f);
    ps = fmaf(r2, ps, -1.6666667e-1f);
    const float sr = fmaf(r * r2, ps, r);
    float pc = fmaf(r2, -2.7557319e-7f, 2.4801587e-5f);  // -1/10!, 1/8!
    pc = fmaf(r2, pc, -1.3888889e-3f);
    pc = fmaf(r2, pc, 4.1666668e-2f);
    pc = fmaf(r2, pc, -0.5f);
    const float cr = fmaf(r2, pc, 1.0f);
    const int q = int(kf);
    const float ss = (q & 1) ? cr : sr;
    const float cc = (q & 1) ? sr : cr;
    *s = (q & 2) ? -ss : ss;
    *c = ((q + 1) & 2) ? -cc : cc;
}

// sincos_poly of an argument whose reduction index rint(x 2/pi) is 0
// (|fl(x 2/pi)| <= 1/2, |x| <= pi/4): the same polynomial operations on x
// itself, without the reduction and the quadrant selection, so the same
// results (a negative zero argument aside).
__device__ inline void sincos_poly_q0(float x, float* s, float* c) {
    const float r2 = x * x;
    float ps = fmaf(r2, 2.7557319e-6f, -1.9841270e-4f);
    ps = fmaf(r2, ps, 8.3333333e-3f);
    ps = fmaf(r2, ps, -1.6666667e-1f);
    *s = fmaf(x * r2, ps, x);
    float pc = fmaf(r2, -2.7557319e-7f, 2.4801587e-5f);
    pc = fmaf(r2, pc, -1.3888889e-3f);
    pc = fmaf(r2, pc, 4.1666668e-2f);
    pc = fmaf(r2, pc, -0.5f);
    *c = fmaf(r2, pc, 1.0f);
}

// sin and cos of a float argument: sincos_poly, and the library sincosf
// beyond |x| = 64.
__device__ inline void sincos_fast(float x, float* s, float* c) {
    if (fabsf(x) > 64.0f) {
        sincosf(x, s, c);
        return;
    }
    sincos_poly(x, s, c);
}

// sin and cos of |x| <= pi + 1e-3 (a phase already reduced by reduce_phase)
__device__ inline void sincos_reduced(float x, float* s, float* c) {
    sincos_poly(x, s, c);
}

// mag * exp(i phi) with the phase formed in FP64 and reduced before the float sincos
__device__ inline cfloat polar_reduced(float mag, double phi) {
    float s, c;
    sincos_reduced(reduce_phase(phi), &s, &c);
    return {mag * c, mag * s};
}

__device__ inline cfloat polar_f(float mag, float phase) {
    float s, c;
    sincosf(phase, &s, &c);
    return {mag * c, mag * s};
}

__device__ inline cfloat cmul(cfloat a, cfloat b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// a + b * c  with fused multiply-adds
__device__ inline cfloat cfma(cfloat b, cfloat c, cfloat a) {
    return {fmaf(b.re, c.re, fmaf(-b.im, c.im, a.re)), fmaf(b.re, c.im, fmaf(b.im, c.re, a.im))};
}

// a + b * conj(c)
__device__ inline cfloat cfma_conj(cfloat b, cfloat c, cfloat a) {
    return {fmaf(b.re, c.re, fmaf(b.im, c.im, a.re)), fmaf(b.im, c.re, fmaf(-b.re, c.im, a.im))};
}

inline unsigned div_ceil(long long a, long long b) {
    return unsigned((a + b - 1) / b);
}

}  // namespace dev
}  // namespace compas
