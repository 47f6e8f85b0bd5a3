#pragma once
// Pieces shared by the pSSFP isochromat kernels (primal and tangent).
//
// Semantics: reference src/operators/isochromat.cuh:9-128.  Rodrigues'
// rotation is evaluated without sqrt/rcp/sin/cos for the small per-step
// angles: with u = |a|^2, s1(u) = sin(t)/t and s2(u) = (1 - cos t)/t^2 are
// Taylor polynomials in u (|t| <= 1/4: 6 terms, error < 1e-17; |t| <= 1/2:
// 7 terms, error < 5e-17; Horner form), cos t = 1 - u s2(u), and
//   m' = cos t m + s1 (a x m) + s2 (a . m) a,
// which equals the reference's normalised form.  Larger angles use FP64 sincos.

#include "compas/device/common.h"

namespace compas {
namespace kernels {

struct PssfpArgs {
    dev::TissueView tissue;
    const cfloat* RF;  // [nTR]
    const cfloat* gamma_dt_RF;  // [nRF]
    const float* z;  // [nz]
    int nTR, nRF, nz;
    float dt_ex, dt_inv, dt_pr;
    float g_ex, g_pr;
    cfloat* echos;  // [nTR][nvoxels]
};

// A pSSFP simulation in a work table (request server): group 0 of a voxel
// block writes a.echos, group g > 0 writes partial + (g - 1) * nTR * nv.
struct PssfpTableJob {
    PssfpArgs a;
    cfloat* partial;
};

struct Vec3d {
    double x, y, z;
};

// Angles up to sqrt(SERIES_LIMIT) = 1/4 rad use the six-term Taylor series,
// up to sqrt(SERIES_LIMIT_WIDE) = 1/2 rad the seven-term one (first omitted
// terms 1e-17 and 5e-17 relative); larger angles FP64 sincos.
constexpr double SERIES_LIMIT = 0.0625;
constexpr double SERIES_LIMIT_WIDE = 0.25;

// Shorter series for the small angles of most RF sub-steps: four terms up to
// sqrt(SERIES_LIMIT_4) = 1/27 rad, five up to sqrt(SERIES_LIMIT_5) = 1/9 rad
// (first omitted terms u^4/9! <= 9.8e-18 and u^5/11! <= 7.2e-18, relative).
constexpr double SERIES_LIMIT_4 = 1.0 / 729.0;
constexpr double SERIES_LIMIT_5 = 1.0 / 81.0;

__device__ inline void rotation_series_4(double u, double& s1, double& s2) {
    s1 = fma(u, -1.0 / 5040.0, 1.0 / 120.0);
    s1 = fma(u, s1, -1.0 / 6.0);
    s1 = fma(u, s1, 1.0);
    s2 = fma(u, -1.0 / 40320.0, 1.0 / 720.0);
    s2 = fma(u, s2, -1.0 / 24.0);
    s2 = fma(u, s2, 0.5);
}

__device__ inline void rotation_series_5(double u, double& s1, double& s2) {
    s1 = fma(u, 1.0 / 362880.0, -1.0 / 5040.0);
    s1 = fma(u, s1, 1.0 / 120.0);
    s1 = fma(u, s1, -1.0 / 6.0);
    s1 = fma(u, s1, 1.0);
    s2 = fma(u, 1.0 / 3628800.0, -1.0 / 40320.0);
    s2 = fma(u, s2, 1.0 / 720.0);
    s2 = fma(u, s2, -1.0 / 24.0);
    s2 = fma(u, s2, 0.5);
}

// s1 = sin(t)/t and s2 = (1 - cos t)/t^2 of t = sqrt(u), u <= SERIES_LIMIT:
// Taylor series in u, six terms each (first omitted terms below 1e-17), in
// Horner form: five FMAs per series and no powers of u; the ZPL isochromats
// of a lane give independent chains that hide the FMA latency.
__device__ inline void rotation_series(double u, double& s1, double& s2) {
    s1 = fma(u, -1.0 / 39916800.0, 1.0 / 362880.0);
    s1 = fma(u, s1, -1.0 / 5040.0);
    s1 = fma(u, s1, 1.0 / 120.0);
    s1 = fma(u, s1, -1.0 / 6.0);
    s1 = fma(u, s1, 1.0);
    s2 = fma(u, -1.0 / 479001600.0, 1.0 / 3628800.0);
    s2 = fma(u, s2, -1.0 / 40320.0);
    s2 = fma(u, s2, 1.0 / 720.0);
    s2 = fma(u, s2, -1.0 / 24.0);
    s2 = fma(u, s2, 0.5);
}

// the same with one more term each, for u <= SERIES_LIMIT_WIDE
__device__ inline void rotation_series_wide(double u, double& s1, double& s2) {
    s1 = fma(u, 1.0 / 6227020800.0, -1.0 / 39916800.0);
    s1 = fma(u, s1, 1.0 / 362880.0);
    s1 = fma(u, s1, -1.0 / 5040.0);
    s1 = fma(u, s1, 1.0 / 120.0);
    s1 = fma(u, s1, -1.0 / 6.0);
    s1 = fma(u, s1, 1.0);
    s2 = fma(u, 1.0 / 87178291200.0, -1.0 / 479001600.0);
    s2 = fma(u, s2, 1.0 / 3628800.0);
    s2 = fma(u, s2, -1.0 / 40320.0);
    s2 = fma(u, s2, 1.0 / 720.0);
    s2 = fma(u, s2, -1.0 / 24.0);
    s2 = fma(u, s2, 0.5);
}

// The same two series summed over the shared powers of u (4 extra multiplies,
// a shorter dependency chain and fewer live values): the form of the tangent
// passes, whose state already fills the registers and where the Horner
// chains raise the register count past 256.
__device__ inline void rotation_series_powers(double u, double& s1, double& s2) {
    const double u2 = u * u, u3 = u2 * u, u4 = u2 * u2, u5 = u4 * u;
    s1 = fma(u, -1.0 / 6.0, 1.0);
    s1 = fma(u2, 1.0 / 120.0, s1);
    s1 = fma(u3, -1.0 / 5040.0, s1);
    s1 = fma(u4, 1.0 / 362880.0, s1);
    s1 = fma(u5, -1.0 / 39916800.0, s1);
    s2 = fma(u, -1.0 / 24.0, 0.5);
    s2 = fma(u2, 1.0 / 720.0, s2);
    s2 = fma(u3, -1.0 / 40320.0, s2);
    s2 = fma(u4, 1.0 / 3628800.0, s2);
    s2 = fma(u5, -1.0 / 479001600.0, s2);
}

__device__ inline void rotation_series_wide_powers(double u, double& s1, double& s2) {
    const double u2 = u * u, u3 = u2 * u, u4 = u2 * u2, u5 = u4 * u, u6 = u3 * u3;
    s1 = fma(u, -1.0 / 6.0, 1.0);
    s1 = fma(u2, 1.0 / 120.0, s1);
    s1 = fma(u3, -1.0 / 5040.0, s1);
    s1 = fma(u4, 1.0 / 362880.0, s1);
    s1 = fma(u5, -1.0 / 39916800.0, s1);
    s1 = fma(u6, 1.0 / 6227020800.0, s1);
    s2 = fma(u, -1.0 / 24.0, 0.5);
    s2 = fma(u2, 1.0 / 720.0, s2);
    s2 = fma(u3, -1.0 / 40320.0, s2);
    s2 = fma(u4, 1.0 / 3628800.0, s2);
    s2 = fma(u5, -1.0 / 479001600.0, s2);
    s2 = fma(u6, 1.0 / 87178291200.0, s2);
}

// the same for any u: the series up to SERIES_LIMIT_WIDE, FP64 sincos above
__device__ inline void rotation_any(double u, double& s1, double& s2) {
    if (u <= SERIES_LIMIT_WIDE) {
        rotation_series_wide(u, s1, s2);
    } else {
        const double t = sqrt(u);
        double s, c;
        sincos(t, &s, &c);
        s1 = s / t;
        s2 = (1.0 - c) / u;
    }
}

// m' = cos t m + s1 (a x m) + s2 (a . m) a, with cos t = 1 - u s2
__device__ inline void apply_rotation(Vec3d& m, double ax, double ay, double az, double u, double s1, double s2) {
    const double c = fma(-u, s2, 1.0);
    const double cx = ay * m.z - az * m.y;
    const double cy = az * m.x - ax * m.z;
    const double cz = ax * m.y - ay * m.x;
    const double d = s2 * (ax * m.x + ay * m.y + az * m.z);
    m.x = fma(c, m.x, fma(s1, cx, d * ax));
    m.y = fma(c, m.y, fma(s1, cy, d * ay));
    m.z = fma(c, m.z, fma(s1, cz, d * az));
}

__device__ inline void relax3(Vec3d& m, double E1, double E2, double R1) {
    m.x *= E2;
    m.y *= E2;
    m.z = fma(m.z, E1, R1);
}

__device__ inline void rotate_z(Vec3d& m, double c, double s) {
    const double x = c * m.x - s * m.y;
    const double y = c * m.y + s * m.x;
    m.x = x;
    m.y = y;
}

constexpr int RF_CHUNK = 256;  // RF sub-steps staged in LDS at a time

}  // namespace kernels
}  // namespace compas
