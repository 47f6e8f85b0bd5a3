// Forward-mode (tangent) derivatives of the pSSFP isochromat simulation.
//
// The reference differentiates by forward finite differences in FP32
// (src/simulate/derivative.cu:64-94, delta = 1e-4), one more simulation per
// parameter.  This kernel runs the FP64 isochromat recursion of pssfp.hip
// together with its tangents with respect to any subset of the tissue
// parameters {T1, T2, B1, B0}, in one pass: every isochromat m carries one
// dm/dp per parameter through each operator of isochromat.cuh:9-128:
//   relaxation  m.xy *= E2, m.z = E1 m.z + 1 - E1
//               T1: dE1/dT1 = (dt/T1^2) E1;  T2: dE2/dT2 = (dt/T2^2) E2
//   RF sub-step m' = Rot(a) m, a = (Re z B1, -Im z B1, -g z - 2 pi dt B0)
//               dm' = Rot(a) dm + (dRot/da . da/dp) m, where da/dB1 =
//               (Re z, -Im z, 0) and da/dB0 = (0, 0, -2 pi dt_ex)
//   precession  m' = Rz(theta) m, dtheta/dB0 = -2 pi dt_pr
//   inversion   m.z = cos(pi B1) (B1 only)
// With u = |a|^2, Rot(a) m = c m + s1 (a x m) + s2 (a . m) a and
// c = 1 - u s2, so the axis tangents need s1'(u) and s2'(u) besides s1, s2:
// their Taylor series (|a| <= 1/2) or, above, the closed forms
// s1' = (c - s1) / (2u), s2' = (s1 - 2 s2) / (2u).
// The result is the exact derivative of the discrete recursion (no step).
//
// One pass shares among its tangents the primal recursion, the rotation
// coefficients c, s1, s2, s1', s2' and the products a x m, a . m of every
// sub-step; each tangent adds only its own rotation (and, for B1 and B0, the
// axis term).  The pass can also write the echoes it carries, so a caller
// that needs the signal and all its derivatives runs the recursion once.
//
// Layout: as the primal kernel (one wave per block of 64 voxels and group of
// ZPL <= 4 slice positions, partial sums over z added by a second pass).

#include <algorithm>
#include <vector>

#include "compas/device/common.h"
#include "compas/device/pssfp_common.h"
#include "compas/device/tangent_set.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

// Output planes of one launch: group blockIdx.y == 0 writes first[k]; group
// j > 0 writes rest + ((j - 1) * planes + k) * nTR * nvoxels.
struct TangentPlanes {
    cfloat* first[5];
    cfloat* rest;
};

// s1'(u), s2'(u) from the Taylor series (term-wise derivatives, one more
// term each; first omitted terms below 1e-18), u <= SERIES_LIMIT; summed over
// the powers of u like rotation_series_powers
__device__ inline void rotation_series_derivative(double u, double& s1p, double& s2p) {
    const double u2 = u * u, u3 = u2 * u, u4 = u2 * u2, u5 = u4 * u;
    s1p = fma(u, 2.0 / 120.0, -1.0 / 6.0);
    s1p = fma(u2, -3.0 / 5040.0, s1p);
    s1p = fma(u3, 4.0 / 362880.0, s1p);
    s1p = fma(u4, -5.0 / 39916800.0, s1p);
    s1p = fma(u5, 6.0 / 6227020800.0, s1p);
    s2p = fma(u, 2.0 / 720.0, -1.0 / 24.0);
    s2p = fma(u2, -3.0 / 40320.0, s2p);
    s2p = fma(u3, 4.0 / 3628800.0, s2p);
    s2p = fma(u4, -5.0 / 479001600.0, s2p);
    s2p = fma(u5, 6.0 / 87178291200.0, s2p);
}

// the same with one more term each, for u <= SERIES_LIMIT_WIDE (first omitted
// terms below 2e-18)
__device__ inline void rotation_series_derivative_wide(double u, double& s1p, double& s2p) {
    const double u2 = u * u, u3 = u2 * u, u4 = u2 * u2, u5 = u4 * u, u6 = u3 * u3;
    s1p = fma(u, 2.0 / 120.0, -1.0 / 6.0);
    s1p = fma(u2, -3.0 / 5040.0, s1p);
    s1p = fma(u3, 4.0 / 362880.0, s1p);
    s1p = fma(u4, -5.0 / 39916800.0, s1p);
    s1p = fma(u5, 6.0 / 6227020800.0, s1p);
    s1p = fma(u6, -7.0 / 1307674368000.0, s1p);
    s2p = fma(u, 2.0 / 720.0, -1.0 / 24.0);
    s2p = fma(u2, -3.0 / 40320.0, s2p);
    s2p = fma(u3, 4.0 / 3628800.0, s2p);
    s2p = fma(u4, -5.0 / 479001600.0, s2p);
    s2p = fma(u5, 6.0 / 87178291200.0, s2p);
    s2p = fma(u6, -7.0 / 20922789888000.0, s2p);
}

// s1, s2 and s1', s2' of the short tiers (u <= SERIES_LIMIT_4: four terms
// of s1, s2, their term-wise derivatives with one more term; u <=
// SERIES_LIMIT_5: five), summed over the powers of u; first omitted terms
// below 1e-17 relative as for the longer tiers.
__device__ inline void rotation_series_4_powers(double u, double& s1, double& s2) {
    const double u2 = u * u, u3 = u2 * u;
    s1 = fma(u, -1.0 / 6.0, 1.0);
    s1 = fma(u2, 1.0 / 120.0, s1);
    s1 = fma(u3, -1.0 / 5040.0, s1);
    s2 = fma(u, -1.0 / 24.0, 0.5);
    s2 = fma(u2, 1.0 / 720.0, s2);
    s2 = fma(u3, -1.0 / 40320.0, s2);
}

__device__ inline void rotation_series_derivative_4(double u, double& s1p, double& s2p) {
    const double u2 = u * u, u3 = u2 * u;
    s1p = fma(u, 2.0 / 120.0, -1.0 / 6.0);
    s1p = fma(u2, -3.0 / 5040.0, s1p);
    s1p = fma(u3, 4.0 / 362880.0, s1p);
    s2p = fma(u, 2.0 / 720.0, -1.0 / 24.0);
    s2p = fma(u2, -3.0 / 40320.0, s2p);
    s2p = fma(u3, 4.0 / 3628800.0, s2p);
}

__device__ inline void rotation_series_5_powers(double u, double& s1, double& s2) {
    const double u2 = u * u, u3 = u2 * u, u4 = u2 * u2;
    s1 = fma(u, -1.0 / 6.0, 1.0);
    s1 = fma(u2, 1.0 / 120.0, s1);
    s1 = fma(u3, -1.0 / 5040.0, s1);
    s1 = fma(u4, 1.0 / 362880.0, s1);
    s2 = fma(u, -1.0 / 24.0, 0.5);
    s2 = fma(u2, 1.0 / 720.0, s2);
    s2 = fma(u3, -1.0 / 40320.0, s2);
    s2 = fma(u4, 1.0 / 3628800.0, s2);
}

__device__ inline void rotation_series_derivative_5(double u, double& s1p, double& s2p) {
    const double u2 = u * u, u3 = u2 * u, u4 = u2 * u2;
    s1p = fma(u, 2.0 / 120.0, -1.0 / 6.0);
    s1p = fma(u2, -3.0 / 5040.0, s1p);
    s1p = fma(u3, 4.0 / 362880.0, s1p);
    s1p = fma(u4, -5.0 / 39916800.0, s1p);
    s2p = fma(u, 2.0 / 720.0, -1.0 / 24.0);
    s2p = fma(u2, -3.0 / 40320.0, s2p);
    s2p = fma(u3, 4.0 / 3628800.0, s2p);
    s2p = fma(u4, -5.0 / 479001600.0, s2p);
}

enum { TIER_SERIES = 0, TIER_WIDE = 1, TIER_ANY = 2, TIER_4 = 3, TIER_5 = 4 };

// Rot(a) as a matrix: c I + s1 [a]x + s2 a a^T
struct RotationMatrix {
    double r00, r01, r02, r10, r11, r12, r20, r21, r22;
    __device__ Vec3d apply(const Vec3d& v) const {
        return {fma(r00, v.x, fma(r01, v.y, r02 * v.z)), fma(r10, v.x, fma(r11, v.y, r12 * v.z)),
                fma(r20, v.x, fma(r21, v.y, r22 * v.z))};
    }
};

__device__ inline RotationMatrix rotation_matrix(double ax, double ay, double az, double c, double s1, double s2) {
    const double bx = s2 * ax, by = s2 * ay, bz = s2 * az;
    const double pxy = bx * ay, pxz = bx * az, pyz = by * az;
    const double qx = s1 * ax, qy = s1 * ay, qz = s1 * az;
    return {fma(bx, ax, c), pxy - qz, pxz + qy, pxy + qz, fma(by, ay, c), pyz - qx, pxz - qy, pyz + qx, fma(bz, az, c)};
}

// Rot(a) m and the tangents dm[k]: Rot(a) dm[k], plus (dRot/da . da/dp) m
// for B1 (da = (zre, -zim, 0)) and B0 (da = (0, 0, daz)).
// TIER: the four-, five- or six-term series (u <= SERIES_LIMIT_4,
// SERIES_LIMIT_5, SERIES_LIMIT), the seven-term one (u <= SERIES_LIMIT_WIDE),
// or any u (the seven-term series, closed forms above).
template<int MASK, int TIER>
__device__ inline void rotate_all(
    Vec3d& m,
    Vec3d* dm,
    double ax,
    double ay,
    double az,
    double u,
    double zre,
    double zim,
    double daz) {
    using S = TangentSet<MASK>;
    double s1, s2, s1p = 0.0, s2p = 0.0;
    if (TIER == TIER_4) {
        rotation_series_4_powers(u, s1, s2);
        if constexpr (S::axis) rotation_series_derivative_4(u, s1p, s2p);
    } else if (TIER == TIER_5) {
        rotation_series_5_powers(u, s1, s2);
        if constexpr (S::axis) rotation_series_derivative_5(u, s1p, s2p);
    } else if (TIER == TIER_SERIES) {
        rotation_series_powers(u, s1, s2);
        if constexpr (S::axis) rotation_series_derivative(u, s1p, s2p);
    } else if (TIER == TIER_WIDE || u <= SERIES_LIMIT_WIDE) {
        rotation_series_wide_powers(u, s1, s2);
        if constexpr (S::axis) rotation_series_derivative_wide(u, s1p, s2p);
    } else {
        rotation_any(u, s1, s2);
        if constexpr (S::axis) {
            const double c = fma(-u, s2, 1.0);
            s1p = (c - s1) / (2.0 * u);
            s2p = (s1 - 2.0 * s2) / (2.0 * u);
        }
    }
    const double c = fma(-u, s2, 1.0);
    if constexpr (S::nt >= 2 && !S::axis) {
        // several vectors through one rotation: the matrix c I + s1 [a]x +
        // s2 a a^T (18 operations) and 9 per vector, against about 19 per
        // vector for the Rodrigues form
        const RotationMatrix R = rotation_matrix(ax, ay, az, c, s1, s2);
#pragma unroll
        for (int k = 0; k < S::nt; k++) dm[k] = R.apply(dm[k]);
        m = R.apply(m);
        return;
    }
    const double cx = ay * m.z - az * m.y;
    const double cy = az * m.x - ax * m.z;
    const double cz = ax * m.y - ay * m.x;
    const double dot = ax * m.x + ay * m.y + az * m.z;
    const double h = s2 * dot;

    // axis terms, from the state before the rotation:
    // dc m + ds1 (a x m) + s1 (da x m) + d(s2 (a . m)) a + s2 (a . m) da
    Vec3d kb1 {0.0, 0.0, 0.0}, kb0 {0.0, 0.0, 0.0};
    if constexpr (S::b1) {
        const double dax = zre, day = -zim;
        const double du = 2.0 * fma(ax, dax, ay * day);
        const double dc = -fma(u, s2p, s2) * du, ds1 = s1p * du, ds2 = s2p * du;
        const double g = fma(ds2, dot, s2 * fma(dax, m.x, day * m.y));
        kb1.x = dc * m.x + ds1 * cx + s1 * (day * m.z) + g * ax + h * dax;
        kb1.y = dc * m.y + ds1 * cy - s1 * (dax * m.z) + g * ay + h * day;
        kb1.z = dc * m.z + ds1 * cz + s1 * (dax * m.y - day * m.x) + g * az;
    }
    if constexpr (S::b0) {
        const double du = 2.0 * az * daz;
        const double dc = -fma(u, s2p, s2) * du, ds1 = s1p * du, ds2 = s2p * du;
        const double g = fma(ds2, dot, s2 * (daz * m.z));
        kb0.x = dc * m.x + ds1 * cx - s1 * (daz * m.y) + g * ax;
        kb0.y = dc * m.y + ds1 * cy + s1 * (daz * m.x) + g * ay;
        kb0.z = dc * m.z + ds1 * cz + g * az + h * daz;
    }

    if constexpr (S::nt >= 2) {  // with axis tangents: the matrix, plus their terms
        const RotationMatrix R = rotation_matrix(ax, ay, az, c, s1, s2);
#pragma unroll
        for (int k = 0; k < S::nt; k++) {
            Vec3d t = R.apply(dm[k]);
            if (S::b1 && k == S::ib1) {
                t.x += kb1.x;
                t.y += kb1.y;
                t.z += kb1.z;
            }
            if (S::b0 && k == S::ib0) {
                t.x += kb0.x;
                t.y += kb0.y;
                t.z += kb0.z;
            }
            dm[k] = t;
        }
        m = R.apply(m);
        return;
    }
#pragma unroll
    for (int k = 0; k < S::nt; k++) {
        const Vec3d d = dm[k];
        const double ex = ay * d.z - az * d.y;
        const double ey = az * d.x - ax * d.z;
        const double ez = ax * d.y - ay * d.x;
        const double ed = s2 * (ax * d.x + ay * d.y + az * d.z);
        Vec3d t {fma(c, d.x, fma(s1, ex, ed * ax)), fma(c, d.y, fma(s1, ey, ed * ay)),
                 fma(c, d.z, fma(s1, ez, ed * az))};
        if (S::b1 && k == S::ib1) {
            t.x += kb1.x;
            t.y += kb1.y;
            t.z += kb1.z;
        }
        if (S::b0 && k == S::ib0) {
            t.x += kb0.x;
            t.y += kb0.y;
            t.z += kb0.z;
        }
        dm[k] = t;
    }

    m.x = fma(c, m.x, fma(s1, cx, h * ax));
    m.y = fma(c, m.y, fma(s1, cy, h * ay));
    m.z = fma(c, m.z, fma(s1, cz, h * az));
}

// Relaxation and its tangents; dE1 = dE1/dT1 and dE2 = dE2/dT2 over the
// interval (used by the T1 and T2 tangents only).
template<int MASK>
__device__ inline void relax_all(Vec3d& m, Vec3d* dm, double E1, double E2, double R1, double dE1, double dE2) {
    using S = TangentSet<MASK>;
#pragma unroll
    for (int k = 0; k < S::nt; k++) {
        if (S::t2 && k == S::it2) {
            dm[k].x = fma(dm[k].x, E2, m.x * dE2);
            dm[k].y = fma(dm[k].y, E2, m.y * dE2);
        } else {
            dm[k].x *= E2;
            dm[k].y *= E2;
        }
        if (S::t1 && k == S::it1) dm[k].z = fma(dm[k].z, E1, (m.z - 1.0) * dE1);  // d(1 - E1) = -dE1
        else dm[k].z *= E1;
    }
    relax3(m, E1, E2, R1);
}

// Rz(theta) and its tangents; the B0 tangent gains dtheta/dB0 J Rz m
// (J = d/dtheta of a rotation about z, which commutes with Rz).
template<int MASK>
__device__ inline void rotate_z_all(Vec3d& m, Vec3d* dm, double c, double s, double dtheta) {
    using S = TangentSet<MASK>;
    rotate_z(m, c, s);
#pragma unroll
    for (int k = 0; k < S::nt; k++) rotate_z(dm[k], c, s);
    if constexpr (S::b0) {
        dm[S::ib0].x = fma(-dtheta, m.y, dm[S::ib0].x);
        dm[S::ib0].y = fma(dtheta, m.x, dm[S::ib0].y);
    }
}

// grid (ceil(nv / 64), groups in this launch), block 64: one wave per voxel
// block and group of ZPL slice positions, as the primal kernel.
template<int ZPL, int MASK>
__global__ void __launch_bounds__(64) pssfp_iso_tangents(PssfpArgs a, int group0, TangentPlanes out) {
    using S = TangentSet<MASK>;
    constexpr int NT = S::nt, NP = S::planes, OFS = int(S::echo);
    static_assert(NT >= 1, "a tangent pass carries at least one tangent");
    __shared__ double2 grf[RF_CHUNK];
    const int lane = threadIdx.x;
    const int group = group0 + int(blockIdx.y);
    const int voxel = blockIdx.x * 64 + lane;
    const bool active = voxel < a.tissue.nvoxels;
    const int vx = active ? voxel : 0;
    const size_t plane = size_t(a.nTR) * size_t(a.tissue.nvoxels);
    cfloat* dst[NP];
#pragma unroll
    for (int k = 0; k < NP; k++)
        dst[k] = (blockIdx.y == 0) ? out.first[k] : out.rest + (size_t(blockIdx.y - 1) * NP + k) * plane;

    const double T1 = a.tissue.T1(vx), T2 = a.tissue.T2(vx);
    const double B1 = a.tissue.B1(vx), B0 = a.tissue.B0(vx);
    const double dt_ex = a.dt_ex, dt_pr = a.dt_pr, dt_inv = a.dt_inv;
    const double E1_ex = exp(-dt_ex / T1), E2_ex = exp(-dt_ex / T2);
    const double E1_pr = exp(-dt_pr / T1), E2_pr = exp(-dt_pr / T2);
    const double E1_inv = exp(-dt_inv / T1), E2_inv = exp(-dt_inv / T2);
    const double R1_ex = 1.0 - E1_ex, R1_pr = 1.0 - E1_pr, R1_inv = 1.0 - E1_inv;
    const double off_ex = dt_ex * dev::TWO_PI * B0;
    const double off_pr = dt_pr * dev::TWO_PI * B0;
    // d/dT of exp(-dt/T) = (dt/T^2) exp(-dt/T)
    double dE1_ex = 0.0, dE1_pr = 0.0, dE1_inv = 0.0, dE2_ex = 0.0, dE2_pr = 0.0, dE2_inv = 0.0;
    if constexpr (S::t1) {
        const double k = 1.0 / (T1 * T1);
        dE1_ex = dt_ex * k * E1_ex;
        dE1_pr = dt_pr * k * E1_pr;
        dE1_inv = dt_inv * k * E1_inv;
    }
    if constexpr (S::t2) {
        const double k = 1.0 / (T2 * T2);
        dE2_ex = dt_ex * k * E2_ex;
        dE2_pr = dt_pr * k * E2_pr;
        dE2_inv = dt_inv * k * E2_inv;
    }
    const double daz = -dev::TWO_PI * dt_ex;  // d a_z / d B0
    const double dth = -dev::TWO_PI * dt_pr;  // d theta_pr / d B0

    Vec3d m[ZPL], dm[ZPL][NT];
    double az[ZPL], cpr[ZPL], spr[ZPL];
    double az2max = 0.0;  // max over the lane's slice positions of a_z^2
    bool valid[ZPL];
#pragma unroll
    for (int i = 0; i < ZPL; i++) {
        const int iz = group * ZPL + i;
        valid[i] = iz < a.nz;
        const double zz = valid[i] ? double(a.z[iz]) : 0.0;
        az[i] = -(double(a.g_ex) * zz) - off_ex;
        az2max = fmax(az2max, az[i] * az[i]);
        sincos(-(double(a.g_pr) * zz) - off_pr, &spr[i], &cpr[i]);
        m[i] = {0.0, 0.0, cos(M_PI * B1)};
#pragma unroll
        for (int k = 0; k < NT; k++) dm[i][k] = {0.0, 0.0, 0.0};
        if constexpr (S::b1) dm[i][S::ib1].z = -M_PI * sin(M_PI * B1);
        relax_all<MASK>(m[i], dm[i], E1_inv, E2_inv, R1_inv, dE1_inv, dE2_inv);
    }

    auto stage_rf = [&](int k0) {
        for (int k = lane; k < RF_CHUNK && k0 + k < a.nRF; k += 64) {
            const cfloat g = a.gamma_dt_RF[k0 + k];
            grf[k] = make_double2(double(g.re), double(g.im));
        }
    };
    const bool rf_resident = a.nRF <= RF_CHUNK;
    if (rf_resident) stage_rf(0);
    __syncthreads();

    auto excite = [&](double tre, double tim) {
        for (int k0 = 0; k0 < a.nRF; k0 += RF_CHUNK) {
            if (!rf_resident) {
                __syncthreads();
                stage_rf(k0);
                __syncthreads();
            }
            const int kn = min(RF_CHUNK, a.nRF - k0);
            double2 g_next = grf[0];
            for (int k = 0; k < kn; k++) {
                // the next sub-step's waveform sample is read from LDS a step ahead
                const double2 g = g_next;
                if (k + 1 < kn) g_next = grf[k + 1];
                const double zre = tre * g.x - tim * g.y;
                const double zim = tre * g.y + tim * g.x;
                const double ax = zre * B1, ay = -zim * B1;
                const double axy2 = fma(ax, ax, ay * ay);
                double u[ZPL];
#pragma unroll
                for (int i = 0; i < ZPL; i++) u[i] = fma(az[i], az[i], axy2);
                // the lane's largest u: a_z is fixed per slice position
                const double umax = axy2 + az2max;
                // wave-uniform choice of the series, as in the primal kernel
                if (!__any(umax > SERIES_LIMIT_4)) {
#pragma unroll
                    for (int i = 0; i < ZPL; i++) {
                        rotate_all<MASK, TIER_4>(m[i], dm[i], ax, ay, az[i], u[i], zre, zim, daz);
                        relax_all<MASK>(m[i], dm[i], E1_ex, E2_ex, R1_ex, dE1_ex, dE2_ex);
                    }
                } else if (!__any(umax > SERIES_LIMIT_5)) {
#pragma unroll
                    for (int i = 0; i < ZPL; i++) {
                        rotate_all<MASK, TIER_5>(m[i], dm[i], ax, ay, az[i], u[i], zre, zim, daz);
                        relax_all<MASK>(m[i], dm[i], E1_ex, E2_ex, R1_ex, dE1_ex, dE2_ex);
                    }
                } else if (!__any(umax > SERIES_LIMIT)) {
#pragma unroll
                    for (int i = 0; i < ZPL; i++) {
                        rotate_all<MASK, TIER_SERIES>(m[i], dm[i], ax, ay, az[i], u[i], zre, zim, daz);
                        relax_all<MASK>(m[i], dm[i], E1_ex, E2_ex, R1_ex, dE1_ex, dE2_ex);
                    }
                } else if (!__any(umax > SERIES_LIMIT_WIDE)) {
#pragma unroll
                    for (int i = 0; i < ZPL; i++) {
                        rotate_all<MASK, TIER_WIDE>(m[i], dm[i], ax, ay, az[i], u[i], zre, zim, daz);
                        relax_all<MASK>(m[i], dm[i], E1_ex, E2_ex, R1_ex, dE1_ex, dE2_ex);
                    }
                } else {
#pragma unroll
                    for (int i = 0; i < ZPL; i++) {
                        rotate_all<MASK, TIER_ANY>(m[i], dm[i], ax, ay, az[i], u[i], zre, zim, daz);
                        relax_all<MASK>(m[i], dm[i], E1_ex, E2_ex, R1_ex, dE1_ex, dE2_ex);
                    }
                }
            }
        }
    };

    // alpha/2 preparation (pssfp_kernels.cuh:72-75), then the prephaser with twice the gradient
    {
        const cfloat rf0 = a.RF[0];
        excite(-0.5 * double(rf0.re), -0.5 * double(rf0.im));
#pragma unroll
        for (int i = 0; i < ZPL; i++) {
            const int iz = group * ZPL + i;
            const double zz = valid[i] ? double(a.z[iz]) : 0.0;
            double s, c;
            sincos(-(2.0 * double(a.g_pr) * zz) - off_pr, &s, &c);
            rotate_z_all<MASK>(m[i], dm[i], c, s, dth);
            relax_all<MASK>(m[i], dm[i], E1_pr, E2_pr, R1_pr, dE1_pr, dE2_pr);
        }
    }

    for (int n = 0; n < a.nTR; n++) {
        const cfloat rf = a.RF[n];
        excite(double(rf.re), double(rf.im));
        double sx[NP], sy[NP];
#pragma unroll
        for (int k = 0; k < NP; k++) sx[k] = sy[k] = 0.0;
#pragma unroll
        for (int i = 0; i < ZPL; i++) {
            rotate_z_all<MASK>(m[i], dm[i], cpr[i], spr[i], dth);
            relax_all<MASK>(m[i], dm[i], E1_pr, E2_pr, R1_pr, dE1_pr, dE2_pr);
            if (valid[i]) {
                if constexpr (S::echo) {
                    sx[0] += m[i].x;
                    sy[0] += m[i].y;
                }
#pragma unroll
                for (int k = 0; k < NT; k++) {
                    sx[OFS + k] += dm[i][k].x;
                    sy[OFS + k] += dm[i][k].y;
                }
            }
        }
        if (active) {
            const size_t at = size_t(n) * size_t(a.tissue.nvoxels) + size_t(voxel);
#pragma unroll
            for (int k = 0; k < NP; k++) dst[k][at] = {float(sx[k]), float(sy[k])};
        }
#pragma unroll
        for (int i = 0; i < ZPL; i++) {
            rotate_z_all<MASK>(m[i], dm[i], cpr[i], spr[i], dth);
            relax_all<MASK>(m[i], dm[i], E1_pr, E2_pr, R1_pr, dE1_pr, dE2_pr);
        }
    }
}

}  // namespace kernels

namespace {

using kernels::TM_B0;
using kernels::TM_B1;
using kernels::TM_ECHO;
using kernels::TM_T1;
using kernels::TM_T2;

using TangentLauncher = void (*)(kernels::PssfpArgs, int, kernels::TangentPlanes, dim3, hipStream_t);

template<int ZPL, int MASK>
void launch_tangents(kernels::PssfpArgs a, int group0, kernels::TangentPlanes out, dim3 grid, hipStream_t s) {
    kernels::pssfp_iso_tangents<ZPL, MASK><<<grid, 64, 0, s>>>(a, group0, out);
}

template<int ZPL, int MASK>
int resident_waves() {
    static const int cached = [] {
        int n = 0;
        COMPAS_HIP_CHECK(
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernels::pssfp_iso_tangents<ZPL, MASK>, 64, 0));
        return std::max(1, n);
    }();
    return cached;
}

// ZPL and launch layout of one pass; the partial planes of a group hold all
// `planes` outputs, so the layout's memory cap sees planes * nTR.  With four
// tangents ZPL stops at 2: at 3 and 4 the state (ZPL x 15 doubles) no longer
// fits the 256 VGPRs and the compiler parks it in AGPRs.
template<int MASK>
void tangent_layout(const CompasContext& ctx, int nz, int nv, int nTR, PssfpLayout* L, TangentLauncher* fn) {
    constexpr bool wide = kernels::TangentSet<MASK>::nt >= 4;
    int waves[5] = {0, resident_waves<1, MASK>(), resident_waves<2, MASK>(), 0, 0};
    if constexpr (!wide) {
        waves[3] = resident_waves<3, MASK>();
        waves[4] = resident_waves<4, MASK>();
    }
    *L = pssfp_layout(ctx, nz, nv, nTR * kernels::TangentSet<MASK>::planes, waves, wide ? 2 : 4);
    if constexpr (wide) {
        *fn = L->zpl == 1 ? &launch_tangents<1, MASK> : &launch_tangents<2, MASK>;
    } else {
        switch (L->zpl) {
            case 1: *fn = &launch_tangents<1, MASK>; break;
            case 2: *fn = &launch_tangents<2, MASK>; break;
            case 3: *fn = &launch_tangents<3, MASK>; break;
            default: *fn = &launch_tangents<4, MASK>; break;
        }
    }
}

// The field sets with a kernel of their own: each field alone, T1 + T2 (the
// four-parameter products) and all four (the six-parameter products), the
// last two also with the echoes.
bool layout_for_mask(int mask, const CompasContext& ctx, int nz, int nv, int nTR, PssfpLayout* L, TangentLauncher* fn) {
    switch (mask) {
        case TM_T1: tangent_layout<TM_T1>(ctx, nz, nv, nTR, L, fn); return true;
        case TM_T2: tangent_layout<TM_T2>(ctx, nz, nv, nTR, L, fn); return true;
        case TM_B1: tangent_layout<TM_B1>(ctx, nz, nv, nTR, L, fn); return true;
        case TM_B0: tangent_layout<TM_B0>(ctx, nz, nv, nTR, L, fn); return true;
        case TM_T1 | TM_T2: tangent_layout<TM_T1 | TM_T2>(ctx, nz, nv, nTR, L, fn); return true;
        case TM_T1 | TM_T2 | TM_ECHO: tangent_layout<TM_T1 | TM_T2 | TM_ECHO>(ctx, nz, nv, nTR, L, fn); return true;
        case TM_T1 | TM_T2 | TM_B1 | TM_B0: tangent_layout<TM_T1 | TM_T2 | TM_B1 | TM_B0>(ctx, nz, nv, nTR, L, fn); return true;
        case TM_T1 | TM_T2 | TM_B1 | TM_B0 | TM_ECHO:
            tangent_layout<TM_T1 | TM_T2 | TM_B1 | TM_B0 | TM_ECHO>(ctx, nz, nv, nTR, L, fn);
            return true;
        default: return false;
    }
}

// Passes that cover `mask` exactly, each with a kernel of its own.
std::vector<int> split_mask(int mask) {
    const int fields = mask & ~TM_ECHO;
    const int echo = mask & TM_ECHO;
    if (fields == (TM_T1 | TM_T2 | TM_B1 | TM_B0)) return {mask};
    std::vector<int> passes;
    int rest = fields;
    if ((fields & (TM_T1 | TM_T2)) == (TM_T1 | TM_T2)) {
        passes.push_back(TM_T1 | TM_T2 | echo);
        rest &= ~(TM_T1 | TM_T2);
    } else if (echo) {
        passes.push_back(TM_ECHO);  // the primal kernel
    }
    for (int bit : {TM_T1, TM_T2, TM_B1, TM_B0})
        if (rest & bit) passes.push_back(bit);
    return passes;
}

// One pass of a mask with a kernel: planes[k] in the order of TangentSet.
void run_pass(const CompasContext& ctx, int mask, const kernels::PssfpArgs& a, cfloat* const* planes) {
    const int nz = a.nz, nv = a.tissue.nvoxels, nTR = a.nTR;
    const int np = __builtin_popcount(unsigned(mask));
    PssfpLayout L;
    TangentLauncher fn;
    if (!layout_for_mask(mask, ctx, nz, nv, nTR, &L, &fn)) COMPAS_ERROR("no tangent kernel for this field set");
    const int64_t plane = int64_t(nTR) * nv;
    const int slots = L.groups > L.per_launch ? L.per_launch : L.groups - 1;
    Array<cfloat, 2> part;  // slot s, plane k at (s * np + k) * plane
    if (slots > 0) part = ctx.empty<cfloat, 2>({int64_t(slots) * np, plane});
    const unsigned vblocks = unsigned(dev::div_ceil(nv, 64));
    for (int g0 = 0; g0 < L.groups; g0 += L.per_launch) {
        const int count = std::min(L.per_launch, L.groups - g0);
        kernels::TangentPlanes out {};
        for (int k = 0; k < np; k++) out.first[k] = g0 == 0 ? planes[k] : part.data() + k * plane;
        out.rest = g0 == 0 ? part.data() : part.data() + np * plane;
        {
            KernelScope ks(ctx, "pssfp_iso_tangents");
            fn(a, g0, out, dim3(vblocks, unsigned(count)), ctx.stream());
            check_launch("pssfp_iso_tangents");
        }
        const int summed = g0 == 0 ? count - 1 : count;
        for (int k = 0; k < np; k++)
            pssfp_sum_groups(ctx, planes[k], part.data() + k * plane, summed, plane, np * plane);
    }
}

}  // namespace

std::vector<Array<cfloat, 2>> simulate_magnetization_tangents(
    const CompasContext& ctx,
    const std::vector<int>& fields,
    TissueParameters parameters,
    pSSFPSequence seq,
    Array<cfloat, 2>* echos) {
    const int nTR = int(seq.RF_train.size(0));
    const int nv = parameters.nvoxels;
    const int nz = int(seq.z.size(0));
    COMPAS_CHECK(nTR > 0);
    int mask = 0;
    for (int f : fields) {
        int bit = 0;
        switch (f) {
            case TissueParameterField::T1: bit = TM_T1; break;
            case TissueParameterField::T2: bit = TM_T2; break;
            case TissueParameterField::B1: bit = TM_B1; break;
            case TissueParameterField::B0: bit = TM_B0; break;
            default: COMPAS_ERROR("tangent derivatives are available for T1, T2, B1 and B0");
        }
        if (mask & bit) COMPAS_ERROR("a field is listed twice");
        mask |= bit;
    }
    // (with no voxels every field reads as absent: the empty result is returned below)
    if (nv > 0 && (((mask & TM_B1) && !parameters.has_b1) || ((mask & TM_B0) && !parameters.has_b0)))
        COMPAS_ERROR("cannot differentiate with respect to an absent parameter");
    ctx.activate();
    std::vector<Array<cfloat, 2>> d;
    for (size_t i = 0; i < fields.size(); i++) d.push_back(ctx.empty<cfloat, 2>({nTR, nv}));
    if (echos) *echos = ctx.empty<cfloat, 2>({nTR, nv});
    if (nv == 0) return d;
    if (nz == 0) {
        for (auto& x : d) ctx.fill_zero(x.data(), x.nbytes());
        if (echos) ctx.fill_zero(echos->data(), echos->nbytes());
        return d;
    }

    // the output array of each bit
    cfloat* by_bit[32] = {};
    for (size_t i = 0; i < fields.size(); i++) {
        const int bit = fields[i] == TissueParameterField::T1 ? TM_T1
            : fields[i] == TissueParameterField::T2           ? TM_T2
            : fields[i] == TissueParameterField::B1           ? TM_B1
                                                              : TM_B0;
        by_bit[bit] = d[i].data();
    }
    if (echos) by_bit[TM_ECHO] = echos->data();

    kernels::PssfpArgs a;
    a.tissue = dev::tissue_view(parameters);
    a.RF = seq.RF_train.data();
    a.gamma_dt_RF = seq.gamma_dt_RF.data();
    a.z = seq.z.data();
    a.nz = nz;
    a.nTR = nTR;
    a.nRF = int(seq.gamma_dt_RF.size(0));
    a.dt_ex = seq.dt.ex;
    a.dt_inv = seq.dt.inv;
    a.dt_pr = seq.dt.pr;
    a.g_ex = seq.gamma_dt_GRz.ex;
    a.g_pr = seq.gamma_dt_GRz.pr;
    a.echos = nullptr;

    for (int pass : split_mask(mask | (echos ? TM_ECHO : 0))) {
        if (pass == TM_ECHO) {
            *echos = simulate_magnetization(ctx, parameters, seq);
            continue;
        }
        cfloat* planes[5];
        int np = 0;
        for (int bit : {TM_ECHO, TM_T1, TM_T2, TM_B1, TM_B0})
            if (pass & bit) planes[np++] = by_bit[bit];
        run_pass(ctx, pass, a, planes);
    }
    return d;
}

Array<cfloat, 2> simulate_magnetization_tangent(
    const CompasContext& ctx, int field, TissueParameters parameters, pSSFPSequence seq) {
    return simulate_magnetization_tangents(ctx, std::vector<int> {field}, parameters, seq, nullptr)[0];
}

}  // namespace compas
