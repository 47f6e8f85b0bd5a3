// Forward-mode (tangent) derivatives of the pSSFP isochromat simulation.
//
// The reference differentiates by forward finite differences in FP32
// (src/simulate/derivative.cu:64-94, delta = 1e-4).  This kernel runs the FP64
// isochromat recursion of pssfp.hip together with its tangent with respect to
// one tissue parameter p in {T1, T2, B1, B0}: every isochromat m carries dm/dp
// through each operator of isochromat.cuh:9-128:
//   relaxation  m.xy *= E2, m.z = E1 m.z + 1 - E1
//               T1: dE1/dT1 = (dt/T1^2) E1;  T2: dE2/dT2 = (dt/T2^2) E2
//   RF sub-step m' = Rot(a) m, a = (Re z B1, -Im z B1, -g z - 2 pi dt B0)
//               dm' = Rot(a) dm + (dRot/da . da/dp) m, where da/dB1 =
//               (Re z, -Im z, 0) and da/dB0 = (0, 0, -2 pi dt_ex)
//   precession  m' = Rz(theta) m, dtheta/dB0 = -2 pi dt_pr
//   inversion   m.z = cos(pi B1) (B1 only)
// With u = |a|^2, Rot(a) m = c m + s1 (a x m) + s2 (a . m) a and
// c = 1 - u s2, so the axis tangent needs s1'(u) and s2'(u) besides s1, s2:
// their Taylor series (|a| <= 1/4) or, above, the closed forms
// s1' = (c - s1) / (2u), s2' = (s1 - 2 s2) / (2u).
// The result is the exact derivative of the discrete recursion (no step).
//
// Layout: as the primal kernel (64 voxels per workgroup, one wave per group
// of ZPL slice positions, sum over z through LDS once per TR).

#include "compas/device/common.h"
#include "compas/device/pssfp_common.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

enum { PTAN_T1 = 0, PTAN_T2 = 1, PTAN_B1 = 2, PTAN_B0 = 3 };

// Rot(a) m and its tangent: Rot(a) dm, plus (dRot/da . da) m when AXIS.
template<bool AXIS>
__device__ inline void rotate_tangent(
    Vec3d& m,
    Vec3d& dm,
    double ax,
    double ay,
    double az,
    double axy2,
    double dax,
    double day,
    double daz) {
    const double u = fma(az, az, axy2);
    double s1, s2, s1p = 0.0, s2p = 0.0;
    if (u <= 0.0625) {
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
        if constexpr (AXIS) {
            // term-wise derivatives, one more term each (first omitted < 1e-18)
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
    } else {
        const double t = sqrt(u);
        double s, c;
        sincos(t, &s, &c);
        s1 = s / t;
        s2 = (1.0 - c) / u;
        if constexpr (AXIS) {
            s1p = (c - s1) / (2.0 * u);
            s2p = (s1 - 2.0 * s2) / (2.0 * u);
        }
    }
    const double c = fma(-u, s2, 1.0);

    // tangent of the state through the same rotation
    const double ex = ay * dm.z - az * dm.y;
    const double ey = az * dm.x - ax * dm.z;
    const double ez = ax * dm.y - ay * dm.x;
    const double ed = s2 * (ax * dm.x + ay * dm.y + az * dm.z);
    double tx = fma(c, dm.x, fma(s1, ex, ed * ax));
    double ty = fma(c, dm.y, fma(s1, ey, ed * ay));
    double tz = fma(c, dm.z, fma(s1, ez, ed * az));

    const double cx = ay * m.z - az * m.y;
    const double cy = az * m.x - ax * m.z;
    const double cz = ax * m.y - ay * m.x;
    const double dot = ax * m.x + ay * m.y + az * m.z;
    if constexpr (AXIS) {
        const double du = 2.0 * (ax * dax + ay * day + az * daz);
        const double dc = -(s2 + u * s2p) * du;
        const double ds1 = s1p * du, ds2 = s2p * du;
        // (da x m), (da . m)
        const double fx = day * m.z - daz * m.y;
        const double fy = daz * m.x - dax * m.z;
        const double fz = dax * m.y - day * m.x;
        const double fd = dax * m.x + day * m.y + daz * m.z;
        const double g = fma(ds2, dot, s2 * fd);  // d(s2 (a . m))
        const double h = s2 * dot;
        tx += dc * m.x + ds1 * cx + s1 * fx + g * ax + h * dax;
        ty += dc * m.y + ds1 * cy + s1 * fy + g * ay + h * day;
        tz += dc * m.z + ds1 * cz + s1 * fz + g * az + h * daz;
    }
    dm = {tx, ty, tz};

    const double d = s2 * dot;
    m.x = fma(c, m.x, fma(s1, cx, d * ax));
    m.y = fma(c, m.y, fma(s1, cy, d * ay));
    m.z = fma(c, m.z, fma(s1, cz, d * az));
}

// Relaxation and its tangent; dE1/dE2 are the parameter derivatives of E1/E2
// (zero for the fields that do not enter them).
template<int FIELD>
__device__ inline void relax_tangent(Vec3d& m, Vec3d& dm, double E1, double E2, double R1, double dE) {
    if constexpr (FIELD == PTAN_T2) {
        dm.x = fma(dm.x, E2, m.x * dE);
        dm.y = fma(dm.y, E2, m.y * dE);
    } else {
        dm.x *= E2;
        dm.y *= E2;
    }
    if constexpr (FIELD == PTAN_T1) dm.z = fma(dm.z, E1, (m.z - 1.0) * dE);  // d(1 - E1) = -dE1
    else dm.z *= E1;
    relax3(m, E1, E2, R1);
}

// Rz(theta) and its tangent (dtheta = 0 unless B0)
template<int FIELD>
__device__ inline void rotate_z_tangent(Vec3d& m, Vec3d& dm, double c, double s, double dtheta) {
    rotate_z(m, c, s);
    rotate_z(dm, c, s);
    if constexpr (FIELD == PTAN_B0) {
        dm.x = fma(-dtheta, m.y, dm.x);
        dm.y = fma(dtheta, m.x, dm.y);
    }
}

template<int ZPL, int FIELD>
__global__ void __launch_bounds__(512) pssfp_iso_tangent(PssfpArgs a, cfloat* __restrict__ dechos) {
    __shared__ double2 partial[2][8][64];
    __shared__ double2 grf[RF_CHUNK];
    const int lane = threadIdx.x & 63;
    const int group = threadIdx.x >> 6;
    const int groups = blockDim.x >> 6;
    const int voxel = blockIdx.x * 64 + lane;
    const bool active = voxel < a.tissue.nvoxels;
    const int vx = active ? voxel : 0;

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
    double dE_ex = 0.0, dE_pr = 0.0, dE_inv = 0.0;
    if constexpr (FIELD == PTAN_T1) {
        const double k = 1.0 / (T1 * T1);
        dE_ex = dt_ex * k * E1_ex;
        dE_pr = dt_pr * k * E1_pr;
        dE_inv = dt_inv * k * E1_inv;
    }
    if constexpr (FIELD == PTAN_T2) {
        const double k = 1.0 / (T2 * T2);
        dE_ex = dt_ex * k * E2_ex;
        dE_pr = dt_pr * k * E2_pr;
        dE_inv = dt_inv * k * E2_inv;
    }
    const double daz = -dev::TWO_PI * dt_ex;  // d a_z / d B0
    const double dth = -dev::TWO_PI * dt_pr;  // d theta_pr / d B0

    Vec3d m[ZPL], dm[ZPL];
    double az[ZPL], cpr[ZPL], spr[ZPL];
    bool valid[ZPL];
#pragma unroll
    for (int i = 0; i < ZPL; i++) {
        const int iz = group * ZPL + i;
        valid[i] = iz < a.nz;
        const double zz = valid[i] ? double(a.z[iz]) : 0.0;
        az[i] = -(double(a.g_ex) * zz) - off_ex;
        sincos(-(double(a.g_pr) * zz) - off_pr, &spr[i], &cpr[i]);
        m[i] = {0.0, 0.0, cos(M_PI * B1)};
        dm[i] = {0.0, 0.0, FIELD == PTAN_B1 ? -M_PI * sin(M_PI * B1) : 0.0};
        relax_tangent<FIELD>(m[i], dm[i], E1_inv, E2_inv, R1_inv, dE_inv);
    }

    auto stage_rf = [&](int k0) {
        for (int k = threadIdx.x; k < RF_CHUNK && k0 + k < a.nRF; k += blockDim.x) {
            const cfloat g = a.gamma_dt_RF[k0 + k];
            grf[k] = make_double2(double(g.re), double(g.im));
        }
    };
    const bool rf_resident = a.nRF <= RF_CHUNK;
    if (rf_resident) stage_rf(0);
    __syncthreads();

    constexpr bool AXIS = (FIELD == PTAN_B1 || FIELD == PTAN_B0);
    auto excite = [&](double tre, double tim) {
        for (int k0 = 0; k0 < a.nRF; k0 += RF_CHUNK) {
            if (!rf_resident) {
                __syncthreads();
                stage_rf(k0);
                __syncthreads();
            }
            const int kn = min(RF_CHUNK, a.nRF - k0);
            for (int k = 0; k < kn; k++) {
                const double2 g = grf[k];
                const double zre = tre * g.x - tim * g.y;
                const double zim = tre * g.y + tim * g.x;
                const double ax = zre * B1, ay = -zim * B1;
                const double axy2 = fma(ax, ax, ay * ay);
                const double dax = (FIELD == PTAN_B1) ? zre : 0.0;
                const double day = (FIELD == PTAN_B1) ? -zim : 0.0;
                const double dz = (FIELD == PTAN_B0) ? daz : 0.0;
#pragma unroll
                for (int i = 0; i < ZPL; i++) {
                    rotate_tangent<AXIS>(m[i], dm[i], ax, ay, az[i], axy2, dax, day, dz);
                    relax_tangent<FIELD>(m[i], dm[i], E1_ex, E2_ex, R1_ex, dE_ex);
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
            rotate_z_tangent<FIELD>(m[i], dm[i], c, s, dth);
            relax_tangent<FIELD>(m[i], dm[i], E1_pr, E2_pr, R1_pr, dE_pr);
        }
    }

    for (int n = 0; n < a.nTR; n++) {
        const cfloat rf = a.RF[n];
        excite(double(rf.re), double(rf.im));
        double sx = 0.0, sy = 0.0;
#pragma unroll
        for (int i = 0; i < ZPL; i++) {
            rotate_z_tangent<FIELD>(m[i], dm[i], cpr[i], spr[i], dth);
            relax_tangent<FIELD>(m[i], dm[i], E1_pr, E2_pr, R1_pr, dE_pr);
            if (valid[i]) {
                sx += dm[i].x;
                sy += dm[i].y;
            }
        }
        partial[n & 1][group][lane] = make_double2(sx, sy);
        __syncthreads();
        if (group == 0 && active) {
            double tx = 0.0, ty = 0.0;
            for (int g = 0; g < groups; g++) {
                const double2 p = partial[n & 1][g][lane];
                tx += p.x;
                ty += p.y;
            }
            dechos[size_t(n) * size_t(a.tissue.nvoxels) + voxel] = {float(tx), float(ty)};
        }
#pragma unroll
        for (int i = 0; i < ZPL; i++) {
            rotate_z_tangent<FIELD>(m[i], dm[i], cpr[i], spr[i], dth);
            relax_tangent<FIELD>(m[i], dm[i], E1_pr, E2_pr, R1_pr, dE_pr);
        }
    }
}

}  // namespace kernels

namespace {

using PssfpTangentLauncher = void (*)(kernels::PssfpArgs, cfloat*, unsigned, unsigned, hipStream_t);

template<int ZPL, int FIELD>
void launch_pssfp_tangent(kernels::PssfpArgs a, cfloat* d, unsigned blocks, unsigned threads, hipStream_t s) {
    kernels::pssfp_iso_tangent<ZPL, FIELD><<<blocks, threads, 0, s>>>(a, d);
}

template<int FIELD>
PssfpTangentLauncher pssfp_tangent_launcher(int zpl) {
    switch (zpl) {
        case 1: return &launch_pssfp_tangent<1, FIELD>;
        case 2: return &launch_pssfp_tangent<2, FIELD>;
        case 3: return &launch_pssfp_tangent<3, FIELD>;
        default: return &launch_pssfp_tangent<4, FIELD>;
    }
}

// Slice positions per lane (1..4, primal + tangent state of four isochromats
// fit the register file) with at most 8 waves per workgroup: the least padding.
int choose_tangent_zpl(int nz) {
    int best = 0, best_pad = 1 << 30;
    for (int zpl = 4; zpl >= 1; zpl--) {
        const int g = (nz + zpl - 1) / zpl;
        if (g > 8) continue;
        const int pad = g * zpl - nz;
        if (pad < best_pad) {
            best_pad = pad;
            best = zpl;
        }
    }
    return best;
}

}  // namespace

Array<cfloat, 2> simulate_magnetization_tangent(
    const CompasContext& ctx, int field, TissueParameters parameters, pSSFPSequence seq) {
    const int nTR = int(seq.RF_train.size(0));
    const int nv = parameters.nvoxels;
    const int nz = int(seq.z.size(0));
    COMPAS_CHECK(nTR > 0);
    if (field != TissueParameterField::T1 && field != TissueParameterField::T2 &&
        field != TissueParameterField::B1 && field != TissueParameterField::B0)
        COMPAS_ERROR("tangent derivatives are available for T1, T2, B1 and B0");
    if ((field == TissueParameterField::B1 && !parameters.has_b1) ||
        (field == TissueParameterField::B0 && !parameters.has_b0))
        COMPAS_ERROR("cannot differentiate with respect to an absent parameter");
    ctx.activate();
    auto d = ctx.empty<cfloat, 2>({nTR, nv});
    if (nv == 0) return d;
    if (nz == 0) {
        ctx.fill_zero(d.data(), d.nbytes());
        return d;
    }

    kernels::PssfpArgs a;
    a.tissue = dev::tissue_view(parameters);
    a.RF = seq.RF_train.data();
    a.gamma_dt_RF = seq.gamma_dt_RF.data();
    a.nTR = nTR;
    a.nRF = int(seq.gamma_dt_RF.size(0));
    a.dt_ex = seq.dt.ex;
    a.dt_inv = seq.dt.inv;
    a.dt_pr = seq.dt.pr;
    a.g_ex = seq.gamma_dt_GRz.ex;
    a.g_pr = seq.gamma_dt_GRz.pr;
    a.echos = nullptr;

    // more than 32 slice positions: passes of <= 32 accumulate into a scratch sum
    const int chunk = 32;  // 8 waves x 4 isochromats per lane
    Array<cfloat, 2> part;
    for (int z0 = 0; z0 < nz; z0 += chunk) {
        const int nzc = std::min(chunk, nz - z0);
        a.z = seq.z.data() + z0;
        a.nz = nzc;
        const int zpl = choose_tangent_zpl(nzc);
        const int groups = (nzc + zpl - 1) / zpl;
        cfloat* out = d.data();
        if (z0 > 0) {
            if (part.is_empty()) part = ctx.empty<cfloat, 2>({nTR, nv});
            out = part.data();
        }
        PssfpTangentLauncher fn;
        switch (field) {
            case TissueParameterField::T1: fn = pssfp_tangent_launcher<kernels::PTAN_T1>(zpl); break;
            case TissueParameterField::T2: fn = pssfp_tangent_launcher<kernels::PTAN_T2>(zpl); break;
            case TissueParameterField::B1: fn = pssfp_tangent_launcher<kernels::PTAN_B1>(zpl); break;
            default: fn = pssfp_tangent_launcher<kernels::PTAN_B0>(zpl); break;
        }
        {
            KernelScope ks(ctx, "pssfp_iso_tangent");
            fn(a, out, dev::div_ceil(nv, 64), unsigned(64 * groups), ctx.stream());
            check_launch("pssfp_iso_tangent");
        }
        if (z0 > 0) add_into(ctx, d.data(), part.data(), int64_t(nTR) * nv);
    }
    return d;
}

}  // namespace compas
