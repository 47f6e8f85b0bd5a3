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
// Layout: as the primal kernel (one wave per block of 64 voxels and group of
// ZPL <= 4 slice positions, partial sums over z added by a second pass).

#include <algorithm>

#include "compas/device/common.h"
#include "compas/device/pssfp_common.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

enum { PTAN_T1 = 0, PTAN_T2 = 1, PTAN_B1 = 2, PTAN_B0 = 3 };

// s1'(u), s2'(u) from the Taylor series (term-wise derivatives, one more
// term each; first omitted terms below 1e-18), u <= SERIES_LIMIT
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

// Rot(a) m and its tangent: Rot(a) dm, plus (dRot/da . da) m when AXIS.
// ANY: u may exceed SERIES_LIMIT (closed forms there); else the series only.
template<bool AXIS, bool ANY>
__device__ inline void rotate_tangent(
    Vec3d& m,
    Vec3d& dm,
    double ax,
    double ay,
    double az,
    double u,
    double dax,
    double day,
    double daz) {
    double s1, s2, s1p = 0.0, s2p = 0.0;
    if (!ANY || u <= SERIES_LIMIT) {
        rotation_series(u, s1, s2);
        if constexpr (AXIS) rotation_series_derivative(u, s1p, s2p);
    } else {
        rotation_any(u, s1, s2);
        if constexpr (AXIS) {
            const double c = fma(-u, s2, 1.0);
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

// grid (ceil(nv / 64), groups in this launch), block 64: one wave per voxel
// block and group of ZPL slice positions, as the primal kernel; blockIdx.y
// == 0 writes out0, blockIdx.y = j > 0 writes rest + (j - 1) * nTR * nv.
template<int ZPL, int FIELD>
__global__ void __launch_bounds__(64) pssfp_iso_tangent(PssfpArgs a, int group0, cfloat* out0, cfloat* rest) {
    __shared__ double2 grf[RF_CHUNK];
    const int lane = threadIdx.x;
    const int group = group0 + int(blockIdx.y);
    const int voxel = blockIdx.x * 64 + lane;
    const bool active = voxel < a.tissue.nvoxels;
    const int vx = active ? voxel : 0;
    const size_t plane = size_t(a.nTR) * size_t(a.tissue.nvoxels);
    cfloat* out = (blockIdx.y == 0) ? out0 : rest + size_t(blockIdx.y - 1) * plane;

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
        for (int k = lane; k < RF_CHUNK && k0 + k < a.nRF; k += 64) {
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
                double u[ZPL];
                bool large = false;
#pragma unroll
                for (int i = 0; i < ZPL; i++) {
                    u[i] = fma(az[i], az[i], axy2);
                    large = large || (u[i] > SERIES_LIMIT);
                }
                if (__any(large)) {  // wave-uniform: the all-small case is branch-free
#pragma unroll
                    for (int i = 0; i < ZPL; i++) {
                        rotate_tangent<AXIS, true>(m[i], dm[i], ax, ay, az[i], u[i], dax, day, dz);
                        relax_tangent<FIELD>(m[i], dm[i], E1_ex, E2_ex, R1_ex, dE_ex);
                    }
                } else {
#pragma unroll
                    for (int i = 0; i < ZPL; i++) {
                        rotate_tangent<AXIS, false>(m[i], dm[i], ax, ay, az[i], u[i], dax, day, dz);
                        relax_tangent<FIELD>(m[i], dm[i], E1_ex, E2_ex, R1_ex, dE_ex);
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
        if (active) out[size_t(n) * size_t(a.tissue.nvoxels) + voxel] = {float(sx), float(sy)};
#pragma unroll
        for (int i = 0; i < ZPL; i++) {
            rotate_z_tangent<FIELD>(m[i], dm[i], cpr[i], spr[i], dth);
            relax_tangent<FIELD>(m[i], dm[i], E1_pr, E2_pr, R1_pr, dE_pr);
        }
    }
}

}  // namespace kernels

namespace {

using PssfpTangentLauncher = void (*)(kernels::PssfpArgs, int, cfloat*, cfloat*, dim3, hipStream_t);

template<int ZPL, int FIELD>
void launch_pssfp_tangent(kernels::PssfpArgs a, int group0, cfloat* out0, cfloat* rest, dim3 grid, hipStream_t s) {
    kernels::pssfp_iso_tangent<ZPL, FIELD><<<grid, 64, 0, s>>>(a, group0, out0, rest);
}

template<int FIELD>
PssfpTangentLauncher pssfp_tangent_launcher(int zpl) {
    switch (zpl) {
        case 1: return &launch_pssfp_tangent<1, FIELD>;
        case 2: return &launch_pssfp_tangent<2, FIELD>;
        case 3: return &launch_pssfp_tangent<3, FIELD>;
        default: return &launch_pssfp_tangent<4, FIELD>;  // primal + tangent of 4 fit the registers
    }
}

template<int ZPL, int FIELD>
int resident_waves() {
    static const int cached = [] {
        int n = 0;
        COMPAS_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernels::pssfp_iso_tangent<ZPL, FIELD>, 64, 0));
        return std::max(1, n);
    }();
    return cached;
}

template<int FIELD>
void tangent_layout(const CompasContext& ctx, int nz, int nv, int nTR, PssfpLayout* L, PssfpTangentLauncher* fn) {
    const int waves[5] = {0,
                          resident_waves<1, FIELD>(),
                          resident_waves<2, FIELD>(),
                          resident_waves<3, FIELD>(),
                          resident_waves<4, FIELD>()};
    *L = pssfp_layout(ctx, nz, nv, nTR, waves, 4);
    *fn = pssfp_tangent_launcher<FIELD>(L->zpl);
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
    // (with no voxels every field reads as absent: the empty result is returned below)
    if (nv > 0 && ((field == TissueParameterField::B1 && !parameters.has_b1) ||
                   (field == TissueParameterField::B0 && !parameters.has_b0)))
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

    PssfpLayout L;
    PssfpTangentLauncher fn;
    switch (field) {
        case TissueParameterField::T1: tangent_layout<kernels::PTAN_T1>(ctx, nz, nv, nTR, &L, &fn); break;
        case TissueParameterField::T2: tangent_layout<kernels::PTAN_T2>(ctx, nz, nv, nTR, &L, &fn); break;
        case TissueParameterField::B1: tangent_layout<kernels::PTAN_B1>(ctx, nz, nv, nTR, &L, &fn); break;
        default: tangent_layout<kernels::PTAN_B0>(ctx, nz, nv, nTR, &L, &fn); break;
    }
    const int64_t plane = int64_t(nTR) * nv;
    const int slots = L.groups > L.per_launch ? L.per_launch : L.groups - 1;
    Array<cfloat, 2> part;
    if (slots > 0) part = ctx.empty<cfloat, 2>({slots, plane});
    const unsigned vblocks = unsigned(dev::div_ceil(nv, 64));
    for (int g0 = 0; g0 < L.groups; g0 += L.per_launch) {
        const int count = std::min(L.per_launch, L.groups - g0);
        cfloat* out0 = g0 == 0 ? d.data() : part.data();
        cfloat* rest = g0 == 0 ? part.data() : part.data() + plane;
        {
            KernelScope ks(ctx, "pssfp_iso_tangent");
            fn(a, g0, out0, rest, dim3(vblocks, unsigned(count)), ctx.stream());
            check_launch("pssfp_iso_tangent");
        }
        pssfp_sum_groups(ctx, d.data(), part.data(), g0 == 0 ? count - 1 : count, plane);
    }
    return d;
}

}  // namespace compas
