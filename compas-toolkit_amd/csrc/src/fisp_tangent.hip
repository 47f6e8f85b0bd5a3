// Forward-mode (tangent) derivatives of the FISP EPG simulation.
//
// The reference differentiates by forward finite differences in FP32
// (src/simulate/derivative.cu:29-59, delta = 1e-4), which loses about three
// digits to cancellation.  This kernel runs the recursion of fisp.hip together
// with its tangent with respect to one tissue parameter p in {T1, T2, B1, B0}:
// every state (F+, F-, Z) carries dF+/dp, dF-/dp, dZ/dp, updated by the same
// per-TR matrix plus the derivative of that matrix applied to the state.  The
// result is the exact derivative of the discrete recursion (no step size).
//
// How each parameter enters one TR (M = D(TR) R(alpha), rows 0/1 of M mix the
// transverse components with the complex decay d = E2 exp(i 2 pi t B0), row 2
// carries E1; regrowth G on Z_0):
//   T2: dd/dT2 = (t/T2^2) d, so dM rows 0/1 = (TR/T2^2) M rows 0/1: the new
//       F+/F- gain (TR/T2^2) times their primal values;
//   B0: dd/dB0 = i 2 pi t d: F+ gains i 2 pi TR F+, F- gains -i 2 pi TR F-;
//   T1: dE1/dT1 = (t/T1^2) E1: Z gains (TR/T1^2) (M row 2 applied to the
//       state), Z_0 gains dG/dT1;
//   B1: alpha = B1 |RF| (rad): dM = D(TR) dR/dalpha |RF|, one more matrix
//       application per state.
// The TE sample, the inversion (T1 only: the inversion ignores B1) and the
// end-of-repetition decay over TW differentiate the same way.
//
// Layout: LPV lanes per voxel (a power of two up to 32) with SPL = 4 states
// each, so the primal and tangent states (12 floats per state) stay in
// registers; max_state <= 128.

#include <cfloat>
#include <cmath>
#include <vector>

#include "compas/device/common.h"
#include "compas/device/fisp_common.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

enum { TAN_T1 = 0, TAN_T2 = 1, TAN_B1 = 2, TAN_B0 = 3 };

// i z
__device__ inline cfloat times_i(cfloat z) {
    return {-z.im, z.re};
}

template<int SPL, int LPV, int FIELD>
__global__ void __launch_bounds__(64) fisp_epg_tangent(FispArgs a, cfloat* __restrict__ dechos) {
    const int sub = (LPV > 1) ? int(threadIdx.x % LPV) : 0;
    const int voxel = (blockIdx.x * 64 + threadIdx.x) / LPV;
    if (voxel >= a.tissue.nvoxels) return;  // whole LPV groups leave together

    const float T1 = a.tissue.T1(voxel);
    const float T2 = a.tissue.T2(voxel);
    const float B1 = a.tissue.B1(voxel);
    const float B0 = a.tissue.B0(voxel);

    const float E1_TE = relax(a.TE, T1), E2_TE = relax(a.TE, T2);
    const float E1_TI = relax(a.TI, T1);
    const float E1_TRmTE = relax(a.TR - a.TE, T1), E2_TRmTE = relax(a.TR - a.TE, T2);
    const float E1_W = relax(a.TW, T1), E2_W = relax(a.TW, T2);
    const cfloat d_TE = off_resonance(a.TE, B0) * E2_TE;
    const cfloat d_TRmTE = off_resonance(a.TR - a.TE, B0) * E2_TRmTE;
    const cfloat d_W = off_resonance(a.TW, B0) * E2_W;
    const cfloat d_TR = dev::cmul(d_TE, d_TRmTE);
    const float E1_TR = E1_TE * E1_TRmTE;
    const float G = E1_TRmTE * (1.0f - E1_TE) + (1.0f - E1_TRmTE);

    // derivative factors (d/dT of exp(-t/T) = (t/T^2) exp(-t/T))
    const float iT1s = 1.0f / (T1 * T1), iT2s = 1.0f / (T2 * T2);
    const float dE1_TE = a.TE * iT1s * E1_TE, dE1_TRmTE = (a.TR - a.TE) * iT1s * E1_TRmTE;
    const float dE1_TI = a.TI * iT1s * E1_TI, dE1_W = a.TW * iT1s * E1_W;
    const float dG = dE1_TRmTE * (1.0f - E1_TE) - E1_TRmTE * dE1_TE - dE1_TRmTE;
    const float kTR = (FIELD == TAN_T2) ? a.TR * iT2s : dev::TWO_PI_F * a.TR;  // T2 / B0 factor over TR
    const float kTE = (FIELD == TAN_T2) ? a.TE * iT2s : dev::TWO_PI_F * a.TE;
    const float kTW = (FIELD == TAN_T2) ? a.TW * iT2s : dev::TWO_PI_F * a.TW;
    const float lTR = a.TR * iT1s;  // T1 factor of M row 2

    const bool owns_zero = (sub == 0);
    const int k0 = sub * SPL;

    for (int slice = 0; slice < a.nslices; slice++) {
        const float4* tab = a.table + size_t(slice) * size_t(a.nTR);
        cfloat P[SPL], Q[SPL], Z[SPL], dP[SPL], dQ[SPL], dZ[SPL];
#pragma unroll
        for (int i = 0; i < SPL; i++) P[i] = Q[i] = Z[i] = dP[i] = dQ[i] = dZ[i] = {0, 0};
        if (owns_zero) Z[0] = {1.0f, 0.0f};

        for (int rep = 0; rep < a.repetitions; rep++) {
            const bool sample = (rep == a.repetitions - 1);
            if (a.inversion) {  // invert (ignores B1), spoil, decay(TI), regrow
#pragma unroll
                for (int i = 0; i < SPL; i++) {
                    P[i] = Q[i] = dP[i] = dQ[i] = {0, 0};
                    if constexpr (FIELD == TAN_T1) dZ[i] = dZ[i] * (-E1_TI) - Z[i] * dE1_TI;
                    else dZ[i] = dZ[i] * (-E1_TI);
                    Z[i] = Z[i] * (-E1_TI);
                }
                if (owns_zero) {
                    Z[0].re += 1.0f - E1_TI;
                    if constexpr (FIELD == TAN_T1) dZ[0].re -= dE1_TI;
                }
            }

            for (int n = 0; n < a.nTR; n++) {
                const float4 t = tab[n];  // wave-uniform
                const float alpha = t.x * B1;
                float sx, cx;
                dev::sincos_fast(0.5f * alpha, &sx, &cx);
                const float sa = 2.0f * sx * cx;  // sin(alpha)
                const float omc = 2.0f * sx * sx;  // 1 - cos(alpha)
                const float cp = t.y, sp = t.z;
                const float R00 = 1.0f - omc * sp * sp, R01 = omc * cp * sp, R02 = sa * sp;
                const float R10 = R01, R11 = 1.0f - omc * cp * cp, R12 = -sa * cp;
                const float R20 = -sa * sp, R21 = sa * cp, R22 = 1.0f - omc;
                const float dr = d_TR.re, di = d_TR.im;
                const float M00 = dr * R00 - di * R10, M01 = dr * R01 - di * R11, M02 = dr * R02 - di * R12;
                const float M10 = di * R00 + dr * R10, M11 = di * R01 + dr * R11, M12 = di * R02 + dr * R12;
                const float M20 = E1_TR * R20, M21 = E1_TR * R21, M22 = E1_TR * R22;
                const float h00 = 0.5f * M00, h01 = 0.5f * M01;
                const float h10 = 0.5f * M10, h11 = 0.5f * M11;
                const float h20 = 0.5f * M20, h21 = 0.5f * M21;
                // dR/dB1 = |RF| dR/dalpha (d sin(alpha) = cos(alpha), d(1 - cos(alpha)) = sin(alpha))
                float N00 = 0, N01 = 0, N02 = 0, N10 = 0, N11 = 0, N12 = 0, N20 = 0, N21 = 0, N22 = 0;
                float S00 = 0, S01 = 0, S02 = 0, S10 = 0, S11 = 0, S12 = 0;
                if constexpr (FIELD == TAN_B1) {
                    const float ca = 1.0f - omc;
                    const float dR00 = -sa * sp * sp * t.x, dR01 = sa * cp * sp * t.x, dR02 = ca * sp * t.x;
                    const float dR10 = dR01, dR11 = -sa * cp * cp * t.x, dR12 = -ca * cp * t.x;
                    const float dR20 = -ca * sp * t.x, dR21 = ca * cp * t.x, dR22 = -sa * t.x;
                    N00 = 0.5f * (dr * dR00 - di * dR10);
                    N01 = 0.5f * (dr * dR01 - di * dR11);
                    N02 = dr * dR02 - di * dR12;
                    N10 = 0.5f * (di * dR00 + dr * dR10);
                    N11 = 0.5f * (di * dR01 + dr * dR11);
                    N12 = di * dR02 + dr * dR12;
                    N20 = 0.5f * E1_TR * dR20;
                    N21 = 0.5f * E1_TR * dR21;
                    N22 = E1_TR * dR22;
                    // rows 0/1 of D(TE) dR for the sample
                    const float er = d_TE.re, ei = d_TE.im;
                    S00 = 0.5f * (er * dR00 - ei * dR10);
                    S01 = 0.5f * (er * dR01 - ei * dR11);
                    S02 = er * dR02 - ei * dR12;
                    S10 = 0.5f * (ei * dR00 + er * dR10);
                    S11 = 0.5f * (ei * dR01 + er * dR11);
                    S12 = ei * dR02 + er * dR12;
                }

                if (sample && owns_zero) {
                    const float er = d_TE.re, ei = d_TE.im;
                    const float T00 = 0.5f * (er * R00 - ei * R10), T01 = 0.5f * (er * R01 - ei * R11);
                    const float T02 = er * R02 - ei * R12;
                    const float T10 = 0.5f * (ei * R00 + er * R10), T11 = 0.5f * (ei * R01 + er * R11);
                    const float T12 = ei * R02 + er * R12;
                    const cfloat S0 = P[0] + Q[0], D0 = P[0] - Q[0], z0 = Z[0];
                    const cfloat dS0 = dP[0] + dQ[0], dD0 = dP[0] - dQ[0], dz0 = dZ[0];
                    // primal echo (needed by the T2 / B0 terms) and the tangent of the state term
                    const float xr = fmaf(T00, S0.re, fmaf(T01, D0.im, T02 * z0.re));
                    const float xi = fmaf(T00, S0.im, fmaf(-T01, D0.re, T02 * z0.im));
                    const float yr = fmaf(T10, S0.re, fmaf(T11, D0.im, T12 * z0.re));
                    const float yi = fmaf(T10, S0.im, fmaf(-T11, D0.re, T12 * z0.im));
                    const cfloat echo = {xr - yi, xi + yr};
                    const float dxr = fmaf(T00, dS0.re, fmaf(T01, dD0.im, T02 * dz0.re));
                    const float dxi = fmaf(T00, dS0.im, fmaf(-T01, dD0.re, T02 * dz0.im));
                    const float dyr = fmaf(T10, dS0.re, fmaf(T11, dD0.im, T12 * dz0.re));
                    const float dyi = fmaf(T10, dS0.im, fmaf(-T11, dD0.re, T12 * dz0.im));
                    cfloat decho = {dxr - dyi, dxi + dyr};
                    if constexpr (FIELD == TAN_T2) decho = decho + echo * kTE;
                    if constexpr (FIELD == TAN_B0) decho = decho + times_i(echo) * kTE;
                    if constexpr (FIELD == TAN_B1) {
                        const float sxr = fmaf(S00, S0.re, fmaf(S01, D0.im, S02 * z0.re));
                        const float sxi = fmaf(S00, S0.im, fmaf(-S01, D0.re, S02 * z0.im));
                        const float syr = fmaf(S10, S0.re, fmaf(S11, D0.im, S12 * z0.re));
                        const float syi = fmaf(S10, S0.im, fmaf(-S11, D0.re, S12 * z0.im));
                        decho = decho + cfloat {sxr - syi, sxi + syr};
                    }
                    for (int u = 0; u < a.usf; u++) {
                        cfloat* out = dechos + size_t(n * a.usf + u) * size_t(a.tissue.nvoxels) + voxel;
                        *out = (slice == 0) ? decho : (*out + decho);
                    }
                }

                cfloat Pn[SPL], Qn[SPL], dPn[SPL], dQn[SPL];
#pragma unroll
                for (int i = 0; i < SPL; i++) {
                    const cfloat S = P[i] + Q[i], D = P[i] - Q[i], z = Z[i];
                    const float xr = fmaf(h00, S.re, fmaf(h01, D.im, M02 * z.re));
                    const float xi = fmaf(h00, S.im, fmaf(-h01, D.re, M02 * z.im));
                    const float yr = fmaf(h10, S.re, fmaf(h11, D.im, M12 * z.re));
                    const float yi = fmaf(h10, S.im, fmaf(-h11, D.re, M12 * z.im));
                    const float zr = fmaf(h20, S.re, fmaf(h21, D.im, M22 * z.re));
                    const float zi = fmaf(h20, S.im, fmaf(-h21, D.re, M22 * z.im));
                    Pn[i] = {xr - yi, xi + yr};
                    Qn[i] = {xr + yi, xi - yr};

                    const cfloat dS = dP[i] + dQ[i], dD = dP[i] - dQ[i], dz = dZ[i];
                    float txr = fmaf(h00, dS.re, fmaf(h01, dD.im, M02 * dz.re));
                    float txi = fmaf(h00, dS.im, fmaf(-h01, dD.re, M02 * dz.im));
                    float tyr = fmaf(h10, dS.re, fmaf(h11, dD.im, M12 * dz.re));
                    float tyi = fmaf(h10, dS.im, fmaf(-h11, dD.re, M12 * dz.im));
                    float tzr = fmaf(h20, dS.re, fmaf(h21, dD.im, M22 * dz.re));
                    float tzi = fmaf(h20, dS.im, fmaf(-h21, dD.re, M22 * dz.im));
                    if constexpr (FIELD == TAN_B1) {
                        txr += fmaf(N00, S.re, fmaf(N01, D.im, N02 * z.re));
                        txi += fmaf(N00, S.im, fmaf(-N01, D.re, N02 * z.im));
                        tyr += fmaf(N10, S.re, fmaf(N11, D.im, N12 * z.re));
                        tyi += fmaf(N10, S.im, fmaf(-N11, D.re, N12 * z.im));
                        tzr += fmaf(N20, S.re, fmaf(N21, D.im, N22 * z.re));
                        tzi += fmaf(N20, S.im, fmaf(-N21, D.re, N22 * z.im));
                    }
                    if constexpr (FIELD == TAN_T1) {
                        tzr = fmaf(lTR, zr, tzr);
                        tzi = fmaf(lTR, zi, tzi);
                    }
                    dPn[i] = {txr - tyi, txi + tyr};
                    dQn[i] = {txr + tyi, txi - tyr};
                    if constexpr (FIELD == TAN_T2) {
                        dPn[i] = dPn[i] + Pn[i] * kTR;
                        dQn[i] = dQn[i] + Qn[i] * kTR;
                    }
                    if constexpr (FIELD == TAN_B0) {
                        dPn[i] = dPn[i] + times_i(Pn[i]) * kTR;
                        dQn[i] = dQn[i] - times_i(Qn[i]) * kTR;
                    }
                    Z[i] = {zr, zi};
                    dZ[i] = {tzr, tzi};
                }
                if (owns_zero) {
                    Z[0].re += G;
                    if constexpr (FIELD == TAN_T1) dZ[0].re += dG;
                }

                // dephasing, primal and tangent alike
                cfloat p_in = {0, 0}, q_in = {0, 0}, dp_in = {0, 0}, dq_in = {0, 0};
                if constexpr (LPV > 1) {
                    p_in.re = __shfl_up(Pn[SPL - 1].re, 1, LPV);
                    p_in.im = __shfl_up(Pn[SPL - 1].im, 1, LPV);
                    q_in.re = __shfl_down(Qn[0].re, 1, LPV);
                    q_in.im = __shfl_down(Qn[0].im, 1, LPV);
                    dp_in.re = __shfl_up(dPn[SPL - 1].re, 1, LPV);
                    dp_in.im = __shfl_up(dPn[SPL - 1].im, 1, LPV);
                    dq_in.re = __shfl_down(dQn[0].re, 1, LPV);
                    dq_in.im = __shfl_down(dQn[0].im, 1, LPV);
                    if (sub == LPV - 1) q_in = dq_in = {0, 0};
                }
#pragma unroll
                for (int i = SPL - 1; i >= 1; i--) {
                    P[i] = Pn[i - 1];
                    dP[i] = dPn[i - 1];
                }
                P[0] = p_in;
                dP[0] = dp_in;
#pragma unroll
                for (int i = 0; i < SPL - 1; i++) {
                    Q[i] = Qn[i + 1];
                    dQ[i] = dQn[i + 1];
                }
                Q[SPL - 1] = q_in;
                dQ[SPL - 1] = dq_in;
#pragma unroll
                for (int i = 0; i < SPL; i++) {
                    if (k0 + i >= a.N - 1) Q[i] = dQ[i] = {0, 0};
                }
                if (owns_zero) {
                    P[0] = conj(Q[0]);
                    dP[0] = conj(dQ[0]);
                }
            }

            if (a.wait_spoiling) {
#pragma unroll
                for (int i = 0; i < SPL; i++) P[i] = Q[i] = dP[i] = dQ[i] = {0, 0};
            }
#pragma unroll
            for (int i = 0; i < SPL; i++) {
                cfloat tP = dev::cmul(dP[i], d_W), tQ = dev::cmul(dQ[i], conj(d_W)), tZ = dZ[i] * E1_W;
                const cfloat pW = dev::cmul(P[i], d_W), qW = dev::cmul(Q[i], conj(d_W));
                if constexpr (FIELD == TAN_T2) {
                    tP = tP + pW * kTW;
                    tQ = tQ + qW * kTW;
                }
                if constexpr (FIELD == TAN_B0) {
                    tP = tP + times_i(pW) * kTW;
                    tQ = tQ - times_i(qW) * kTW;
                }
                if constexpr (FIELD == TAN_T1) tZ = tZ + Z[i] * dE1_W;
                P[i] = pW;
                Q[i] = qW;
                Z[i] = Z[i] * E1_W;
                dP[i] = tP;
                dQ[i] = tQ;
                dZ[i] = tZ;
            }
            if (owns_zero) {
                Z[0].re += 1.0f - E1_W;
                if constexpr (FIELD == TAN_T1) dZ[0].re -= dE1_W;
            }
        }
    }
}

}  // namespace kernels

namespace {

using TangentLauncher = void (*)(kernels::FispArgs, cfloat*, unsigned, hipStream_t);

template<int SPL, int LPV, int FIELD>
void launch_tangent(kernels::FispArgs a, cfloat* d, unsigned blocks, hipStream_t s) {
    kernels::fisp_epg_tangent<SPL, LPV, FIELD><<<blocks, 64, 0, s>>>(a, d);
}

// Four states per lane keep primal + tangent state (48 floats) and the
// next-state temporaries under 128 VGPRs (four waves per SIMD); eight states
// per lane need 236+ and leave one.
template<int FIELD>
TangentLauncher tangent_launcher(int lpv) {
    switch (lpv) {
        case 1: return &launch_tangent<4, 1, FIELD>;
        case 2: return &launch_tangent<4, 2, FIELD>;
        case 4: return &launch_tangent<4, 4, FIELD>;
        case 8: return &launch_tangent<4, 8, FIELD>;
        case 16: return &launch_tangent<4, 16, FIELD>;
        default: return &launch_tangent<4, 32, FIELD>;
    }
}

}  // namespace

Array<cfloat, 2> simulate_magnetization_tangent(
    const CompasContext& ctx, int field, TissueParameters parameters, FISPSequence seq) {
    const int nTR = int(seq.RF_train.size(0));
    const int nv = parameters.nvoxels;
    COMPAS_CHECK(seq.max_state >= 1);
    if (seq.max_state > 128) COMPAS_ERROR("max_state cannot exceed 128");
    COMPAS_CHECK(seq.sliceprofiles.size(1) == nTR);
    COMPAS_CHECK(seq.undersampling_factor >= 1);
    COMPAS_CHECK(seq.repetitions >= 1);
    if (field != TissueParameterField::T1 && field != TissueParameterField::T2 &&
        field != TissueParameterField::B1 && field != TissueParameterField::B0)
        COMPAS_ERROR("tangent derivatives are available for T1, T2, B1 and B0");
    // (with no voxels every field reads as absent: the empty result is returned below)
    if (nv > 0 && ((field == TissueParameterField::B1 && !parameters.has_b1) ||
                   (field == TissueParameterField::B0 && !parameters.has_b0)))
        COMPAS_ERROR("cannot differentiate with respect to an absent parameter");
    const int nslices = int(seq.sliceprofiles.size(0));
    const int nr = nTR * seq.undersampling_factor;
    ctx.activate();
    auto d = ctx.empty<cfloat, 2>({nr, nv});
    if (nv == 0 || nr == 0) return d;
    if (nslices == 0) {
        ctx.fill_zero(d.data(), d.nbytes());
        return d;
    }
    auto table = fisp_rf_table(ctx, seq);

    kernels::FispArgs a;
    a.tissue = dev::tissue_view(parameters);
    a.table = table.data();
    a.nTR = nTR;
    a.nslices = nslices;
    a.TR = seq.TR;
    a.TE = seq.TE;
    a.TW = seq.TW;
    a.TI = seq.TI;
    a.N = seq.max_state;
    a.usf = seq.undersampling_factor;
    a.repetitions = seq.repetitions;
    a.inversion = seq.inversion_prepulse;
    a.wait_spoiling = seq.wait_spoiling;
    a.echos = nullptr;

    // lanes per voxel: the smallest power of two that leaves <= 4 states per lane
    int lpv = 1;
    while (lpv < 32 && (seq.max_state + lpv - 1) / lpv > 4) lpv *= 2;
    TangentLauncher fn;
    switch (field) {
        case TissueParameterField::T1: fn = tangent_launcher<kernels::TAN_T1>(lpv); break;
        case TissueParameterField::T2: fn = tangent_launcher<kernels::TAN_T2>(lpv); break;
        case TissueParameterField::B1: fn = tangent_launcher<kernels::TAN_B1>(lpv); break;
        default: fn = tangent_launcher<kernels::TAN_B0>(lpv); break;
    }
    KernelScope ks(ctx, "fisp_epg_tangent");
    fn(a, d.data(), dev::div_ceil(int64_t(nv) * lpv, 64), ctx.stream());
    check_launch("fisp_epg_tangent");
    return d;
}

}  // namespace compas

namespace compas {

std::vector<Array<cfloat, 2>> simulate_magnetization_tangents(
    const CompasContext& ctx,
    const std::vector<int>& fields,
    TissueParameters parameters,
    FISPSequence seq,
    Array<cfloat, 2>* echos) {
    for (size_t i = 0; i < fields.size(); i++)
        for (size_t j = 0; j < i; j++)
            if (fields[i] == fields[j]) COMPAS_ERROR("a field is listed twice");
    std::vector<Array<cfloat, 2>> d;
    for (int f : fields) d.push_back(simulate_magnetization_tangent(ctx, f, parameters, seq));
    if (echos) *echos = simulate_magnetization(ctx, parameters, seq);
    return d;
}

}  // namespace compas
