// Forward-mode (tangent) derivatives of the FISP EPG simulation.
//
// The reference differentiates by forward finite differences in FP32
// (src/simulate/derivative.cu:29-59, delta = 1e-4), which loses about three
// digits to cancellation and costs one more simulation per parameter.  This
// kernel runs the recursion of fisp.hip together with its tangents with
// respect to any set of tissue parameters p in {T1, T2, B1, B0}, in one pass:
// every state (F+, F-, Z) carries dF+/dp, dF-/dp, dZ/dp per parameter,
// updated by the same per-TR matrix plus the derivative of that matrix
// applied to the state.  The result is the exact derivative of the discrete
// recursion (no step size).  The per-TR matrices, the primal state and its
// next value are shared by all tangents of a pass; the echoes can be written
// by the same pass.
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
// Layout: LPV lanes per voxel (a power of two up to 64) with SPL states each:
// 4 with up to two tangents, 2 with more, so the primal and tangent states
// (6 floats per state and vector) stay in registers; max_state <= 128.

#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "compas/device/common.h"
#include "compas/device/fisp_common.h"
#include "compas/device/tangent_set.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

// i z
__device__ inline cfloat times_i(cfloat z) {
    return {-z.im, z.re};
}

struct FispPlanes {
    cfloat* p[5];  // [nTR * usf][nvoxels] each, in the order of TangentSet
};

// The three rows of a 3x3 real matrix applied to an EPG state (F+, F-, Z)
// in the (S = F+ + F-, D = F+ - F-) form the recursion uses: rows 0/1 act on
// S with half weights h, on D through i, row 2 likewise (fisp.hip).
struct Rows {
    float h00, h01, m02, h10, h11, m12, h20, h21, m22;
};

__device__ inline void apply_rows(const Rows& r, cfloat S, cfloat D, cfloat z, cfloat& Pn, cfloat& Qn, cfloat& Zn) {
    const float xr = fmaf(r.h00, S.re, fmaf(r.h01, D.im, r.m02 * z.re));
    const float xi = fmaf(r.h00, S.im, fmaf(-r.h01, D.re, r.m02 * z.im));
    const float yr = fmaf(r.h10, S.re, fmaf(r.h11, D.im, r.m12 * z.re));
    const float yi = fmaf(r.h10, S.im, fmaf(-r.h11, D.re, r.m12 * z.im));
    const float zr = fmaf(r.h20, S.re, fmaf(r.h21, D.im, r.m22 * z.re));
    const float zi = fmaf(r.h20, S.im, fmaf(-r.h21, D.re, r.m22 * z.im));
    Pn = {xr - yi, xi + yr};
    Qn = {xr + yi, xi - yr};
    Zn = {zr, zi};
}

// rows 0/1 only (the TE sample: the echo is F+ of state 0)
__device__ inline cfloat sample_rows(const Rows& r, cfloat S, cfloat D, cfloat z) {
    const float xr = fmaf(r.h00, S.re, fmaf(r.h01, D.im, r.m02 * z.re));
    const float xi = fmaf(r.h00, S.im, fmaf(-r.h01, D.re, r.m02 * z.im));
    const float yr = fmaf(r.h10, S.re, fmaf(r.h11, D.im, r.m12 * z.re));
    const float yi = fmaf(r.h10, S.im, fmaf(-r.h11, D.re, r.m12 * z.im));
    return {xr - yi, xi + yr};
}

// rows 0/1 of D(d) R and row 2 of E1 R, with the half weights on columns 0/1
__device__ inline Rows decay_rows(
    cfloat d, float E1, float R00, float R01, float R02, float R10, float R11, float R12, float R20, float R21,
    float R22) {
    Rows r;
    r.h00 = 0.5f * (d.re * R00 - d.im * R10);
    r.h01 = 0.5f * (d.re * R01 - d.im * R11);
    r.m02 = d.re * R02 - d.im * R12;
    r.h10 = 0.5f * (d.im * R00 + d.re * R10);
    r.h11 = 0.5f * (d.im * R01 + d.re * R11);
    r.m12 = d.im * R02 + d.re * R12;
    r.h20 = 0.5f * E1 * R20;
    r.h21 = 0.5f * E1 * R21;
    r.m22 = E1 * R22;
    return r;
}

// The recursion and its tangents for one voxel (lane `sub` of its LPV
// lanes).  FORM: the sincos of the half flip angles (FispHalfAngle);
// SIMPLE: the simple-sequence form of fisp_wave_form (no test in the TR
// loop, each TR's outputs stored once through running row pointers);
// direct_echo: fisp_direct_echo of the wave.
template<int SPL, int LPV, int MASK, int FORM, bool SIMPLE>
__device__ __forceinline__ void fisp_tangents_voxel(const FispArgs& a, const FispPlanes& out, int voxel, int sub,
                                                    bool direct_echo) {
    using TS = TangentSet<MASK>;
    constexpr int NT = TS::nt, NP = TS::planes, OFS = int(TS::echo);
    static_assert(NT >= 1, "a tangent pass carries at least one tangent");

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
    const float lTR = a.TR * iT1s;  // T1: factor of M row 2
    // T2: d/dT2 of the decay d over t is (t/T2^2) d; B0: i 2 pi t d
    const float k2TR = a.TR * iT2s, k2TE = a.TE * iT2s, k2TW = a.TW * iT2s;
    const float k0TR = dev::TWO_PI_F * a.TR, k0TE = dev::TWO_PI_F * a.TE, k0TW = dev::TWO_PI_F * a.TW;

    const bool owns_zero = (sub == 0);
    const int k0 = sub * SPL;

    // F+_0 at TE from the updated state, as in the primal kernel: F+_0(TE) =
    // F+_0(TR) / d_TRmTE, and for the tangents d/dp [F+_0(TR) / d_TRmTE] with
    // d_TRmTE depending on T2 (factor (TR - TE)/T2^2) and B0 (i 2 pi (TR - TE)).
    // A wave with a lane whose d_TRmTE cannot be inverted samples with the
    // operator D(TE) R as before.
    const float dd = d_TRmTE.re * d_TRmTE.re + d_TRmTE.im * d_TRmTE.im;
    const cfloat inv_TRmTE = direct_echo ? cfloat {0.0f, 0.0f}
                                         : cfloat {d_TRmTE.re / dd, -d_TRmTE.im / dd};
    const float k2TRmTE = (a.TR - a.TE) * iT2s, k0TRmTE = dev::TWO_PI_F * (a.TR - a.TE);

    for (int slice = 0; slice < a.nslices; slice++) {
        const float4* tab = a.table + size_t(slice) * size_t(a.nTR);
        cfloat P[SPL], Q[SPL], Z[SPL], dP[NT][SPL], dQ[NT][SPL], dZ[NT][SPL];
#pragma unroll
        for (int i = 0; i < SPL; i++) {
            P[i] = Q[i] = Z[i] = {0, 0};
#pragma unroll
            for (int k = 0; k < NT; k++) dP[k][i] = dQ[k][i] = dZ[k][i] = {0, 0};
        }
        if (owns_zero) Z[0] = {1.0f, 0.0f};

        for (int rep = 0; rep < a.repetitions; rep++) {
            const bool sample = (rep == a.repetitions - 1);
            if (a.inversion) {  // invert (ignores B1), spoil, decay(TI), regrow
#pragma unroll
                for (int i = 0; i < SPL; i++) {
                    P[i] = Q[i] = {0, 0};
#pragma unroll
                    for (int k = 0; k < NT; k++) {
                        dP[k][i] = dQ[k][i] = {0, 0};
                        dZ[k][i] = dZ[k][i] * (-E1_TI);
                        if (TS::t1 && k == TS::it1) dZ[k][i] = dZ[k][i] - Z[i] * dE1_TI;
                    }
                    Z[i] = Z[i] * (-E1_TI);
                }
                if (owns_zero) {
                    Z[0].re += 1.0f - E1_TI;
                    if constexpr (TS::t1) dZ[TS::it1][0].re -= dE1_TI;
                }
            }

            // the RF table entry of the next TR is loaded one TR ahead (the
            // table holds one entry more than the TRs of all slices)
            float4 t_next = (a.nTR > 0) ? tab[0] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            // SIMPLE: each output plane's row of the TR, and the voxel's byte offset in it
            cfloat* rows[NP];
#pragma unroll
            for (int k = 0; k < NP; k++) rows[k] = out.p[k];
            const uint32_t voff = uint32_t(voxel) * uint32_t(sizeof(cfloat));
            for (int n = 0; n < a.nTR; n++) {
                const float4 t = t_next;  // wave-uniform
                t_next = tab[n + 1];
                const float alpha = t.x * B1;
                float sx, cx;
                half_angle_sincos<FORM>(0.5f * alpha, &sx, &cx);
                const float sa = 2.0f * sx * cx;  // sin(alpha)
                const float omc = 2.0f * sx * sx;  // 1 - cos(alpha)
                const float cp = t.y, sp = t.z;
                const float R00 = 1.0f - omc * sp * sp, R01 = omc * cp * sp, R02 = sa * sp;
                const float R10 = R01, R11 = 1.0f - omc * cp * cp, R12 = -sa * cp;
                const float R20 = -sa * sp, R21 = sa * cp, R22 = 1.0f - omc;
                const Rows M = decay_rows(d_TR, E1_TR, R00, R01, R02, R10, R11, R12, R20, R21, R22);
                // B1: alpha = B1 |RF|, dR/dB1 = |RF| dR/dalpha (d sin = cos, d(1 - cos) = sin)
                Rows N {}, Sb {};
                if constexpr (TS::b1) {
                    const float ca = 1.0f - omc;
                    const float dR00 = -sa * sp * sp * t.x, dR01 = sa * cp * sp * t.x, dR02 = ca * sp * t.x;
                    const float dR10 = dR01, dR11 = -sa * cp * cp * t.x, dR12 = -ca * cp * t.x;
                    const float dR20 = -ca * sp * t.x, dR21 = ca * cp * t.x, dR22 = -sa * t.x;
                    N = decay_rows(d_TR, E1_TR, dR00, dR01, dR02, dR10, dR11, dR12, dR20, dR21, dR22);
                    if (!SIMPLE && direct_echo) Sb = decay_rows(d_TE, 0.0f, dR00, dR01, dR02, dR10, dR11, dR12, dR20, dR21, dR22);
                }

                if (!SIMPLE && direct_echo && sample && owns_zero) {
                    const Rows T = decay_rows(d_TE, 0.0f, R00, R01, R02, R10, R11, R12, R20, R21, R22);
                    const cfloat S0 = P[0] + Q[0], D0 = P[0] - Q[0];
                    const cfloat echo = sample_rows(T, S0, D0, Z[0]);
                    cfloat v[NP];
                    if constexpr (TS::echo) v[0] = echo;
#pragma unroll
                    for (int k = 0; k < NT; k++) {
                        cfloat dv = sample_rows(T, dP[k][0] + dQ[k][0], dP[k][0] - dQ[k][0], dZ[k][0]);
                        if (TS::t2 && k == TS::it2) dv = dv + echo * k2TE;
                        if (TS::b0 && k == TS::ib0) dv = dv + times_i(echo) * k0TE;
                        if (TS::b1 && k == TS::ib1) dv = dv + sample_rows(Sb, S0, D0, Z[0]);
                        v[OFS + k] = dv;
                    }
                    for (int u = 0; u < a.usf; u++) {
                        const size_t at = size_t(n * a.usf + u) * size_t(a.tissue.nvoxels) + voxel;
#pragma unroll
                        for (int k = 0; k < NP; k++) out.p[k][at] = (slice == 0) ? v[k] : (out.p[k][at] + v[k]);
                    }
                }

                cfloat Pn[SPL], Qn[SPL], dPn[NT][SPL], dQn[NT][SPL];
#pragma unroll
                for (int i = 0; i < SPL; i++) {
                    const cfloat S = P[i] + Q[i], D = P[i] - Q[i];
                    cfloat Zn;
                    apply_rows(M, S, D, Z[i], Pn[i], Qn[i], Zn);
#pragma unroll
                    for (int k = 0; k < NT; k++) {
                        cfloat tp, tq, tz;
                        apply_rows(M, dP[k][i] + dQ[k][i], dP[k][i] - dQ[k][i], dZ[k][i], tp, tq, tz);
                        if (TS::b1 && k == TS::ib1) {
                            cfloat np, nq, nz;
                            apply_rows(N, S, D, Z[i], np, nq, nz);
                            tp = tp + np;
                            tq = tq + nq;
                            tz = tz + nz;
                        }
                        if (TS::t1 && k == TS::it1) tz = tz + Zn * lTR;
                        if (TS::t2 && k == TS::it2) {
                            tp = tp + Pn[i] * k2TR;
                            tq = tq + Qn[i] * k2TR;
                        }
                        if (TS::b0 && k == TS::ib0) {
                            tp = tp + times_i(Pn[i]) * k0TR;
                            tq = tq - times_i(Qn[i]) * k0TR;
                        }
                        dPn[k][i] = tp;
                        dQn[k][i] = tq;
                        dZ[k][i] = tz;
                    }
                    Z[i] = Zn;
                }
                if (owns_zero) {
                    Z[0].re += G;
                    if constexpr (TS::t1) dZ[TS::it1][0].re += dG;
                }
                if ((SIMPLE || (!direct_echo && sample)) && owns_zero) {
                    const cfloat echo = dev::cmul(Pn[0], inv_TRmTE);
                    cfloat v[NP];
                    if constexpr (TS::echo) v[0] = echo;
#pragma unroll
                    for (int k = 0; k < NT; k++) {
                        cfloat dv = dev::cmul(dPn[k][0], inv_TRmTE);
                        if (TS::t2 && k == TS::it2) dv = dv - echo * k2TRmTE;
                        if (TS::b0 && k == TS::ib0) dv = dv - times_i(echo) * k0TRmTE;
                        v[OFS + k] = dv;
                    }
                    if constexpr (SIMPLE) {
#pragma unroll
                        for (int k = 0; k < NP; k++)
                            *reinterpret_cast<cfloat*>(reinterpret_cast<char*>(rows[k]) + voff) = v[k];
                    } else {
                        for (int u = 0; u < a.usf; u++) {
                            const size_t at = size_t(n * a.usf + u) * size_t(a.tissue.nvoxels) + voxel;
#pragma unroll
                            for (int k = 0; k < NP; k++) out.p[k][at] = (slice == 0) ? v[k] : (out.p[k][at] + v[k]);
                        }
                    }
                }
                if constexpr (SIMPLE) {
#pragma unroll
                    for (int k = 0; k < NP; k++) rows[k] += a.tissue.nvoxels;
                }

                // dephasing: F+ moves up one state, F- down one, primal and tangents alike
                auto shift = [&](cfloat (&p)[SPL], cfloat (&q)[SPL], const cfloat (&pn)[SPL], const cfloat (&qn)[SPL]) {
                    cfloat p_in = {0, 0}, q_in = {0, 0};
                    if constexpr (LPV > 1) {
                        p_in.re = __shfl_up(pn[SPL - 1].re, 1, LPV);
                        p_in.im = __shfl_up(pn[SPL - 1].im, 1, LPV);
                        q_in.re = __shfl_down(qn[0].re, 1, LPV);
                        q_in.im = __shfl_down(qn[0].im, 1, LPV);
                        if (sub == LPV - 1) q_in = {0, 0};
                    }
#pragma unroll
                    for (int i = SPL - 1; i >= 1; i--) p[i] = pn[i - 1];
                    p[0] = p_in;
#pragma unroll
                    for (int i = 0; i < SPL - 1; i++) q[i] = qn[i + 1];
                    q[SPL - 1] = q_in;
#pragma unroll
                    for (int i = 0; i < SPL; i++)
                        if (k0 + i >= a.N - 1) q[i] = {0, 0};
                    if (owns_zero) p[0] = conj(q[0]);
                };
                shift(P, Q, Pn, Qn);
#pragma unroll
                for (int k = 0; k < NT; k++) shift(dP[k], dQ[k], dPn[k], dQn[k]);
            }

            if (a.wait_spoiling) {
#pragma unroll
                for (int i = 0; i < SPL; i++) {
                    P[i] = Q[i] = {0, 0};
#pragma unroll
                    for (int k = 0; k < NT; k++) dP[k][i] = dQ[k][i] = {0, 0};
                }
            }
#pragma unroll
            for (int i = 0; i < SPL; i++) {
                const cfloat pW = dev::cmul(P[i], d_W), qW = dev::cmul(Q[i], conj(d_W));
#pragma unroll
                for (int k = 0; k < NT; k++) {
                    cfloat tP = dev::cmul(dP[k][i], d_W), tQ = dev::cmul(dQ[k][i], conj(d_W));
                    cfloat tZ = dZ[k][i] * E1_W;
                    if (TS::t2 && k == TS::it2) {
                        tP = tP + pW * k2TW;
                        tQ = tQ + qW * k2TW;
                    }
                    if (TS::b0 && k == TS::ib0) {
                        tP = tP + times_i(pW) * k0TW;
                        tQ = tQ - times_i(qW) * k0TW;
                    }
                    if (TS::t1 && k == TS::it1) tZ = tZ + Z[i] * dE1_W;
                    dP[k][i] = tP;
                    dQ[k][i] = tQ;
                    dZ[k][i] = tZ;
                }
                P[i] = pW;
                Q[i] = qW;
                Z[i] = Z[i] * E1_W;
            }
            if (owns_zero) {
                Z[0].re += 1.0f - E1_W;
                if constexpr (TS::t1) dZ[TS::it1][0].re -= dE1_W;
            }
        }
    }
}

// grid ceil(nvoxels LPV / 64), block 64: the recursion and tangents in the
// form the wave needs (fisp_wave_form).
template<int SPL, int LPV, int MASK>
__global__ void __launch_bounds__(64) fisp_epg_tangents(FispArgs a, FispPlanes out) {
    const int sub = (LPV > 1) ? int(threadIdx.x % LPV) : 0;
    const int voxel = (blockIdx.x * 64 + threadIdx.x) / LPV;
    const float max_angle = fisp_table_max_angle(a);  // while every lane is active
    if (voxel >= a.tissue.nvoxels) return;  // whole LPV groups leave together
    const FispWaveForm f = fisp_wave_form(a, voxel, max_angle);
    if (f.wide) {
        fisp_tangents_voxel<SPL, LPV, MASK, HALF_WIDE, false>(a, out, voxel, sub, f.direct_echo);
    } else if (!f.simple) {
        fisp_tangents_voxel<SPL, LPV, MASK, HALF_POLY, false>(a, out, voxel, sub, f.direct_echo);
    } else {
        fisp_tangents_voxel<SPL, LPV, MASK, HALF_POLY, true>(a, out, voxel, sub, false);
    }
}

}  // namespace kernels

namespace {

using kernels::TM_B0;
using kernels::TM_B1;
using kernels::TM_ECHO;
using kernels::TM_T1;
using kernels::TM_T2;

using TangentLauncher = void (*)(kernels::FispArgs, kernels::FispPlanes, unsigned, hipStream_t);

template<int SPL, int LPV, int MASK>
void launch_tangents(kernels::FispArgs a, kernels::FispPlanes o, unsigned blocks, hipStream_t s) {
    kernels::fisp_epg_tangents<SPL, LPV, MASK><<<blocks, 64, 0, s>>>(a, o);
}

// States per lane: four keep primal + two tangents (36 floats of state per
// lane) and the next-state temporaries within 96-134 VGPRs; with four
// tangents four states take 217 VGPRs (two waves per SIMD) and two states 136
// (three waves, twice the lanes per voxel and the per-TR matrices computed
// by twice the lanes).  COMPAS_FISP_TANGENT_SPL=2|4 picks for four tangents.
template<int MASK, int SPL>
TangentLauncher launcher_spl(int lpv) {
    switch (lpv) {
        case 1: return &launch_tangents<SPL, 1, MASK>;
        case 2: return &launch_tangents<SPL, 2, MASK>;
        case 4: return &launch_tangents<SPL, 4, MASK>;
        case 8: return &launch_tangents<SPL, 8, MASK>;
        case 16: return &launch_tangents<SPL, 16, MASK>;
        case 32: return &launch_tangents<SPL, 32, MASK>;
        default: return &launch_tangents<SPL, 64, MASK>;
    }
}

int wide_states_per_lane() {
    static const int spl = [] {
        const char* e = std::getenv("COMPAS_FISP_TANGENT_SPL");
        return (e && std::atoi(e) == 2) ? 2 : 4;
    }();
    return spl;
}

template<int MASK>
TangentLauncher launcher_for(int lpv, int* spl) {
    if constexpr (kernels::TangentSet<MASK>::nt <= 2) {
        *spl = 4;
        return launcher_spl<MASK, 4>(lpv);
    } else {
        *spl = wide_states_per_lane();
        return *spl == 2 ? launcher_spl<MASK, 2>(lpv) : launcher_spl<MASK, 4>(lpv);
    }
}

// field sets with a kernel of their own (as the pSSFP passes)
bool launcher_for_mask(int mask, int lpv, TangentLauncher* fn, int* spl) {
    switch (mask) {
#define COMPAS_FISP_MASK(M)                   \
    case M: *fn = launcher_for<M>(lpv, spl); return true;
        COMPAS_FISP_MASK(TM_T1)
        COMPAS_FISP_MASK(TM_T2)
        COMPAS_FISP_MASK(TM_B1)
        COMPAS_FISP_MASK(TM_B0)
        COMPAS_FISP_MASK(TM_T1 | TM_T2)
        COMPAS_FISP_MASK(TM_T1 | TM_T2 | TM_ECHO)
        COMPAS_FISP_MASK(TM_T1 | TM_T2 | TM_B1 | TM_B0)
        COMPAS_FISP_MASK(TM_T1 | TM_T2 | TM_B1 | TM_B0 | TM_ECHO)
#undef COMPAS_FISP_MASK
        default: return false;
    }
}

// the smallest power of two of lanes per voxel that leaves <= spl states per lane
int lanes_per_voxel(int max_state, int spl) {
    int lpv = 1;
    while (lpv < 64 && (max_state + lpv - 1) / lpv > spl) lpv *= 2;
    return lpv;
}

}  // namespace

std::vector<Array<cfloat, 2>> simulate_magnetization_tangents(
    const CompasContext& ctx,
    const std::vector<int>& fields,
    TissueParameters parameters,
    FISPSequence seq,
    Array<cfloat, 2>* echos) {
    const int nTR = int(seq.RF_train.size(0));
    const int nv = parameters.nvoxels;
    COMPAS_CHECK(seq.max_state >= 1);
    if (seq.max_state > 128) COMPAS_ERROR("max_state cannot exceed 128");
    COMPAS_CHECK(seq.sliceprofiles.size(1) == nTR);
    COMPAS_CHECK(seq.undersampling_factor >= 1);
    COMPAS_CHECK(seq.repetitions >= 1);
    int mask = 0;
    std::vector<int> bits;
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
        bits.push_back(bit);
    }
    // (with no voxels every field reads as absent: the empty result is returned below)
    if (nv > 0 && (((mask & TM_B1) && !parameters.has_b1) || ((mask & TM_B0) && !parameters.has_b0)))
        COMPAS_ERROR("cannot differentiate with respect to an absent parameter");
    const int nslices = int(seq.sliceprofiles.size(0));
    const int nr = nTR * seq.undersampling_factor;
    ctx.activate();
    std::vector<Array<cfloat, 2>> d;
    for (size_t i = 0; i < fields.size(); i++) d.push_back(ctx.empty<cfloat, 2>({nr, nv}));
    if (echos) *echos = ctx.empty<cfloat, 2>({nr, nv});
    if (nv == 0 || nr == 0 || fields.empty()) {
        if (echos && nv > 0 && nr > 0) *echos = simulate_magnetization(ctx, parameters, seq);
        return d;
    }
    if (nslices == 0) {
        for (auto& x : d) ctx.fill_zero(x.data(), x.nbytes());
        if (echos) ctx.fill_zero(echos->data(), echos->nbytes());
        return d;
    }
    cfloat* by_bit[32] = {};
    for (size_t i = 0; i < bits.size(); i++) by_bit[bits[i]] = d[i].data();
    if (echos) by_bit[TM_ECHO] = echos->data();

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

    // passes with a kernel of their own: all four, or T1 + T2, then single fields
    const int fields_mask = mask;
    std::vector<int> passes;
    const int with_echo = echos ? TM_ECHO : 0;
    if (fields_mask == (TM_T1 | TM_T2 | TM_B1 | TM_B0)) {
        passes.push_back(fields_mask | with_echo);
    } else {
        int rest = fields_mask;
        if ((rest & (TM_T1 | TM_T2)) == (TM_T1 | TM_T2)) {
            passes.push_back(TM_T1 | TM_T2 | with_echo);
            rest &= ~(TM_T1 | TM_T2);
        } else if (with_echo) {
            passes.push_back(TM_ECHO);  // the plain simulation
        }
        for (int bit : {TM_T1, TM_T2, TM_B1, TM_B0})
            if (rest & bit) passes.push_back(bit);
    }
    for (int pass : passes) {
        if (pass == TM_ECHO) {
            *echos = simulate_magnetization(ctx, parameters, seq);
            continue;
        }
        kernels::FispPlanes o {};
        int np = 0;
        for (int bit : {TM_ECHO, TM_T1, TM_T2, TM_B1, TM_B0})
            if (pass & bit) o.p[np++] = by_bit[bit];
        TangentLauncher fn = nullptr;
        int spl = 4;
        launcher_for_mask(pass, 1, &fn, &spl);
        const int lpv = lanes_per_voxel(seq.max_state, spl);
        launcher_for_mask(pass, lpv, &fn, &spl);
        KernelScope ks(ctx, "fisp_epg_tangents");
        fn(a, o, unsigned(dev::div_ceil(int64_t(nv) * lpv, 64)), ctx.stream());
        check_launch("fisp_epg_tangents");
    }
    return d;
}

Array<cfloat, 2> simulate_magnetization_tangent(
    const CompasContext& ctx, int field, TissueParameters parameters, FISPSequence seq) {
    return simulate_magnetization_tangents(ctx, std::vector<int> {field}, parameters, seq, nullptr)[0];
}

}  // namespace compas
