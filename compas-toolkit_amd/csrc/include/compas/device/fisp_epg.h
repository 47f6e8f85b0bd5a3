#pragma once
// The FISP EPG recursion of one voxel, shared by the per-call kernel
// (fisp.hip) and the request server's work-table kernel (fisp_table.hip).
//
// Semantics: reference src/simulate/fisp_kernels.cuh:99-244 and the EPG
// operators of src/operators/epg.cuh:19-322.  The per-TR operator "excite ->
// decay(TE) -> regrow -> decay(TR-TE) -> regrow" is one real 3x3 matrix
// M = D(TR) R(alpha, phi) acting on the state in the real transverse basis
// X = (F+ + F-)/2, Y = (F+ - F-)/(2i); the dephasing shift is a register
// rename (or a lane shuffle when LPV > 1 lanes share a voxel).

#include "compas/device/common.h"
#include "compas/device/fisp_common.h"

namespace compas {
namespace kernels {

// The recursion of one voxel (lane `sub` of its LPV lanes) over every slice
// profile and repetition; writes the voxel's echoes.  FORM: the sincos of
// the half flip angles (FispHalfAngle).  direct_echo: fisp_direct_echo of
// the wave.  SIMPLE: one slice profile, one repetition, no undersampling, no
// direct echo and at most 2^28 voxels, so every TR stores its echo once at a
// 32-bit byte offset from the TR's row, with no test in the loop.
template<int SPL, int LPV, int FORM, bool SIMPLE>
__device__ __forceinline__ void fisp_epg_voxel_form(const FispArgs& a, int voxel, int sub, bool direct_echo) {
    const float T1 = a.tissue.T1(voxel);
    const float T2 = a.tissue.T2(voxel);
    const float B1 = a.tissue.B1(voxel);
    const float B0 = a.tissue.B0(voxel);

    const float E1_TE = relax(a.TE, T1), E2_TE = relax(a.TE, T2);
    const float E1_TI = relax(a.TI, T1);
    const float E1_TRmTE = relax(a.TR - a.TE, T1), E2_TRmTE = relax(a.TR - a.TE, T2);
    const float E1_W = relax(a.TW, T1), E2_W = relax(a.TW, T2);

    // transverse decay+precession factors d = E2 * exp(i 2 pi t B0)
    const cfloat d_TE = off_resonance(a.TE, B0) * E2_TE;
    const cfloat d_TRmTE = off_resonance(a.TR - a.TE, B0) * E2_TRmTE;
    const cfloat d_W = off_resonance(a.TW, B0) * E2_W;
    const cfloat d_TR = dev::cmul(d_TE, d_TRmTE);
    const float E1_TR = E1_TE * E1_TRmTE;
    // regrowth of Z_0 accumulated over TE and TR-TE
    const float G = E1_TRmTE * (1.0f - E1_TE) + (1.0f - E1_TRmTE);

    const bool owns_zero = (sub == 0);  // lane holding state k = 0
    const int k0 = sub * SPL;           // global index of local state 0

    // 1 / d_TRmTE for the echo from the updated state (fisp_direct_echo)
    const float dd = d_TRmTE.re * d_TRmTE.re + d_TRmTE.im * d_TRmTE.im;
    const cfloat inv_TRmTE = direct_echo ? cfloat {0.0f, 0.0f}
                                         : cfloat {d_TRmTE.re / dd, -d_TRmTE.im / dd};

    for (int slice = 0; slice < a.nslices; slice++) {
        const float4* tab = a.table + size_t(slice) * size_t(a.nTR);
        // the first slice of a fully sampled sequence stores each echo once;
        // otherwise every undersampled copy, and later slices add to it
        const bool plain_store = (a.usf == 1) && (slice == 0);
        auto store_echo = [&](int n, cfloat echo) {
            if (plain_store) {
                a.echos[size_t(n) * size_t(a.tissue.nvoxels) + voxel] = echo;
            } else {
                for (int u = 0; u < a.usf; u++) {
                    cfloat* out = a.echos + size_t(n * a.usf + u) * size_t(a.tissue.nvoxels) + voxel;
                    *out = (slice == 0) ? echo : (*out + echo);
                }
            }
        };
        cfloat P[SPL], Q[SPL], Z[SPL];
#pragma unroll
        for (int i = 0; i < SPL; i++) {
            P[i] = {0, 0};
            Q[i] = {0, 0};
            Z[i] = {0, 0};
        }
        if (owns_zero) Z[0] = {1.0f, 0.0f};

        for (int rep = 0; rep < a.repetitions; rep++) {
            const bool sample = (rep == a.repetitions - 1);
            if (a.inversion) {  // invert (ignores B1), spoil, decay(TI), regrow
#pragma unroll
                for (int i = 0; i < SPL; i++) {
                    P[i] = {0, 0};
                    Q[i] = {0, 0};
                    Z[i] = Z[i] * (-E1_TI);
                }
                if (owns_zero) Z[0].re += 1.0f - E1_TI;
            }

            // the RF table entry of the next TR is loaded one TR ahead (the
            // table holds one entry more than the TRs of all slices)
            float4 t_next = (a.nTR > 0) ? tab[0] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            // SIMPLE: the TR's echo row, and the voxel's byte offset in it
            cfloat* row = a.echos;
            const uint32_t voff = uint32_t(voxel) * uint32_t(sizeof(cfloat));
            for (int n = 0; n < a.nTR; n++) {
                const float4 t = t_next;  // wave-uniform
                t_next = tab[n + 1];
                const float alpha = t.x * B1;
                float sx, cx;
                half_angle_sincos<FORM>(0.5f * alpha, &sx, &cx);
                const float sa = 2.0f * sx * cx;  // sin(alpha)
                const float omc = 2.0f * sx * sx;  // 1 - cos(alpha), without cancellation
                const float cp = t.y, sp = t.z;
                // rotation about (cos phi, sin phi, 0) by alpha, real XYZ basis
                const float R00 = 1.0f - omc * sp * sp, R01 = omc * cp * sp, R02 = sa * sp;
                const float R10 = R01, R11 = 1.0f - omc * cp * cp, R12 = -sa * cp;
                const float R20 = -sa * sp, R21 = sa * cp, R22 = 1.0f - omc;

                // M = D(TR) R ; rows 0/1 mix with the complex decay d = dr + i di
                const float dr = d_TR.re, di = d_TR.im;
                const float M00 = dr * R00 - di * R10, M01 = dr * R01 - di * R11, M02 = dr * R02 - di * R12;
                const float M10 = di * R00 + dr * R10, M11 = di * R01 + dr * R11, M12 = di * R02 + dr * R12;
                const float M20 = E1_TR * R20, M21 = E1_TR * R21, M22 = E1_TR * R22;
                // X = S/2, Y = -i D/2 with S = F+ + F-, D = F+ - F-
                const float h00 = 0.5f * M00, h01 = 0.5f * M01;
                const float h10 = 0.5f * M10, h11 = 0.5f * M11;
                const float h20 = 0.5f * M20, h21 = 0.5f * M21;

                // S = F+ + F- and D = F+ - F- of local state i.  With one lane
                // per voxel (SPL = N) two of them are known: F+_0 = conj(F-_0)
                // (S_0 = (2 Re F-_0, 0), D_0 = (0, -2 Im F-_0), the same values
                // as the sums) and F-_{N-1} = 0, so F+_0 and F-_{N-1} are never
                // stored.
                auto sums = [&](int i, cfloat& Si, cfloat& Di) {
                    if (LPV == 1 && i == SPL - 1 && SPL > 1) {
                        Si = P[i];
                        Di = P[i];
                    } else {
                        Si = P[i] + Q[i];
                        Di = P[i] - Q[i];
                    }
                };

                // sample F+_0 at TE:  rows 0/1 of D(TE) R applied to state 0
                if (!SIMPLE && direct_echo && sample && owns_zero) {
                    const float er = d_TE.re, ei = d_TE.im;
                    const float T00 = er * R00 - ei * R10, T01 = er * R01 - ei * R11, T02 = er * R02 - ei * R12;
                    const float T10 = ei * R00 + er * R10, T11 = ei * R01 + er * R11, T12 = ei * R02 + er * R12;
                    cfloat S0, D0;
                    sums(0, S0, D0);
                    const cfloat z0 = Z[0];
                    const float xr = fmaf(0.5f * T00, S0.re, fmaf(0.5f * T01, D0.im, T02 * z0.re));
                    const float xi = fmaf(0.5f * T00, S0.im, fmaf(-0.5f * T01, D0.re, T02 * z0.im));
                    const float yr = fmaf(0.5f * T10, S0.re, fmaf(0.5f * T11, D0.im, T12 * z0.re));
                    const float yi = fmaf(0.5f * T10, S0.im, fmaf(-0.5f * T11, D0.re, T12 * z0.im));
                    const cfloat echo = {xr - yi, xi + yr};
                    store_echo(n, echo);
                }

                // the states are updated from the top down, and each state's
                // outputs go straight to their dephased slots (F+ of state i to
                // i + 1, F- to i - 1), written after the value they replace has
                // been read: the sums of state i - 1 are formed before state i's
                // outputs are stored, so no state moves between registers
                cfloat pn_top = {0, 0}, qn_bottom = {0, 0}, pn0 = {0, 0};
                cfloat S, D;
                sums(SPL - 1, S, D);
#pragma unroll
                for (int i = SPL - 1; i >= 0; i--) {
                    cfloat Sn = {0, 0}, Dn = {0, 0};
                    if (i > 0) sums(i - 1, Sn, Dn);
                    const cfloat z = Z[i];
                    const float xr = fmaf(h00, S.re, fmaf(h01, D.im, M02 * z.re));
                    const float yr = fmaf(h10, S.re, fmaf(h11, D.im, M12 * z.re));
                    const float zr = fmaf(h20, S.re, fmaf(h21, D.im, M22 * z.re));
                    float xi, yi, zi;
                    if (false) {
                        xi = M02 * z.im;
                        yi = M12 * z.im;
                        zi = M22 * z.im;
                    } else {
                        xi = fmaf(h00, S.im, fmaf(-h01, D.re, M02 * z.im));
                        yi = fmaf(h10, S.im, fmaf(-h11, D.re, M12 * z.im));
                        zi = fmaf(h20, S.im, fmaf(-h21, D.re, M22 * z.im));
                    }
                    const cfloat pn = {xr - yi, xi + yr}, qn = {xr + yi, xi - yr};
                    if (i == 0) pn0 = pn;
                    if (i + 1 < SPL) {
                        P[i + 1] = pn;  // dephasing: F+_{i+1} <- F+_i
                    } else {
                        pn_top = pn;
                    }
                    if (i > 0) {
                        Q[i - 1] = qn;  // dephasing: F-_{i-1} <- F-_i
                    } else {
                        qn_bottom = qn;
                    }
                    Z[i] = {zr, zi};
                    S = Sn;
                    D = Dn;
                }
                if (owns_zero) Z[0].re += G;
                if constexpr (SIMPLE) {
                    if (owns_zero) {
                        *reinterpret_cast<cfloat*>(reinterpret_cast<char*>(row) + voff) = dev::cmul(pn0, inv_TRmTE);
                    }
                    row += a.tissue.nvoxels;
                } else if (!direct_echo && sample && owns_zero) {
                    const cfloat echo = dev::cmul(pn0, inv_TRmTE);
                    store_echo(n, echo);
                }

                // the ends of the dephasing: F+ into state 0 and F- into the top
                // state come from the neighbouring lane (or are zero), F-_{N-1} <- 0,
                // F+_0 <- conj(F-_0)
                cfloat p_in = {0, 0}, q_in = {0, 0};
                if constexpr (LPV > 1) {
                    p_in.re = __shfl_up(pn_top.re, 1, LPV);
                    p_in.im = __shfl_up(pn_top.im, 1, LPV);
                    q_in.re = __shfl_down(qn_bottom.re, 1, LPV);
                    q_in.im = __shfl_down(qn_bottom.im, 1, LPV);
                    if (sub == LPV - 1) q_in = {0, 0};
                }
                P[0] = p_in;
                Q[SPL - 1] = q_in;
                if constexpr (LPV > 1) {
#pragma unroll
                    for (int i = 0; i < SPL; i++) {
                        if (k0 + i >= a.N - 1) Q[i] = {0, 0};
                    }
                }
                if (owns_zero) P[0] = conj(Q[0]);
            }

            if (a.wait_spoiling) {
#pragma unroll
                for (int i = 0; i < SPL; i++) {
                    P[i] = {0, 0};
                    Q[i] = {0, 0};
                }
            }
#pragma unroll
            for (int i = 0; i < SPL; i++) {
                P[i] = dev::cmul(P[i], d_W);
                Q[i] = dev::cmul(Q[i], conj(d_W));
                Z[i] = Z[i] * E1_W;
            }
            if (owns_zero) Z[0].re += 1.0f - E1_W;
        }
    }
}

// fisp_epg_voxel_form in the form the wave needs (fisp_wave_form), decided
// once per wave so that the TR loop carries no test for it.
template<int SPL, int LPV>
__device__ __forceinline__ void fisp_epg_voxel(const FispArgs& a, int voxel, int sub, float max_angle) {
    const FispWaveForm f = fisp_wave_form(a, voxel, max_angle);
    if (f.wide) {
        fisp_epg_voxel_form<SPL, LPV, HALF_WIDE, false>(a, voxel, sub, f.direct_echo);
    } else if (!f.simple) {
        fisp_epg_voxel_form<SPL, LPV, HALF_POLY, false>(a, voxel, sub, f.direct_echo);
    } else if (!f.q0) {
        fisp_epg_voxel_form<SPL, LPV, HALF_POLY, true>(a, voxel, sub, false);
    } else {
        fisp_epg_voxel_form<SPL, LPV, HALF_Q0, true>(a, voxel, sub, false);
    }
}

}  // namespace kernels
}  // namespace compas
