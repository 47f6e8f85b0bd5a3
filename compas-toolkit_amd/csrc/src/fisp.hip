// FISP extended-phase-graph simulation on gfx950.
//
// Semantics: reference src/simulate/fisp_kernels.cuh:99-244, fisp.cu:7-140 and
// the EPG operators of src/operators/epg.cuh:19-322 (see the oracle for the
// literal restatement).
//
// Design (MI355X):
//   * One voxel per lane (or LPV lanes per voxel for large state counts or
//     small voxel counts).  The F+/F-/Z states of a voxel live in VGPRs; the
//     dephasing shift is a register rename inside the unrolled update.
//   * The per-TR operator "excite -> decay(TE) -> regrow -> decay(TR-TE) ->
//     regrow" is one real 3x3 matrix M = D(TR) R(alpha, phi) acting on the
//     state written in the real transverse basis X = (F+ + F-)/2,
//     Y = (F+ - F-)/(2i): 18 FMAs per state instead of ~40 for the complex
//     F+/F-/Z form; the basis change costs 8 adds.  The TE sample F+_0(TE) uses
//     the first two rows of D(TE) R on state 0 only.
//   * The RF table (alpha per unit B1, cos phi, sin phi) is precomputed per
//     (slice, TR) and read with wave-uniform (scalar) loads.
//   * Echoes are written directly in [nreadouts][nvoxels] layout (coalesced
//     over voxels), including the undersampling expansion, so no transpose
//     pass is needed.

#include <cfloat>
#include <cmath>

#include <cstdlib>
#include <mutex>
#include <sstream>

#include "compas/device/common.h"
#include "compas/device/fisp_common.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

__global__ void fisp_rf_table(const cfloat* RF, const cfloat* profiles, int nTR, int nslices, float4* table) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nTR * nslices) return;
    int n = i % nTR;
    cfloat p = profiles[i], r = RF[n];
    cfloat rf = {p.re * r.re - p.im * r.im, p.re * r.im + p.im * r.re};
    float a = hypotf(rf.re, rf.im);
    float c = 1.0f, s = 0.0f;
    if (a > FLT_MIN) {  // epg.cuh:64-67
        c = rf.re / a;
        s = rf.im / a;
    }
    table[i] = make_float4(dev::DEG2RAD_F * a, c, s, 0.0f);
}

template<int SPL, int LPV>
__global__ void __launch_bounds__(64) fisp_epg(FispArgs a) {
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

    // F+_0 at TE from the updated state: over TR - TE the transverse state
    // only decays and precesses (F+ times d_TRmTE), so F+_0(TE) = F+_0(TR) /
    // d_TRmTE, one complex multiply per TR instead of a second 2x3 operator.
    // A wave with a lane whose d_TRmTE is too small to invert (T2 of a few
    // microseconds) applies the operator D(TE) R to state 0 instead.
    const float dd = d_TRmTE.re * d_TRmTE.re + d_TRmTE.im * d_TRmTE.im;
    const bool direct_echo = __any(!(dd > 1e-30f));  // wave-uniform
    const cfloat inv_TRmTE = direct_echo ? cfloat {0.0f, 0.0f}
                                         : cfloat {d_TRmTE.re / dd, -d_TRmTE.im / dd};

    for (int slice = 0; slice < a.nslices; slice++) {
        const float4* tab = a.table + size_t(slice) * size_t(a.nTR);
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

            for (int n = 0; n < a.nTR; n++) {
                const float4 t = tab[n];  // wave-uniform
                const float alpha = t.x * B1;
                float sx, cx;
                dev::sincos_fast(0.5f * alpha, &sx, &cx);
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

                // sample F+_0 at TE:  rows 0/1 of D(TE) R applied to state 0
                if (direct_echo && sample && owns_zero) {
                    const float er = d_TE.re, ei = d_TE.im;
                    const float T00 = er * R00 - ei * R10, T01 = er * R01 - ei * R11, T02 = er * R02 - ei * R12;
                    const float T10 = ei * R00 + er * R10, T11 = ei * R01 + er * R11, T12 = ei * R02 + er * R12;
                    const cfloat S0 = P[0] + Q[0], D0 = P[0] - Q[0];
                    const cfloat z0 = Z[0];
                    const float xr = fmaf(0.5f * T00, S0.re, fmaf(0.5f * T01, D0.im, T02 * z0.re));
                    const float xi = fmaf(0.5f * T00, S0.im, fmaf(-0.5f * T01, D0.re, T02 * z0.im));
                    const float yr = fmaf(0.5f * T10, S0.re, fmaf(0.5f * T11, D0.im, T12 * z0.re));
                    const float yi = fmaf(0.5f * T10, S0.im, fmaf(-0.5f * T11, D0.re, T12 * z0.im));
                    const cfloat echo = {xr - yi, xi + yr};
                    for (int u = 0; u < a.usf; u++) {
                        cfloat* out = a.echos + size_t(n * a.usf + u) * size_t(a.tissue.nvoxels) + voxel;
                        *out = (slice == 0) ? echo : (*out + echo);
                    }
                }

                cfloat Pn[SPL], Qn[SPL];
#pragma unroll
                for (int i = 0; i < SPL; i++) {
                    const cfloat S = P[i] + Q[i], D = P[i] - Q[i];
                    const cfloat z = Z[i];
                    const float xr = fmaf(h00, S.re, fmaf(h01, D.im, M02 * z.re));
                    const float xi = fmaf(h00, S.im, fmaf(-h01, D.re, M02 * z.im));
                    const float yr = fmaf(h10, S.re, fmaf(h11, D.im, M12 * z.re));
                    const float yi = fmaf(h10, S.im, fmaf(-h11, D.re, M12 * z.im));
                    const float zr = fmaf(h20, S.re, fmaf(h21, D.im, M22 * z.re));
                    const float zi = fmaf(h20, S.im, fmaf(-h21, D.re, M22 * z.im));
                    Pn[i] = {xr - yi, xi + yr};
                    Qn[i] = {xr + yi, xi - yr};
                    Z[i] = {zr, zi};
                }
                if (owns_zero) Z[0].re += G;
                if (!direct_echo && sample && owns_zero) {
                    const cfloat echo = dev::cmul(Pn[0], inv_TRmTE);
                    for (int u = 0; u < a.usf; u++) {
                        cfloat* out = a.echos + size_t(n * a.usf + u) * size_t(a.tissue.nvoxels) + voxel;
                        *out = (slice == 0) ? echo : (*out + echo);
                    }
                }

                // dephasing: F+_k <- F+_{k-1}, F-_k <- F-_{k+1}, F-_{N-1} <- 0, F+_0 <- conj(F-_0)
                cfloat p_in = {0, 0}, q_in = {0, 0};
                if constexpr (LPV > 1) {
                    p_in.re = __shfl_up(Pn[SPL - 1].re, 1, LPV);
                    p_in.im = __shfl_up(Pn[SPL - 1].im, 1, LPV);
                    q_in.re = __shfl_down(Qn[0].re, 1, LPV);
                    q_in.im = __shfl_down(Qn[0].im, 1, LPV);
                    if (sub == LPV - 1) q_in = {0, 0};
                }
#pragma unroll
                for (int i = SPL - 1; i >= 1; i--) P[i] = Pn[i - 1];
                P[0] = p_in;
#pragma unroll
                for (int i = 0; i < SPL - 1; i++) Q[i] = Qn[i + 1];
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

}  // namespace kernels

using FispLauncher = void (*)(kernels::FispArgs, unsigned, hipStream_t);

template<int SPL, int LPV>
static void launch_fisp(kernels::FispArgs a, unsigned blocks, hipStream_t stream) {
    kernels::fisp_epg<SPL, LPV><<<blocks, 64, 0, stream>>>(a);
}

#define COMPAS_FISP_CASE(S, L) \
    case S:                    \
        return &launch_fisp<S, L>;

static FispLauncher fisp_launcher_lpv1(int spl) {
    switch (spl) {
        COMPAS_FISP_CASE(1, 1)
        COMPAS_FISP_CASE(2, 1)
        COMPAS_FISP_CASE(3, 1)
        COMPAS_FISP_CASE(4, 1)
        COMPAS_FISP_CASE(5, 1)
        COMPAS_FISP_CASE(6, 1)
        COMPAS_FISP_CASE(7, 1)
        COMPAS_FISP_CASE(8, 1)
        COMPAS_FISP_CASE(9, 1)
        COMPAS_FISP_CASE(10, 1)
        COMPAS_FISP_CASE(11, 1)
        COMPAS_FISP_CASE(12, 1)
        COMPAS_FISP_CASE(13, 1)
        COMPAS_FISP_CASE(14, 1)
        COMPAS_FISP_CASE(15, 1)
        COMPAS_FISP_CASE(16, 1)
        COMPAS_FISP_CASE(17, 1)
        COMPAS_FISP_CASE(18, 1)
        COMPAS_FISP_CASE(19, 1)
        COMPAS_FISP_CASE(20, 1)
        COMPAS_FISP_CASE(21, 1)
        COMPAS_FISP_CASE(22, 1)
        COMPAS_FISP_CASE(23, 1)
        COMPAS_FISP_CASE(24, 1)
        COMPAS_FISP_CASE(25, 1)
        COMPAS_FISP_CASE(26, 1)
        COMPAS_FISP_CASE(27, 1)
        COMPAS_FISP_CASE(28, 1)
        COMPAS_FISP_CASE(29, 1)
        COMPAS_FISP_CASE(30, 1)
        COMPAS_FISP_CASE(31, 1)
        COMPAS_FISP_CASE(32, 1)
        default:
            return nullptr;
    }
}

static FispLauncher fisp_launcher_lpv2(int spl) {
    switch (spl) {
        COMPAS_FISP_CASE(1, 2)
        COMPAS_FISP_CASE(2, 2)
        COMPAS_FISP_CASE(3, 2)
        COMPAS_FISP_CASE(4, 2)
        COMPAS_FISP_CASE(5, 2)
        COMPAS_FISP_CASE(6, 2)
        COMPAS_FISP_CASE(7, 2)
        COMPAS_FISP_CASE(8, 2)
        COMPAS_FISP_CASE(9, 2)
        COMPAS_FISP_CASE(10, 2)
        COMPAS_FISP_CASE(11, 2)
        COMPAS_FISP_CASE(12, 2)
        COMPAS_FISP_CASE(13, 2)
        COMPAS_FISP_CASE(14, 2)
        COMPAS_FISP_CASE(15, 2)
        COMPAS_FISP_CASE(16, 2)
        COMPAS_FISP_CASE(32, 2)
        default:
            return nullptr;
    }
}

static FispLauncher fisp_launcher_lpv4(int spl) {
    switch (spl) {
        COMPAS_FISP_CASE(1, 4)
        COMPAS_FISP_CASE(2, 4)
        COMPAS_FISP_CASE(3, 4)
        COMPAS_FISP_CASE(4, 4)
        COMPAS_FISP_CASE(5, 4)
        COMPAS_FISP_CASE(6, 4)
        COMPAS_FISP_CASE(7, 4)
        COMPAS_FISP_CASE(8, 4)
        COMPAS_FISP_CASE(32, 4)
        default:
            return nullptr;
    }
}

Array<float4, 1> fisp_rf_table(const CompasContext& ctx, const FISPSequence& seq) {
    const int nTR = int(seq.RF_train.size(0)), nslices = int(seq.sliceprofiles.size(0));
    auto table = ctx.empty<float4, 1>({int64_t(nslices) * nTR});
    if (table.numel() == 0) return table;
    KernelScope ks(ctx, "fisp_rf_table");
    const int total = nslices * nTR;
    kernels::fisp_rf_table<<<dev::div_ceil(total, 256), 256, 0, ctx.stream()>>>(
        seq.RF_train.data(), seq.sliceprofiles.data(), nTR, nslices, table.data());
    check_launch("fisp_rf_table");
    return table;
}

static int simd_count(int device) {
    static int cached[64] = {0};
    if (device >= 0 && device < 64 && cached[device]) return cached[device];
    hipDeviceProp_t prop;
    COMPAS_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    int n = prop.multiProcessorCount * 4;  // 4 SIMDs per CU on CDNA
    if (device >= 0 && device < 64) cached[device] = n;
    return n;
}

// Lanes per voxel: pick the split with the lowest modelled time,
// ceil(waves / SIMDs) * per-wave cost, cost ~ 26 VALU ops per local state per
// TR plus ~90 ops of per-TR matrix set-up (duplicated in each lane of a voxel).
// A wave alone on its SIMD issues at about half the rate of two co-resident
// ones (50176 voxels x 25 states back to back: 1.90 ms in one lane per voxel,
// 1.45 ms in two), but in the simulate + signal loop the one-lane layout is
// faster (2.00-2.03 against 2.03-2.10 ms per launch after the contraction,
// 2.30 against 2.43 ms per pipelined step): the model keeps one lane there.
static void choose_fisp_layout(int N, int nvoxels, int simds, int* spl, int* lpv) {
    if (N > 64) {
        *spl = 32;
        *lpv = 4;
        return;
    }
    if (N > 32) {
        *spl = 32;
        *lpv = 2;
        return;
    }
    double best = 1e300;
    for (int l : {1, 2, 4}) {
        int s = (N + l - 1) / l;
        if (l == 2 && s > 16) continue;
        if (l == 4 && s > 8) continue;
        const double waves = double(nvoxels) * l / 64.0;
        const double rounds = std::ceil(waves / simds);
        const double cost = rounds * (26.0 * s + 90.0 + (l > 1 ? 12.0 : 0.0));
        if (cost < best - 1e-9) {
            best = cost;
            *spl = s;
            *lpv = l;
        }
    }
}

static void check_fisp_sequence(const FISPSequence& seq) {
    COMPAS_CHECK(seq.max_state >= 1);
    if (seq.max_state > 128) COMPAS_ERROR("max_state cannot exceed 128");  // fisp.cu:100-102
    COMPAS_CHECK(seq.sliceprofiles.size(1) == seq.RF_train.size(0));  // fisp.cu:22
    COMPAS_CHECK(seq.undersampling_factor >= 1);
    COMPAS_CHECK(seq.repetitions >= 1);
}

// Launches the recursion for every voxel: `table` holds the RF table of
// `seq` (fisp_rf_table), `echos` the [nTR * usf][nvoxels] output.
static void launch_fisp_epg(
    const CompasContext& ctx, const TissueParameters& parameters, const FISPSequence& seq, const float4* table,
    cfloat* echos, bool timed);

Array<cfloat, 2> simulate_magnetization(
    const CompasContext& ctx,
    TissueParameters parameters,
    FISPSequence seq) {
    check_fisp_sequence(seq);
    const int nTR = int(seq.RF_train.size(0));
    const int nv = parameters.nvoxels;
    const int nslices = int(seq.sliceprofiles.size(0));
    const int nr = nTR * seq.undersampling_factor;
    ctx.activate();

    auto echos = ctx.empty<cfloat, 2>({nr, nv});
    if (nv == 0 || nr == 0) return echos;
    if (nslices == 0) {
        ctx.fill_zero(echos.data(), echos.nbytes());
        return echos;
    }

    auto table = fisp_rf_table(ctx, seq);
    launch_fisp_epg(ctx, parameters, seq, table.data(), echos.data(), true);
    return echos;
}

// ---------------------------------------------------------------------------
// Allocation-free variant: the RF table lives in a file-scope device array
// instead of the context's pool, and the caller provides the echo buffer.
//
// Each GPU holds its own copy of `scratch` for the life of the process.  In
// host code the name `scratch` is a host-side stand-in, never a device
// address: the per-device address comes from hipGetSymbolAddress, looked up
// once per device with that device current, and only that address is passed
// to kernels.  Calls on one GPU share the array, so calls that overlap on one
// GPU (different streams or threads without ordering) race; calls on one
// context's stream are ordered and safe.

static constexpr int64_t kRfTableScratchEntries = int64_t(1) << 16;  // (slice, TR) pairs: 1 MiB

static __device__ __align__(256) float4 scratch[kRfTableScratchEntries];

static float4* scratch_address(int device) {
    static std::mutex mutex;
    static float4* address[256] = {};
    if (device < 0 || device >= 256) COMPAS_ERROR("device ordinal out of range");
    std::lock_guard<std::mutex> lock(mutex);
    if (address[device] == nullptr) {
        void* p = nullptr;
        COMPAS_HIP_CHECK(hipSetDevice(device));
        COMPAS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (p == nullptr) COMPAS_ERROR("no device address for the RF-table scratch array");
        address[device] = static_cast<float4*>(p);
    }
    return address[device];
}

void simulate_magnetization_into(
    const CompasContext& ctx, const TissueParameters& parameters, const FISPSequence& seq, cfloat* echos,
    int64_t capacity) {
    check_fisp_sequence(seq);
    const int64_t nTR = seq.RF_train.size(0);
    const int64_t nv = parameters.nvoxels;
    const int64_t nslices = seq.sliceprofiles.size(0);
    const int64_t nr = nTR * seq.undersampling_factor;
    if (nslices * nTR > kRfTableScratchEntries) {
        std::stringstream ss;
        ss << "sequence too large for the allocation-free FISP simulation: " << nslices << " slice profiles x " << nTR
           << " TRs = " << nslices * nTR << " RF-table entries, at most " << kRfTableScratchEntries;
        throw CompasException(ss.str());
    }
    if (capacity < nr * nv) {
        std::stringstream ss;
        ss << "echo buffer too small: " << capacity << " complex values for [" << nr << "][" << nv << "]";
        throw CompasException(ss.str());
    }
    if (nr * nv == 0) return;
    if (echos == nullptr) COMPAS_ERROR("null echo buffer");
    ctx.activate();
    if (nslices == 0) {
        ctx.fill_zero(echos, size_t(nr * nv) * sizeof(cfloat));
        return;
    }
    float4* table = scratch_address(ctx.device());
    const int total = int(nslices * nTR);
    kernels::fisp_rf_table<<<dev::div_ceil(total, 256), 256, 0, ctx.stream()>>>(
        seq.RF_train.data(), seq.sliceprofiles.data(), int(nTR), int(nslices), table);
    check_launch("fisp_rf_table");
    launch_fisp_epg(ctx, parameters, seq, table, echos, false);
}

static void launch_fisp_epg(
    const CompasContext& ctx, const TissueParameters& parameters, const FISPSequence& seq, const float4* table,
    cfloat* echos, bool timed) {
    const int nTR = int(seq.RF_train.size(0));
    const int nv = parameters.nvoxels;
    const int nslices = int(seq.sliceprofiles.size(0));

    kernels::FispArgs a;
    a.tissue = dev::tissue_view(parameters);
    a.table = table;
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
    a.echos = echos;

    int spl = 0, lpv = 0;
    choose_fisp_layout(seq.max_state, nv, simd_count(ctx.device()), &spl, &lpv);
    if (const char* e = std::getenv("COMPAS_FISP_LPV")) {  // layout override (1, 2 or 4 lanes per voxel)
        const int l = std::atoi(e);
        const int s = (seq.max_state + l - 1) / l;
        if ((l == 1 && s <= 32) || (l == 2 && (s <= 16 || s == 32)) || (l == 4 && (s <= 8 || s == 32))) {
            spl = (l > 1 && s > (l == 2 ? 16 : 8)) ? 32 : s;
            lpv = l;
        }
    }
    FispLauncher fn = lpv == 1 ? fisp_launcher_lpv1(spl) : lpv == 2 ? fisp_launcher_lpv2(spl) : fisp_launcher_lpv4(spl);
    COMPAS_CHECK(fn != nullptr);
    unsigned blocks = dev::div_ceil(int64_t(nv) * lpv, 64);
    if (timed) {
        KernelScope ks(ctx, "fisp_epg");
        fn(a, blocks, ctx.stream());
        check_launch("fisp_epg");
    } else {
        fn(a, blocks, ctx.stream());
        check_launch("fisp_epg");
    }
}

}  // namespace compas
