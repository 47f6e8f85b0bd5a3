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
#include "compas/device/fisp_epg.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

// table[i] for i < nslices nTR, then a zero entry at nslices nTR (the
// recursion loads the entry of the TR after its last one); nslices nTR + 1
// threads.
__global__ void fisp_rf_table(const cfloat* RF, const cfloat* profiles, int nTR, int nslices, float4* table) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nTR * nslices) return;
    if (i == nTR * nslices) {
        table[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
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
    const float max_angle = fisp_table_max_angle(a);  // while every lane is active
    if (voxel >= a.tissue.nvoxels) return;  // whole LPV groups leave together
    fisp_epg_voxel<SPL, LPV>(a, voxel, sub, max_angle);
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
    auto table = ctx.empty<float4, 1>({int64_t(nslices) * nTR + 1});
    KernelScope ks(ctx, "fisp_rf_table");
    const int total = nslices * nTR + 1;
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

static __device__ __align__(256) float4 scratch[kRfTableScratchEntries + 1];  // and the terminating zero entry

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
    const int total = int(nslices * nTR) + 1;
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
