// pSSFP isochromat (Bloch) simulation on gfx950.
//
// Semantics: reference src/simulate/pssfp_kernels.cuh:14-129, pssfp.cu:7-91
// and src/operators/isochromat.cuh:9-128.
//
// Design (MI355X):
//   * One wave per workgroup: 64 voxels (one per lane) times one group of ZPL
//     slice positions, ZPL independent isochromats per lane (instruction-level
//     parallelism).  Waves share nothing, so the hardware packs them onto the
//     SIMDs freely; a workgroup of all G = ceil(nz / ZPL) groups of a voxel
//     block would have to be co-resident (at 35 positions and ZPL 7: five
//     waves, one workgroup per CU, one SIMD with two waves and three with one,
//     and four rounds of 256 workgroups for 784 blocks, the last nearly empty).
//   * Group 0 writes the echoes; groups 1..G-1 write FP32 partial echoes that
//     a second pass adds in FP64, in group order (deterministic).  That costs
//     (G-1) x 8 x nTR x nv bytes written and read once: 0.5 ms for the sum
//     at 35 positions and 1120 x 50176, against 73 ms for the simulation.
//   * ZPL is chosen per launch from the resident waves per CU of each
//     instance (occupancy API) and the rounds of waves it takes.
//   * The isochromat state is FP64.  On MI355X, FP64 FMA issues at the same
//     rate as non-packed FP32, and the FP32 form of this recursion drifts
//     ~3e-5 (relative L2) from the exact result over 1000+ TRs because each TR
//     applies nRF Rodrigues rotations with little T2 damping.
//   * Rodrigues' rotation is evaluated without sqrt/rcp/sin/cos for the small
//     per-step angles: with u = |a|^2, s1(u) = sin(t)/t and s2(u) =
//     (1 - cos t)/t^2 are Taylor polynomials in u (4, 5, 6 or 7 terms for |t|
//     up to 1/27, 1/9, 1/4, 1/2; error < 1e-17), cos t = 1 - u s2(u), and
//       m' = cos t m + s1 (a x m) + s2 (a . m) a,
//     which equals the reference's normalised form.  Larger angles fall back
//     to FP64 sincos; the choice is made once per RF sub-step for the whole
//     wave, so the common case is branch-free and the compiler interleaves
//     the ZPL independent FP64 dependency chains.
//   * The RF sub-pulse table is staged in LDS as doubles (broadcast reads in
//     the sub-step loop instead of a global load per sub-step).

#include <algorithm>

#include "compas/device/common.h"
#include "compas/device/pssfp_common.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

// One wave: voxels 64 vblock .. 64 vblock + 63 (one per lane) times slice
// positions ZPL group .. ZPL group + ZPL - 1; writes the sum over those
// positions of every TR's echo to out[TR][voxel].  grf: LDS for RF_CHUNK
// waveform samples.
template<int ZPL>
__device__ __forceinline__ void pssfp_iso_block(const PssfpArgs& a, int group, int vblock, cfloat* out, double2* grf) {
    const int lane = threadIdx.x;
    const int voxel = vblock * 64 + lane;
    const bool active = voxel < a.tissue.nvoxels;
    const int vx = active ? voxel : 0;

    const double T1 = a.tissue.T1(vx), T2 = a.tissue.T2(vx);
    const double B1 = a.tissue.B1(vx), B0 = a.tissue.B0(vx);
    const double E1_ex = exp(-double(a.dt_ex) / T1), E2_ex = exp(-double(a.dt_ex) / T2);
    const double E1_pr = exp(-double(a.dt_pr) / T1), E2_pr = exp(-double(a.dt_pr) / T2);
    const double E1_inv = exp(-double(a.dt_inv) / T1), E2_inv = exp(-double(a.dt_inv) / T2);
    const double R1_ex = 1.0 - E1_ex, R1_pr = 1.0 - E1_pr, R1_inv = 1.0 - E1_inv;
    const double off_ex = double(a.dt_ex) * dev::TWO_PI * B0;
    const double off_pr = double(a.dt_pr) * dev::TWO_PI * B0;

    Vec3d m[ZPL];
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
        // m = (0,0,1); invert(p): m_z *= cos(pi B1); decay + regrowth over dt_inv
        m[i] = {0.0, 0.0, cos(M_PI * B1)};
        relax3(m[i], E1_inv, E2_inv, R1_inv);
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
                // wave-uniform choice of the series: the common small-angle
                // cases are straight-line code over the ZPL isochromats
                if (!__any(umax > SERIES_LIMIT_4)) {
#pragma unroll
                    for (int i = 0; i < ZPL; i++) {
                        double s1, s2;
                        rotation_series_4(u[i], s1, s2);
                        apply_rotation(m[i], ax, ay, az[i], u[i], s1, s2);
                        relax3(m[i], E1_ex, E2_ex, R1_ex);
                    }
                } else if (!__any(umax > SERIES_LIMIT_5)) {
#pragma unroll
                    for (int i = 0; i < ZPL; i++) {
                        double s1, s2;
                        rotation_series_5(u[i], s1, s2);
                        apply_rotation(m[i], ax, ay, az[i], u[i], s1, s2);
                        relax3(m[i], E1_ex, E2_ex, R1_ex);
                    }
                } else if (!__any(umax > SERIES_LIMIT)) {
#pragma unroll
                    for (int i = 0; i < ZPL; i++) {
                        double s1, s2;
                        rotation_series(u[i], s1, s2);
                        apply_rotation(m[i], ax, ay, az[i], u[i], s1, s2);
                        relax3(m[i], E1_ex, E2_ex, R1_ex);
                    }
                } else if (!__any(umax > SERIES_LIMIT_WIDE)) {
#pragma unroll
                    for (int i = 0; i < ZPL; i++) {
                        double s1, s2;
                        rotation_series_wide(u[i], s1, s2);
                        apply_rotation(m[i], ax, ay, az[i], u[i], s1, s2);
                        relax3(m[i], E1_ex, E2_ex, R1_ex);
                    }
                } else {
#pragma unroll
                    for (int i = 0; i < ZPL; i++) {
                        double s1, s2;
                        rotation_any(u[i], s1, s2);
                        apply_rotation(m[i], ax, ay, az[i], u[i], s1, s2);
                        relax3(m[i], E1_ex, E2_ex, R1_ex);
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
            rotate_z(m[i], c, s);
            relax3(m[i], E1_pr, E2_pr, R1_pr);
        }
    }

    for (int n = 0; n < a.nTR; n++) {
        const cfloat rf = a.RF[n];
        excite(double(rf.re), double(rf.im));
        double sx = 0.0, sy = 0.0;
#pragma unroll
        for (int i = 0; i < ZPL; i++) {
            rotate_z(m[i], cpr[i], spr[i]);
            relax3(m[i], E1_pr, E2_pr, R1_pr);
            if (valid[i]) {
                sx += m[i].x;
                sy += m[i].y;
            }
        }
        if (active) out[size_t(n) * size_t(a.tissue.nvoxels) + voxel] = {float(sx), float(sy)};
#pragma unroll
        for (int i = 0; i < ZPL; i++) {
            rotate_z(m[i], cpr[i], spr[i]);
            relax3(m[i], E1_pr, E2_pr, R1_pr);
        }
    }
}

// grid (ceil(nv / 64), groups in this launch), block 64.  Group g of the
// launch is group0 + blockIdx.y; blockIdx.y == 0 writes out0, blockIdx.y = j
// > 0 writes rest + (j - 1) * nTR * nv.
template<int ZPL>
__global__ void __launch_bounds__(64) pssfp_iso(PssfpArgs a, int group0, cfloat* out0, cfloat* rest) {
    __shared__ double2 grf[RF_CHUNK];  // gamma_dt_RF as doubles (LDS broadcast reads)
    const size_t plane = size_t(a.nTR) * size_t(a.tissue.nvoxels);
    cfloat* out = (blockIdx.y == 0) ? out0 : rest + size_t(blockIdx.y - 1) * plane;
    pssfp_iso_block<ZPL>(a, group0 + int(blockIdx.y), int(blockIdx.x), out, grf);
}

// Work-table form (request server): workgroup b runs item (job, voxel block,
// group) = items[b].xyz of jobs[job]; group 0 writes the job's echoes, group
// g > 0 the partial plane g - 1.  The jobs of one launch may belong to
// different requests with any voxel, TR and RF counts; a wave's arithmetic
// is that of pssfp_iso for the same (voxel block, group).
template<int ZPL>
__global__ void __launch_bounds__(64) pssfp_iso_table(const PssfpTableJob* __restrict__ jobs, const int4* __restrict__ items) {
    __shared__ double2 grf[RF_CHUNK];
    const int4 it = items[blockIdx.x];
    const PssfpTableJob job = jobs[it.x];
    const size_t plane = size_t(job.a.nTR) * size_t(job.a.tissue.nvoxels);
    cfloat* out = it.z == 0 ? job.a.echos : job.partial + size_t(it.z - 1) * plane;
    pssfp_iso_block<ZPL>(job.a, it.z, it.y, out, grf);
}

// echos[i] += sum_j partial[j][i], accumulated in FP64 in slot order
__global__ void pssfp_sum_groups_kernel(
    cfloat* __restrict__ echos, const cfloat* __restrict__ partial, int count, int64_t n, int64_t stride) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const cfloat e = echos[i];
        double re = e.re, im = e.im;
        for (int j = 0; j < count; j++) {
            const cfloat p = partial[size_t(j) * size_t(stride) + size_t(i)];
            re += p.re;
            im += p.im;
        }
        echos[i] = {float(re), float(im)};
    }
}

}  // namespace kernels

void pssfp_sum_groups(
    const CompasContext& ctx, cfloat* echos, const cfloat* partial, int count, int64_t n, int64_t stride) {
    if (count <= 0 || n == 0) return;
    KernelScope ks(ctx, "pssfp_sum_groups");
    const int64_t blocks = std::min<int64_t>(dev::div_ceil(n, int64_t(256)), 8192);
    kernels::pssfp_sum_groups_kernel<<<unsigned(blocks), 256, 0, ctx.stream()>>>(echos, partial, count, n, stride);
    check_launch("pssfp_sum_groups");
}

PssfpLayout pssfp_layout(const CompasContext& ctx, int nz, int nv, int nTR, const int* waves_per_cu, int max_zpl) {
    hipDeviceProp_t prop;
    COMPAS_HIP_CHECK(hipGetDeviceProperties(&prop, ctx.device()));
    const int64_t cus = prop.multiProcessorCount;
    const int64_t blocks = dev::div_ceil(int64_t(nv), int64_t(64));
    // partial echoes held at once (one plane per group of a launch, launch 0
    // writing its first group into the echoes): at most a sixth of the
    // device's memory (48 GB of 288) and at least 8 GiB, and at least one
    // plane.  A launch split for memory costs whole extra rounds of waves.
    const int64_t plane = int64_t(8) * nTR * std::max(nv, 1);
    const int64_t budget = std::max<int64_t>(int64_t(8) << 30, int64_t(prop.totalGlobalMem / 6));
    const int cap = int(std::max<int64_t>(1, std::min<int64_t>(1 << 15, budget / plane)));
    PssfpLayout best {1, nz, 1};
    double best_cost = 1e300;
    for (int zpl = max_zpl; zpl >= 1; zpl--) {
        const int wpc = std::max(1, waves_per_cu[zpl]);
        const int groups = (nz + zpl - 1) / zpl;
        const int per_launch = std::min(groups, cap);
        // every wave does the same work, issue-bound at (zpl * 50 + 12) FP64
        // instructions per RF sub-step; wpc / 4 waves share a SIMD
        double cost = 0.0;
        for (int g0 = 0; g0 < groups; g0 += per_launch) {
            const int64_t waves = blocks * std::min(per_launch, groups - g0);
            const int64_t rounds = dev::div_ceil(waves, cus * wpc);
            cost += double(rounds) * (wpc / 4.0) * (zpl * 50.0 + 12.0);
        }
        if (cost < best_cost * 0.999) {
            best_cost = cost;
            best = {zpl, groups, per_launch};
        }
    }
    return best;
}

namespace {

using PssfpLauncher = void (*)(kernels::PssfpArgs, int, cfloat*, cfloat*, dim3, hipStream_t);

template<int ZPL>
void launch_pssfp(kernels::PssfpArgs a, int group0, cfloat* out0, cfloat* rest, dim3 grid, hipStream_t s) {
    kernels::pssfp_iso<ZPL><<<grid, 64, 0, s>>>(a, group0, out0, rest);
}

PssfpLauncher pssfp_launcher(int zpl) {
    switch (zpl) {
        case 1: return &launch_pssfp<1>;
        case 2: return &launch_pssfp<2>;
        case 3: return &launch_pssfp<3>;
        case 4: return &launch_pssfp<4>;
        case 5: return &launch_pssfp<5>;
        case 6: return &launch_pssfp<6>;
        default: return &launch_pssfp<7>;  // 8 per lane exceeds 256 VGPRs
    }
}

template<int ZPL>
int resident_waves() {
    static const int cached = [] {
        int n = 0;
        COMPAS_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernels::pssfp_iso<ZPL>, 64, 0));
        return std::max(1, n);
    }();
    return cached;
}

}  // namespace

int pssfp_table_zpl(int nz) {
    // the fewest groups of at most 7 positions, balanced: 35 -> 7 x 5, 10 -> 5 x 2
    const int groups = std::max(1, (nz + 6) / 7);
    return std::max(1, (nz + groups - 1) / groups);
}

template<int ZPL>
static void launch_table(hipStream_t s, const kernels::PssfpTableJob* jobs, const int4* items, int n) {
    kernels::pssfp_iso_table<ZPL><<<unsigned(n), 64, 0, s>>>(jobs, items);
}

void launch_pssfp_table(
    const CompasContext& ctx, int zpl, const kernels::PssfpTableJob* jobs, const int4* items, int n) {
    if (n <= 0) return;
    KernelScope ks(ctx, "pssfp_iso_table");
    switch (zpl) {
        case 1: launch_table<1>(ctx.stream(), jobs, items, n); break;
        case 2: launch_table<2>(ctx.stream(), jobs, items, n); break;
        case 3: launch_table<3>(ctx.stream(), jobs, items, n); break;
        case 4: launch_table<4>(ctx.stream(), jobs, items, n); break;
        case 5: launch_table<5>(ctx.stream(), jobs, items, n); break;
        case 6: launch_table<6>(ctx.stream(), jobs, items, n); break;
        case 7: launch_table<7>(ctx.stream(), jobs, items, n); break;
        default: COMPAS_ERROR("slice positions per lane outside 1..7");
    }
    check_launch("pssfp_iso_table");
}

static PssfpLayout primal_layout(const CompasContext& ctx, int nz, int nv, int nTR) {
    const int waves[8] = {0,
                          resident_waves<1>(),
                          resident_waves<2>(),
                          resident_waves<3>(),
                          resident_waves<4>(),
                          resident_waves<5>(),
                          resident_waves<6>(),
                          resident_waves<7>()};
    return pssfp_layout(ctx, nz, nv, nTR, waves, 7);
}

static int partial_slots(const PssfpLayout& L) {
    return L.groups > L.per_launch ? L.per_launch : L.groups - 1;
}

int64_t pssfp_scratch_bytes(const CompasContext& ctx, int nz, int nv, int nTR) {
    if (nz <= 0 || nv <= 0 || nTR <= 0) return 0;
    return int64_t(partial_slots(primal_layout(ctx, nz, nv, nTR))) * nTR * nv * int64_t(sizeof(cfloat));
}

Array<cfloat, 2> simulate_magnetization(
    const CompasContext& ctx,
    TissueParameters parameters,
    pSSFPSequence seq) {
    const int nTR = int(seq.RF_train.size(0));
    const int nv = parameters.nvoxels;
    const int nz = int(seq.z.size(0));
    COMPAS_CHECK(nTR > 0);  // pssfp.h:176
    ctx.activate();
    auto echos = ctx.empty<cfloat, 2>({nTR, nv});
    if (nv == 0) return echos;
    if (nz == 0) {
        ctx.fill_zero(echos.data(), echos.nbytes());
        return echos;
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
    a.echos = echos.data();

    const PssfpLayout L = primal_layout(ctx, nz, nv, nTR);
    const int64_t plane = int64_t(nTR) * nv;
    const int slots = partial_slots(L);
    Array<cfloat, 2> part;
    if (slots > 0) part = ctx.empty<cfloat, 2>({slots, plane});
    const PssfpLauncher fn = pssfp_launcher(L.zpl);
    const unsigned vblocks = unsigned(dev::div_ceil(nv, 64));
    for (int g0 = 0; g0 < L.groups; g0 += L.per_launch) {
        const int count = std::min(L.per_launch, L.groups - g0);
        cfloat* out0 = g0 == 0 ? echos.data() : part.data();
        cfloat* rest = g0 == 0 ? part.data() : part.data() + plane;
        {
            KernelScope ks(ctx, "pssfp_iso");
            fn(a, g0, out0, rest, dim3(vblocks, unsigned(count)), ctx.stream());
            check_launch("pssfp_iso");
        }
        pssfp_sum_groups(ctx, echos.data(), part.data(), g0 == 0 ? count - 1 : count, plane, plane);
    }
    return echos;
}

}  // namespace compas
