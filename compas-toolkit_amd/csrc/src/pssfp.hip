// pSSFP isochromat (Bloch) simulation on gfx950.
//
// Semantics: reference src/simulate/pssfp_kernels.cuh:14-129, pssfp.cu:7-91
// and src/operators/isochromat.cuh:9-128.
//
// Design (MI355X):
//   * A workgroup covers 64 voxels (one per lane) and G = ceil(nz / ZPL)
//     waves; wave g simulates slice positions [g*ZPL, g*ZPL + ZPL) of its
//     voxel, ZPL independent isochromats per lane (instruction-level
//     parallelism).  The sum over z at each echo goes through LDS, one
//     barrier per TR (double-buffered partials).
//   * The isochromat state is FP64.  On MI355X, FP64 FMA issues at the same
//     rate as non-packed FP32, and the FP32 form of this recursion drifts
//     ~3e-5 (relative L2) from the exact result over 1000+ TRs because each TR
//     applies nRF Rodrigues rotations with little T2 damping.
//   * Rodrigues' rotation is evaluated without sqrt/rcp/sin/cos for the small
//     per-step angles: with u = |a|^2, s1(u) = sin(t)/t and s2(u) =
//     (1 - cos t)/t^2 are Taylor polynomials in u (|t| <= 1/4: 6 terms, error
//     < 1e-17), cos t = 1 - u s2(u), and
//       m' = cos t m + s1 (a x m) + s2 (a . m) a,
//     which equals the reference's normalised form.  Larger angles fall back
//     to FP64 sincos.
//   * The RF sub-pulse table is staged in LDS as doubles (broadcast reads in
//     the sub-step loop instead of a global load per sub-step).

#include "compas/device/common.h"
#include "compas/device/pssfp_common.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

template<int ZPL>
__global__ void __launch_bounds__(512) pssfp_iso(PssfpArgs a) {
    __shared__ double2 partial[2][8][64];
    __shared__ double2 grf[RF_CHUNK];  // gamma_dt_RF as doubles (LDS broadcast reads)
    const int lane = threadIdx.x & 63;
    const int group = threadIdx.x >> 6;
    const int groups = blockDim.x >> 6;
    const int voxel = blockIdx.x * 64 + lane;
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
    bool valid[ZPL];
#pragma unroll
    for (int i = 0; i < ZPL; i++) {
        const int iz = group * ZPL + i;
        valid[i] = iz < a.nz;
        const double zz = valid[i] ? double(a.z[iz]) : 0.0;
        az[i] = -(double(a.g_ex) * zz) - off_ex;
        sincos(-(double(a.g_pr) * zz) - off_pr, &spr[i], &cpr[i]);
        // m = (0,0,1); invert(p): m_z *= cos(pi B1); decay + regrowth over dt_inv
        m[i] = {0.0, 0.0, cos(M_PI * B1)};
        relax3(m[i], E1_inv, E2_inv, R1_inv);
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

    // every thread of the block runs the same excitations, so the barriers of a
    // chunked RF table are uniform
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
#pragma unroll
                for (int i = 0; i < ZPL; i++) {
                    rotate_small(m[i], ax, ay, az[i], axy2);
                    relax3(m[i], E1_ex, E2_ex, R1_ex);
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
        partial[n & 1][group][lane] = make_double2(sx, sy);
        __syncthreads();
        if (group == 0 && active) {
            double tx = 0.0, ty = 0.0;
            for (int g = 0; g < groups; g++) {
                const double2 p = partial[n & 1][g][lane];
                tx += p.x;
                ty += p.y;
            }
            a.echos[size_t(n) * size_t(a.tissue.nvoxels) + voxel] = {float(tx), float(ty)};
        }
#pragma unroll
        for (int i = 0; i < ZPL; i++) {
            rotate_z(m[i], cpr[i], spr[i]);
            relax3(m[i], E1_pr, E2_pr, R1_pr);
        }
    }
}

}  // namespace kernels

using PssfpLauncher = void (*)(kernels::PssfpArgs, unsigned, unsigned, hipStream_t);

template<int ZPL>
static void launch_pssfp(kernels::PssfpArgs a, unsigned blocks, unsigned threads, hipStream_t s) {
    kernels::pssfp_iso<ZPL><<<blocks, threads, 0, s>>>(a);
}

static PssfpLauncher pssfp_launcher(int zpl) {
    switch (zpl) {
        case 1: return &launch_pssfp<1>;
        case 2: return &launch_pssfp<2>;
        case 3: return &launch_pssfp<3>;
        case 4: return &launch_pssfp<4>;
        case 5: return &launch_pssfp<5>;
        case 6: return &launch_pssfp<6>;
        case 7: return &launch_pssfp<7>;
        default: return nullptr;
    }
}

// Slice positions per lane: the smallest ZPL in 1..8 with at most 8 waves
// per workgroup and the least padding, preferring 4..8 isochromats per lane.
static int choose_zpl(int nz) {
    int best = 0, best_pad = 1 << 30;
    for (int zpl = 7; zpl >= 1; zpl--) {  // 8 per lane exceeds 256 VGPRs
        int g = (nz + zpl - 1) / zpl;
        if (g > 8) continue;
        int pad = g * zpl - nz;
        if (zpl < 4 && nz >= 4) pad += 1;  // fewer isochromats per lane cost per-lane overhead
        if (pad < best_pad) {
            best_pad = pad;
            best = zpl;
        }
    }
    return best;
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
    a.nTR = nTR;
    a.nRF = int(seq.gamma_dt_RF.size(0));
    a.dt_ex = seq.dt.ex;
    a.dt_inv = seq.dt.inv;
    a.dt_pr = seq.dt.pr;
    a.g_ex = seq.gamma_dt_GRz.ex;
    a.g_pr = seq.gamma_dt_GRz.pr;
    a.echos = echos.data();

    // more than 56 slice positions: passes of <= 56 accumulate into a scratch
    // sum (rare; the reference supports any nz by batching launches)
    const int chunk = 56;  // 8 waves x 7 isochromats per lane
    Array<cfloat, 2> part;
    for (int z0 = 0; z0 < nz; z0 += chunk) {
        const int nzc = std::min(chunk, nz - z0);
        a.z = seq.z.data() + z0;
        a.nz = nzc;
        const int zpl = choose_zpl(nzc);
        const int groups = (nzc + zpl - 1) / zpl;
        if (z0 > 0) {
            if (part.is_empty()) part = ctx.empty<cfloat, 2>({nTR, nv});
            a.echos = part.data();
        }
        {
            KernelScope ks(ctx, "pssfp_iso");
            pssfp_launcher(zpl)(a, dev::div_ceil(nv, 64), unsigned(64 * groups), ctx.stream());
            check_launch("pssfp_iso");
        }
        if (z0 > 0) add_into(ctx, echos.data(), part.data(), int64_t(nTR) * nv);
    }
    return echos;
}

}  // namespace compas
