// J^H.y on MFMA with the readout reduction fused into the contraction.
//
// Reference: compute_jacobian_hermitian (src/jacobian/hermitian.cu:10-339,
// hermitian_kernels.cuh:30-92 and :314-360).  Per voxel v and coil c
//   G [c][r][v] = sum_s y[c][r][s] conj(E[s][v])
//   G'[c][r][v] = sum_s y[c][r][s] conj(dE/dT2[s][v])
//   u_c  = sum_r G  conj(m)      d1_c = sum_r G conj(dm/dT1)
//   d2_c = sum_r G  conj(dm/dT2) + G' conj(m)
// and JHy = (conj(rho T1 C) d1, conj(rho T2 C) d2, conj(C) u, -i conj(C) u)
// summed over coils.
//
// The reference materialises G per coil and re-reads m, dm/dT1, dm/dT2 once
// per coil in its finalize pass.  Here one kernel computes a tile of G and G'
// on MFMA (rows = readout x coil, columns = voxel x {E, dE}) and reduces it
// over the tile's readouts in the epilogue, so m and both derivatives are read
// from HBM once for all coils and G never leaves the registers.  Rows are
// packed readout-major with the coils of a group innermost (row = r NCP + c),
// which puts all coils of a readout in the same lane (accumulator element
// e % 4), so each magnetisation value loaded serves NCP coils.  Coils run in
// groups of NCP = 4 (one or two coils use the per-coil path instead).  Partial sums
// per (readout tile, coil, voxel) are reduced in FP64 by a small kernel that
// also applies the coil, proton-density and T1/T2 factors.

#include "compas/device/common.h"
#include "compas/device/mfma_tile.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

using namespace mfma;

// (re, im) of x conj(z) contributed by the real part (part 0) or the
// imaginary part (part 1) of x: summing both lanes' contributions gives x conj(z)
__device__ inline float2 conj_factor(cfloat z, int part) {
    return part ? make_float2(z.im, z.re) : make_float2(z.re, -z.im);
}

// Per-lane readout sums of the epilogue (lane = 2 g + part for voxel g of the
// wave's 16, coil c): G conj(m), G conj(dm/dT1), G conj(dm/dT2), G' conj(m).
// Record (wave (w, ow), part, g) starts at red_index; records are padded to
// 8 NCP + 4 floats and the part planes offset by 16 more, so that the lanes'
// 16-byte stores and the voxel threads' loads spread over the banks.
template<int NCP>
struct RedLayout {
    static constexpr int RS = 8 * NCP + 4;  // floats per record
    static constexpr int PLANE = 16 * RS + 16;  // floats per (wave, part) plane
};
template<int WN, int NCP>
__device__ inline int red_index(int w, int ow, int g, int part) {
    return ((w * WN + ow) * 2 + part) * RedLayout<NCP>::PLANE + g * RedLayout<NCP>::RS;
}

template<int P, int WM, int WN, int BKC, int NCP>
union HermitianSmem {
    MfmaSmem<P, WM, WN, BKC> tile;
    __attribute__((aligned(16))) float red[WM * WN * 2 * RedLayout<NCP>::PLANE];
};

// Readout tiles per group of the launch order: a group sweeps every voxel tile
// for GROUP_R readout tiles, so the group's y rows stay in L2 while the
// voxel tiles stream past (and each voxel tile is re-read once per group).
constexpr int GROUP_R = 8;

template<int P, int WM, int WN, int BKC, int NCP>
__global__ void __launch_bounds__(64 * WM * WN, 4) hermitian_mfma(
    const cfloat* __restrict__ y,  // [nc][nr][ns]
    const cfloat* __restrict__ EH,  // [2 nv][ns]: conj(E) rows, then conj(dE/dT2) rows
    const cfloat* __restrict__ m,  // [nr][nv]
    const cfloat* __restrict__ d1,
    const cfloat* __restrict__ d2,
    int c0, int ncg, int nr, int nv, int ns, int tiles_r, int tiles_n,
    const float* __restrict__ scale_y,  // FP16: per y row (c nr + r), else nullptr
    const float* __restrict__ scale_b,  // FP16: per EH row
    float2* __restrict__ partial) {  // [tiles_r][NCP][3][nv]: u, d1, d2
    using Tl = MfmaTile<P, WM, WN, BKC>;
    static_assert(NCP == 1 || NCP == 2 || NCP == 4, "coils per group");
    __shared__ HermitianSmem<P, WM, WN, BKC, NCP> sm;

    const int total = tiles_r * tiles_n;
    const int lid = xcd_logical_id(blockIdx.x, total);
    if (lid >= total) return;
    const int group = lid / (GROUP_R * tiles_n), rem = lid % (GROUP_R * tiles_n);
    const int tr0 = group * GROUP_R, gr = min(GROUP_R, tiles_r - tr0);
    const int tr = tr0 + rem % gr, tn = rem / gr;
    // rows: (readout, coil) with the coil innermost; complex columns of wave wn:
    // 16 voxels with conj(E) (block j = 0) and the same 16 with conj(dE) (j = 1)
    const int m0 = tr * Tl::BM, vox0 = tn * (Tl::BNC / 2);

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WN, wn = wave % WN;

    f32x16 acc[2][2];
    zero(acc);
    {
        const int am = m0 + staged_a_row<P, WM, WN, BKC>(tid);
        const int ar = am / NCP, ac = am % NCP;
        const bool a_ok = ar < nr && ac < ncg;
        const int64_t yrow = int64_t(c0 + ac) * nr + ar;
        const int nl = staged_b_row<P, WM, WN, BKC>(tid);  // complex column within the tile
        const int bv = vox0 + (nl / 32) * 16 + nl % 16, bvar = (nl / 16) % 2;
        const bool b_ok = bv < nv;
        const int64_t brow = int64_t(bvar) * nv + bv;
        const float sa = (scale_y && a_ok) ? scale_y[yrow] : 1.0f;
        const float sb = (scale_b && b_ok) ? scale_b[brow] : 1.0f;
        const bool vec_a = ((reinterpret_cast<uintptr_t>(y) & 15) == 0) && (ns % 2 == 0);
        const bool vec_b = ((reinterpret_cast<uintptr_t>(EH) & 15) == 0) && (ns % 2 == 0);
        mfma_accumulate<P, WM, WN, BKC>(acc, sm.tile, a_ok ? y + yrow * ns : nullptr, sa, vec_a,
                                        b_ok ? EH + brow * ns : nullptr, sb, vec_b, 0, ns);
    }
    // mfma_accumulate ends with a barrier: the tile buffers are free for the sums

    // ---- epilogue: the lane holds G (block j = 0) and G' (j = 1) of one voxel
    // for its readouts; reduce them against m, dm (one load serves both blocks
    // and all NCP coils of the readout)
    const int part = lane & 1;
    const int v = vox0 + wn * 16 + (lane & 31) / 2;
    const bool v_ok = v < nv;
    // readout of (i, q, tt) = r_base + (32 i + 8 q + NCP tt) / NCP
    const int r_base = (m0 + wm * 64 + (lane >> 5) * 4) / NCP;
    const int64_t base = int64_t(r_base) * nv + v;
    float s[NCP][8];
#pragma unroll
    for (int c = 0; c < NCP; c++)
#pragma unroll
        for (int k = 0; k < 8; k++) s[c][k] = 0.0f;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int tt = 0; tt < 4 / NCP; tt++) {
                const int dr = (i * 32 + q * 8 + tt * NCP) / NCP;
                const int r = r_base + dr;
                cfloat mv {0, 0}, a1 {0, 0}, a2 {0, 0};
                if (v_ok && r < nr) {
                    const int64_t idx = base + int64_t(dr) * nv;
                    mv = m[idx];
                    a1 = d1[idx];
                    a2 = d2[idx];
                }
                const float2 fm = conj_factor(mv, part), f1 = conj_factor(a1, part), f2 = conj_factor(a2, part);
#pragma unroll
                for (int c = 0; c < NCP; c++) {
                    const int e = q * 4 + tt * NCP + c;
                    float g = acc[i][0][e], gp = acc[i][1][e];
                    if constexpr (is_fp16(P)) {
                        if (v_ok && r < nr && c < ncg) {
                            const float sy = scale_y[int64_t(c0 + c) * nr + r];
                            g /= sy * scale_b[v];
                            gp /= sy * scale_b[int64_t(nv) + v];
                        }
                    }
                    s[c][0] = fmaf(g, fm.x, s[c][0]);
                    s[c][1] = fmaf(g, fm.y, s[c][1]);
                    s[c][2] = fmaf(g, f1.x, s[c][2]);
                    s[c][3] = fmaf(g, f1.y, s[c][3]);
                    s[c][4] = fmaf(g, f2.x, s[c][4]);
                    s[c][5] = fmaf(g, f2.y, s[c][5]);
                    s[c][6] = fmaf(gp, fm.x, s[c][6]);
                    s[c][7] = fmaf(gp, fm.y, s[c][7]);
                }
            }
    // readouts held by the other half-wave (same voxel, same coil)
#pragma unroll
    for (int c = 0; c < NCP; c++)
#pragma unroll
        for (int k = 0; k < 8; k++) s[c][k] += __shfl_xor(s[c][k], 32);
    if (lane < 32) {
        float* rec = sm.red + red_index<WN, NCP>(wm, wn, lane >> 1, lane & 1);
#pragma unroll
        for (int c = 0; c < NCP; c++)
#pragma unroll
            for (int k = 0; k < 8; k++) rec[8 * c + k] = s[c][k];
    }
    __syncthreads();

    // ---- one thread per (coil, voxel) of the tile: sum the waves along M and
    // the two parts of the voxel; G' conj(m) joins d2
    constexpr int VOX = Tl::BNC / 2;
    for (int o = tid; o < NCP * VOX; o += Tl::THREADS) {
        const int c = o / VOX, vl = o % VOX;
        const int vo = vox0 + vl;
        if (c >= ncg || vo >= nv) continue;
        const int ow = vl / 16, g = vl % 16;
        float t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int w = 0; w < WM; w++)
#pragma unroll
            for (int pt = 0; pt < 2; pt++) {
                const float* e = sm.red + red_index<WN, NCP>(w, ow, g, pt) + 8 * c;
#pragma unroll
                for (int k = 0; k < 8; k++) t[k] += e[k];
            }
        float2* out = partial + (int64_t(tr) * NCP + c) * 3 * int64_t(nv) + vo;
        out[0] = make_float2(t[0], t[1]);
        out[nv] = make_float2(t[2], t[3]);
        out[2 * int64_t(nv)] = make_float2(t[4] + t[6], t[5] + t[7]);
    }
}

// JHy[p][v] (+)= sum over the coils of the group of the FP64-reduced partials
// times the adjoint factors (hermitian_kernels.cuh:314-360)
__global__ void hermitian_combine(
    dev::TissueView t, const float2* __restrict__ partial, int tiles_r, int ncp, int ncg,
    const cfloat* __restrict__ coils, bool first, cfloat* __restrict__ JHy) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int nv = t.nvoxels;
    if (v >= nv) return;
    const cfloat rho = t.rho(v);
    const cfloat rt1 = rho * t.T1(v), rt2 = rho * t.T2(v);
    cfloat o0 {0, 0}, o1 {0, 0}, o2 {0, 0};
    for (int c = 0; c < ncg; c++) {
        double s[6] = {0, 0, 0, 0, 0, 0};
        for (int tr = 0; tr < tiles_r; tr++) {
            const float2* p = partial + (int64_t(tr) * ncp + c) * 3 * int64_t(nv) + v;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float2 x = p[k * int64_t(nv)];
                s[2 * k] += double(x.x);
                s[2 * k + 1] += double(x.y);
            }
        }
        const cfloat u {float(s[0]), float(s[1])}, g1 {float(s[2]), float(s[3])}, g2 {float(s[4]), float(s[5])};
        const cfloat C = coils[int64_t(c) * nv + v];
        o0 = o0 + dev::cmul(conj(dev::cmul(rt1, C)), g1);
        o1 = o1 + dev::cmul(conj(dev::cmul(rt2, C)), g2);
        o2 = o2 + dev::cmul(conj(C), u);
    }
    const cfloat o3 = dev::cmul(o2, cfloat {0.0f, -1.0f});
    if (first) {
        JHy[v] = o0;
        JHy[int64_t(nv) + v] = o1;
        JHy[2 * int64_t(nv) + v] = o2;
        JHy[3 * int64_t(nv) + v] = o3;
    } else {
        JHy[v] += o0;
        JHy[int64_t(nv) + v] += o1;
        JHy[2 * int64_t(nv) + v] += o2;
        JHy[3 * int64_t(nv) + v] += o3;
    }
}

}  // namespace kernels

namespace {

using namespace kernels;

constexpr int H_WM = 2, H_WN = 4, H_BKC = 16;

template<int P, int NCP>
void launch_hermitian(
    const CompasContext& ctx, const dev::TissueView& tv, const cfloat* y, const cfloat* EH, const cfloat* m,
    const cfloat* d1, const cfloat* d2, const cfloat* coils, int c0, int ncg, int nr, int ns,
    const float* scale_y, const float* scale_b, bool first, cfloat* JHy) {
    using Tl = MfmaTile<P, H_WM, H_WN, H_BKC>;
    const int nv = tv.nvoxels;
    const int tiles_r = int(dev::div_ceil(int64_t(nr) * NCP, Tl::BM));
    const int tiles_n = int(dev::div_ceil(2 * int64_t(nv), Tl::BNC));
    auto partial = ctx.empty<float2, 1>({int64_t(tiles_r) * NCP * 3 * nv});
    {
        KernelScope ks(ctx, "hermitian_mfma");
        hermitian_mfma<P, H_WM, H_WN, H_BKC, NCP><<<tiles_r * tiles_n, Tl::THREADS, 0, ctx.stream()>>>(
            y, EH, m, d1, d2, c0, ncg, nr, nv, ns, tiles_r, tiles_n, scale_y, scale_b, partial.data());
        check_launch("hermitian_mfma");
    }
    {
        KernelScope ks(ctx, "hermitian_combine");
        hermitian_combine<<<dev::div_ceil(nv, 256), 256, 0, ctx.stream()>>>(
            tv, partial.data(), tiles_r, NCP, ncg, coils + int64_t(c0) * nv, first, JHy);
        check_launch("hermitian_combine");
    }
}

template<int P>
void hermitian_groups(
    const CompasContext& ctx, const dev::TissueView& tv, const cfloat* y, const cfloat* EH, const cfloat* m,
    const cfloat* d1, const cfloat* d2, const cfloat* coils, int nc, int nr, int ns, const float* scale_y,
    const float* scale_b, cfloat* JHy) {
    // groups of four coils; a last group of fewer coils runs with zero rows
    for (int c0 = 0; c0 < nc; c0 += 4) {
        const int ncg = std::min(4, nc - c0);
        launch_hermitian<P, 4>(ctx, tv, y, EH, m, d1, d2, coils, c0, ncg, nr, ns, scale_y, scale_b, c0 == 0, JHy);
    }
}

}  // namespace

bool jacobian_hermitian_mfma(
    const CompasContext& ctx, const TissueParameters& p, const cfloat* y, const cfloat* EH, const cfloat* m,
    const cfloat* d1, const cfloat* d2, const cfloat* coils, int nc, int nr, int ns,
    ContractionPrecision precision, cfloat* JHy) {
    // with one or two coils the padded coil slots would waste half or more of
    // the MFMA work: the caller's per-coil path is cheaper there
    if (nc < 3) return false;
    const auto tv = dev::tissue_view(p);
    const int nv = p.nvoxels;
    Array<float, 1> sy, sb;
    if (precision == ContractionPrecision::FP16x3 || precision == ContractionPrecision::FP16x1) {
        sy = ctx.empty<float, 1>({int64_t(nc) * nr});
        sb = ctx.empty<float, 1>({2 * int64_t(nv)});
        row_pow2_scales(ctx, y, nc * nr, ns, ns, sy.data());
        row_pow2_scales(ctx, EH, 2 * nv, ns, ns, sb.data());
    }
    switch (precision) {
        case ContractionPrecision::BF16x3:
            hermitian_groups<P_BF16X3>(ctx, tv, y, EH, m, d1, d2, coils, nc, nr, ns, nullptr, nullptr, JHy);
            return true;
        case ContractionPrecision::BF16x1:
            hermitian_groups<P_BF16X1>(ctx, tv, y, EH, m, d1, d2, coils, nc, nr, ns, nullptr, nullptr, JHy);
            return true;
        case ContractionPrecision::FP16x3:
            hermitian_groups<P_FP16X3>(ctx, tv, y, EH, m, d1, d2, coils, nc, nr, ns, sy.data(), sb.data(), JHy);
            return true;
        case ContractionPrecision::FP16x1:
            hermitian_groups<P_FP16X1>(ctx, tv, y, EH, m, d1, d2, coils, nc, nr, ns, sy.data(), sb.data(), JHy);
            return true;
        default:
            return false;
    }
}

}  // namespace compas
