// Spiral sums out[c][r][s] = sum_v C[c][v] (A[r][v] + (s - ns/2) Ab[r][v]) E_rv^s
// (signal: A only; J.v: both amplitudes) with the voxel sum on MFMA.
//
// Per readout this is a complex GEMM, rows (coil, amplitude) x K = voxels x
// N = samples, whose B operand E_rv^s is generated: E^s is a geometric
// sequence in s for each voxel, so a lane that keeps its voxels fixed and walks
// the samples needs one complex multiply per element.  Hence the loop order:
// a wave owns one readout, a range of voxels and a block of up to 512 samples,
// and keeps the accumulators of all 32 sixteen-sample tiles (4 floats per
// lane each) while it walks its voxels 16 at a time (one K step of
// v_mfma_f32_16x16x32_bf16).  Per K step each lane evaluates, for its four
// voxels, E^(s0 + n) from an FP64 phase and E^16 (the tile step) once, then
// per tile: four complex multiplies, the BF16 hi + lo split of four pairs and
// three MFMAs (hi*hi, hi*lo, lo*hi; one with the single-product precision).
//
// Rows of the 16-row A operand (lane group g = l / 16 receives result rows
// 4 g .. 4 g + 3 of its column):
//   J.v    row 4 c + (Re Wa, Im Wa, Re Wb, Im Wb), four coils per group; the
//          two amplitude sums are stored as separate planes and the
//          reduction forms Wa-sum + (s - ns/2) Wb-sum;
//   signal row 2 c + (Re W, Im W), eight coils per group.
// W = C A (or C Ab) as (w_re, -w_im) / (w_im, w_re) pairs along the
// interleaved K.  Voxel ranges of one readout are summed by a second pass
// over the partial planes in FP64 (deterministic).
//
// Layout of v_mfma_f32_16x16x32_bf16 (wave64): lane l supplies row (A) or
// column (B) l % 16 and K elements 8 (l / 16) .. + 7; result element j of lane
// l is row 4 (l / 16) + j, column l % 16.

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "compas/device/common.h"
#include "compas/device/mfma_tile.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

typedef float f32x4s __attribute__((ext_vector_type(4)));


__device__ inline mfma::bf16x8 pack_bf16x8(const uint32_t (&w)[4]) {
    const uint4 v = {w[0], w[1], w[2], w[3]};
    return __builtin_bit_cast(mfma::bf16x8, v);
}

// partial[split][c][r][s] over voxels [v_begin, v_end) of split blockIdx.y;
// grid (nr, splits, sample passes), block 64.  FULL: every pass holds
// SSM_TILES whole tiles (ns a multiple of 16 SSM_TILES), so the tile loop
// carries no test.
template<bool TWO, bool LO, int SSM_TILES, bool FULL>
__global__ void __launch_bounds__(64) spiral_sum_mfma(
    dev::TissueView tissue, const cfloat* __restrict__ A, const cfloat* __restrict__ Ab,
    const cfloat* __restrict__ coils, int nc, const cfloat* __restrict__ delta_k, int nr, int ns, float dt,
    int v_chunk, cfloat* __restrict__ partial) {
    constexpr int P = LO ? mfma::P_BF16X3 : mfma::P_BF16X1;
    const int lane = threadIdx.x;
    const int n = lane & 15, g = lane >> 4;
    const int r = blockIdx.x;
    const int nv = tissue.nvoxels;
    const int v_begin = blockIdx.y * v_chunk, v_end = min(nv, v_begin + v_chunk);
    const int sb = blockIdx.z * 16 * SSM_TILES;  // first sample of the pass
    const int ntiles = min(SSM_TILES, (ns - sb + 15) / 16);
    const size_t row = size_t(r) * nv;
    const cfloat dk = delta_k[r];
    const double two_pi_dt = double(dt) * dev::TWO_PI;

    // A-operand row of this lane
    const int arow = n;
    const int acoil = TWO ? (arow >> 2) : (arow >> 1);
    const bool a_im = TWO ? (arow & 1) : (arow & 1);
    const bool a_b = TWO && (arow & 2);  // row of the second amplitude
    const bool a_ok = acoil < nc;

    f32x4s acc[SSM_TILES];
#pragma unroll
    for (int t = 0; t < SSM_TILES; t++) acc[t] = f32x4s {0.0f, 0.0f, 0.0f, 0.0f};

    const int s_lane = sb + n;  // the lane's sample in tile 0
    for (int vk = v_begin; vk < v_end; vk += 16) {
        // the lane's four voxels of this K step (K elements 8 g .. 8 g + 7).
        // Their phase slope, decay and 16-sample step do not depend on the
        // lane's sample: lane n of the group evaluates them for voxel n % 4,
        // and the group's lanes read them from there
        double theta_own;
        float decay_own;
        cfloat st16_own;
        {
            const int v = vk + 4 * g + (n & 3);
            const int vv = v < v_end ? v : v_begin;
            const float R2 = 1.0f / tissue.T2(vv);
            theta_own = double(dk.re) * double(tissue.x(vv)) + double(dk.im) * double(tissue.y(vv)) +
                two_pi_dt * double(tissue.B0(vv));
            decay_own = -dt * R2;
            st16_own = dev::polar_reduced(expf(16.0f * decay_own), 16.0 * theta_own);
        }
        cfloat ph[4], st16[4];
        uint32_t ahw[4], alw[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int v = vk + 4 * g + j;
            const bool ok = v < v_end;
            const int src = (lane & ~15) | j;
            const double theta = __shfl(theta_own, src);
            const float decay = __shfl(decay_own, src);
            st16[j] = {__shfl(st16_own.re, src), __shfl(st16_own.im, src)};
            ph[j] = ok ? dev::polar_reduced(expf(float(s_lane) * decay), double(s_lane) * theta) : cfloat {0, 0};
            // A: W = C[coil][v] (A or Ab)[r][v] as a (re, -im) or (im, re) pair
            cfloat w = {0.0f, 0.0f};
            if (a_ok && ok) w = dev::cmul(coils[size_t(acoil) * nv + v], a_b ? Ab[row + v] : A[row + v]);
            const float e0 = a_im ? w.im : w.re;
            const float e1 = a_im ? w.re : -w.im;
            mfma::split_pair<P>(e0, e1, ahw[j], alw[j]);
        }
        const mfma::bf16x8 ahi = pack_bf16x8(ahw);
        const mfma::bf16x8 alo = pack_bf16x8(alw);
#pragma unroll
        for (int t = 0; t < SSM_TILES; t++) {
            if (FULL || t < ntiles) {
                uint32_t bhw[4], blw[4];
#pragma unroll
                for (int j = 0; j < 4; j++) mfma::split_pair<P>(ph[j].re, ph[j].im, bhw[j], blw[j]);
                const mfma::bf16x8 bhi = pack_bf16x8(bhw);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, bhi, acc[t], 0, 0, 0);
                if constexpr (LO) {
                    const mfma::bf16x8 blo = pack_bf16x8(blw);
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, blo, acc[t], 0, 0, 0);
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, bhi, acc[t], 0, 0, 0);
                }
#pragma unroll
                for (int j = 0; j < 4; j++) ph[j] = dev::cmul(ph[j], st16[j]);
            }
        }
    }

    // result rows 4 g + j of column n (sample sb + 16 t + n); with two
    // amplitudes the second one goes to its own plane (the reduction applies
    // the (s - ns/2) factor), so the accumulators are only stored here
    const size_t plane = size_t(nc) * nr * ns;
    cfloat* base = partial + size_t(blockIdx.y) * (TWO ? 2 : 1) * plane;
#pragma unroll
    for (int t = 0; t < SSM_TILES; t++) {
        const int s = sb + 16 * t + n;
        if (FULL || (t < ntiles && s < ns)) {
            if constexpr (TWO) {
                const int c = g;
                if (c < nc) {
                    const size_t o = (size_t(c) * nr + r) * ns + s;
                    base[o] = {acc[t][0], acc[t][1]};
                    base[plane + o] = {acc[t][2], acc[t][3]};
                }
            } else {
                const int c0 = 2 * g;
                if (c0 < nc) base[(size_t(c0) * nr + r) * ns + s] = {acc[t][0], acc[t][1]};
                if (c0 + 1 < nc) base[(size_t(c0 + 1) * nr + r) * ns + s] = {acc[t][2], acc[t][3]};
            }
        }
    }
}

// out[i] = sum over the splits (FP64, in order); with two amplitudes each
// split holds the plane of the first and then of the second, and
// out = first + (s - ns/2) second, s = i % ns
template<bool TWO>
__global__ void spiral_sum_reduce(
    const cfloat* __restrict__ partial, int splits, int64_t n, int ns, cfloat* __restrict__ out) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        double re = 0.0, im = 0.0, bre = 0.0, bim = 0.0;
        for (int k = 0; k < splits; k++) {
            const size_t o = size_t(k) * (TWO ? 2 : 1) * size_t(n) + size_t(i);
            const cfloat p = partial[o];
            re += p.re;
            im += p.im;
            if constexpr (TWO) {
                const cfloat q = partial[o + size_t(n)];
                bre += q.re;
                bim += q.im;
            }
        }
        if constexpr (TWO) {
            const double sp = double(i % ns) - 0.5 * double(ns);
            re = fma(sp, bre, re);
            im = fma(sp, bim, im);
        }
        out[i] = {float(re), float(im)};
    }
}

}  // namespace kernels

bool spiral_sum_mfma(
    const CompasContext& ctx, const TissueParameters& p, const cfloat* A, const cfloat* Ab, const cfloat* coils,
    int nc, const SpiralTrajectory& t, ContractionPrecision precision, cfloat* out) {
    if (precision == ContractionPrecision::FP32) return false;
    const bool two = Ab != nullptr;
    if (nc > (two ? 4 : 8)) return false;  // one 16-row group; more coils: the VALU form
    const bool lo = precision == ContractionPrecision::BF16x3 || precision == ContractionPrecision::FP16x3;
    const int nr = t.nreadouts, ns = t.samples_per_readout, nv = p.nvoxels;
    // sample block per pass: 32 tiles of 16 samples (128 accumulator
    // registers, two waves per SIMD).  Measured at 4 coils x 1120 x 512 over
    // 50176 voxels: 16.5 ms per launch, against 18.7 ms with 16 tiles (four
    // waves per SIMD but twice the per-voxel set-up per sample).
    constexpr int tiles = 32;
    const int passes = int(dev::div_ceil(ns, 16 * tiles));
    // voxel splits so that the grid holds about four waves per SIMD
    const int want = int(std::max<int64_t>(1, dev::div_ceil(4 * 1024, int64_t(nr) * passes)));
    const int max_splits = int(dev::div_ceil(nv, 16));
    const int splits0 = std::min(want, max_splits);
    const int v_chunk = int(dev::div_ceil(dev::div_ceil(nv, splits0), 16)) * 16;
    const int splits = int(dev::div_ceil(nv, v_chunk));
    const int64_t plane = int64_t(nc) * nr * ns;
    auto partial = ctx.empty<cfloat, 1>({int64_t(splits) * (two ? 2 : 1) * plane});
    {
        KernelScope ks(ctx, "spiral_sum_mfma");
        const dim3 grid(nr, splits, passes);
        const auto tv = dev::tissue_view(p);
        hipStream_t s = ctx.stream();
        const bool full = ns % (16 * tiles) == 0;
        auto launch = [&](auto two_c, auto lo_c) {
            constexpr bool TWO = decltype(two_c)::value;
            constexpr bool LO = decltype(lo_c)::value;
            if (full) {
                kernels::spiral_sum_mfma<TWO, LO, tiles, true><<<grid, 64, 0, s>>>(
                    tv, A, TWO ? Ab : nullptr, coils, nc, t.delta_k.data(), nr, ns, t.delta_t, v_chunk, partial.data());
            } else {
                kernels::spiral_sum_mfma<TWO, LO, tiles, false><<<grid, 64, 0, s>>>(
                    tv, A, TWO ? Ab : nullptr, coils, nc, t.delta_k.data(), nr, ns, t.delta_t, v_chunk, partial.data());
            }
        };
        if (two) {
            if (lo) launch(std::true_type {}, std::true_type {});
            else launch(std::true_type {}, std::false_type {});
        } else {
            if (lo) launch(std::false_type {}, std::true_type {});
            else launch(std::false_type {}, std::false_type {});
        }
        check_launch("spiral_sum_mfma");
    }
    {
        KernelScope ks(ctx, "spiral_sum_reduce");
        const int64_t blocks = std::min<int64_t>(dev::div_ceil(plane, int64_t(256)), 8192);
        if (two) kernels::spiral_sum_reduce<true><<<unsigned(blocks), 256, 0, ctx.stream()>>>(partial.data(), splits, plane, ns, out);
        else kernels::spiral_sum_reduce<false><<<unsigned(blocks), 256, 0, ctx.stream()>>>(partial.data(), splits, plane, ns, out);
        check_launch("spiral_sum_reduce");
    }
    return true;
}

}  // namespace compas
