// Complex contraction engine: C[m][n] = sum_k A[m][k] * B[n][k].
//
// The reference runs its signal and Jacobian contractions through
// cuBLAS/rocBLAS as four real GEMMs on planar operands
// (src/utils/gemm.cu:147-173).  Here the contraction is one kernel on
// interleaved complex operands:
//   * FP32 path: LDS-tiled VALU kernel, 64x64 complex tile per 256-thread
//     workgroup, 4x4 complex outputs per thread, K-split across workgroups;
//   * split-K partials are summed in FP64 by a reduction kernel that also
//     writes the caller's output layout (e.g. [coil][readout][sample]).

#include <atomic>
#include <cmath>

#include "compas/device/common.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

constexpr int CT_BM = 64, CT_BN = 64, CT_BK = 16;

// out-of-place: partial[split][M][N] (or direct output when direct != nullptr)
__global__ void __launch_bounds__(256) contraction_fp32(
    const cfloat* __restrict__ A,
    const cfloat* __restrict__ B,
    int M, int N, int K, int64_t lda, int64_t ldb,
    int k_chunk,
    float2* __restrict__ partial,
    cfloat* __restrict__ direct, int nsub, int64_t out_outer) {
    __shared__ float2 As[CT_BK][CT_BM + 1];
    __shared__ float2 Bs[CT_BK][CT_BN + 1];

    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int m0 = blockIdx.y * CT_BM, n0 = blockIdx.x * CT_BN;
    const int kb = blockIdx.z * k_chunk;
    const int ke = min(K, kb + k_chunk);

    float acc_re[4][4], acc_im[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            acc_re[i][j] = 0.0f;
            acc_im[i][j] = 0.0f;
        }

    const int lrow = tid >> 2;  // 0..63
    const int lk = (tid & 3) * 4;  // 0,4,8,12

    for (int k0 = kb; k0 < ke; k0 += CT_BK) {
        // stage A and B tiles (zero outside the matrix)
        {
            const int m = m0 + lrow;
            const cfloat* src = A + int64_t(m) * lda;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int k = k0 + lk + i;
                cfloat v = (m < M && k < ke) ? src[k] : cfloat {0, 0};
                As[lk + i][lrow] = make_float2(v.re, v.im);
            }
            const int n = n0 + lrow;
            const cfloat* srcb = B + int64_t(n) * ldb;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int k = k0 + lk + i;
                cfloat v = (n < N && k < ke) ? srcb[k] : cfloat {0, 0};
                Bs[lk + i][lrow] = make_float2(v.re, v.im);
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < CT_BK; kk++) {
            float2 a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; j++) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    acc_re[i][j] = fmaf(a[i].x, b[j].x, fmaf(-a[i].y, b[j].y, acc_re[i][j]));
                    acc_im[i][j] = fmaf(a[i].x, b[j].y, fmaf(a[i].y, b[j].x, acc_im[i][j]));
                }
        }
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int m = m0 + ty * 4 + i;
        if (m >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int n = n0 + tx * 4 + j;
            if (n >= N) continue;
            if (direct) {
                const int64_t o = int64_t(n / nsub) * out_outer + int64_t(m) * nsub + (n % nsub);
                direct[o] = {acc_re[i][j], acc_im[i][j]};
            } else {
                partial[(int64_t(blockIdx.z) * M + m) * N + n] = make_float2(acc_re[i][j], acc_im[i][j]);
            }
        }
    }
}

__global__ void contraction_reduce(
    const float2* __restrict__ partial, int splits, int M, int N,
    cfloat* __restrict__ out, int nsub, int64_t out_outer, bool accumulate) {
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= int64_t(M) * N) return;
    const int m = int(idx / N), n = int(idx % N);
    double re = 0.0, im = 0.0;
    for (int s = 0; s < splits; s++) {
        const float2 p = partial[int64_t(s) * M * N + idx];
        re += double(p.x);
        im += double(p.y);
    }
    const int64_t o = int64_t(n / nsub) * out_outer + int64_t(m) * nsub + (n % nsub);
    if (accumulate) {
        re += double(out[o].re);
        im += double(out[o].im);
    }
    out[o] = {float(re), float(im)};
}

}  // namespace kernels

ContractionPrecision signal_precision(SimulateSignalMethod method) {
    switch (method) {
        case SimulateSignalMethod::Naive:
        case SimulateSignalMethod::Direct:
            return ContractionPrecision::FP32;
        case SimulateSignalMethod::MatmulPedantic:
            return ContractionPrecision::FP16x3;
        case SimulateSignalMethod::Matmul:
        case SimulateSignalMethod::MatmulFast:
            return ContractionPrecision::BF16x3;
        case SimulateSignalMethod::MatmulLow:
            return ContractionPrecision::BF16x1;
    }
    COMPAS_ERROR("invalid value for `SimulateSignalMethod`");
}

ContractionPrecision jacobian_precision(JacobianComputeMethod method) {
    switch (method) {
        case JacobianComputeMethod::Naive:
        case JacobianComputeMethod::Direct:
        case JacobianComputeMethod::Gemm:
            return ContractionPrecision::FP32;
        case JacobianComputeMethod::GemmFast:
            return ContractionPrecision::BF16x3;
        case JacobianComputeMethod::GemmLow:
            return ContractionPrecision::BF16x1;
    }
    COMPAS_ERROR("invalid value for `JacobianComputeMethod`");
}

void complex_contraction_mfma(
    const CompasContext& ctx,
    const cfloat* A,
    const cfloat* B,
    cfloat* out,
    ContractionShape s,
    ContractionPrecision precision,
    bool* handled);

namespace {
std::atomic<int> g_precision_override {-1};
}

void set_contraction_precision_override(int precision) {
    COMPAS_CHECK(precision >= -1 && precision <= int(ContractionPrecision::FP16x1));
    g_precision_override.store(precision);
}

ContractionPrecision effective_precision(ContractionPrecision requested) {
    const int o = g_precision_override.load();
    return o >= 0 ? ContractionPrecision(o) : requested;
}

void complex_contraction(
    const CompasContext& ctx,
    const cfloat* A,
    const cfloat* B,
    cfloat* out,
    ContractionShape s,
    ContractionPrecision precision) {
    COMPAS_CHECK(s.M >= 0 && s.N >= 0 && s.K >= 0 && s.nsub > 0);
    precision = effective_precision(precision);
    if (s.M == 0 || s.N == 0) return;
    if (precision != ContractionPrecision::FP32) {
        bool handled = false;
        complex_contraction_mfma(ctx, A, B, out, s, precision, &handled);
        if (handled) return;
    }
    const int tiles_m = int(dev::div_ceil(s.M, kernels::CT_BM));
    const int tiles_n = int(dev::div_ceil(s.N, kernels::CT_BN));
    const int tiles = tiles_m * tiles_n;
    int splits = std::max(1, int(std::ceil(768.0 / tiles)));
    splits = std::min(splits, std::max(1, s.K / (8 * kernels::CT_BK)));
    int k_chunk = int(dev::div_ceil(dev::div_ceil(s.K, splits), kernels::CT_BK)) * kernels::CT_BK;
    if (k_chunk == 0) k_chunk = kernels::CT_BK;
    splits = std::max(1, int(dev::div_ceil(s.K, k_chunk)));

    dim3 grid(tiles_n, tiles_m, splits);
    if (splits == 1 && !s.accumulate) {
        KernelScope ks(ctx, "contraction_fp32");
        kernels::contraction_fp32<<<grid, 256, 0, ctx.stream()>>>(
            A, B, s.M, s.N, s.K, s.lda, s.ldb, k_chunk, nullptr, out, s.nsub, s.out_outer);
        check_launch("contraction_fp32");
        return;
    }
    auto partial = ctx.empty<float2, 1>({int64_t(splits) * s.M * s.N});
    {
        KernelScope ks(ctx, "contraction_fp32");
        kernels::contraction_fp32<<<grid, 256, 0, ctx.stream()>>>(
            A, B, s.M, s.N, s.K, s.lda, s.ldb, k_chunk, partial.data(), nullptr, s.nsub, s.out_outer);
        check_launch("contraction_fp32");
    }
    {
        KernelScope ks(ctx, "contraction_reduce");
        int64_t total = int64_t(s.M) * s.N;
        kernels::contraction_reduce<<<dev::div_ceil(total, 256), 256, 0, ctx.stream()>>>(
            partial.data(), splits, s.M, s.N, out, s.nsub, s.out_outer, s.accumulate);
        check_launch("contraction_reduce");
    }
}

}  // namespace compas
