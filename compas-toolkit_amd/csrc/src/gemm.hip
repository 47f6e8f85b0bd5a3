// GEMM utility of the compas:: API: result = alpha * lhs * rhs^T + beta * result
// on real [m][k] x [n][k] operands and on planar complex [2][m][k] x [2][n][k]
// operands, plus the planar <-> interleaved converters
// (reference include/compas/utils/gemm.h:7-71, src/utils/gemm.cu:8-254,
// src/utils/gemm_kernels.cuh:9-32).
//
// The reference hands these to cuBLAS/rocBLAS with a compute type chosen by
// GemmComputeMethod.  Here they are project kernels:
//   * FP32 operands.  Pedantic: FP32 VALU kernel (one FMA per product, no
//     reduced precision).  Regular: MFMA with each operand split into BF16
//     hi + lo and hi*hi + hi*lo + lo*hi accumulated in FP32 (about 16 bits
//     per product, FP32-level results).  Fast: MFMA on the BF16 hi parts
//     only (one product; coarser than the reference's TF32).
//   * BF16 operands: every product is exact in FP32, so all methods run one
//     MFMA product per element with FP32 accumulation.
//   * Complex: Re = Lr Rr^T - Li Ri^T and Im = Lr Ri^T + Li Rr^T
//     (gemm.cu:147-173); each plane is ONE launch over the two terms
//     concatenated along K instead of two GEMMs.
//   * K is split across workgroups; partial sums are reduced in FP64, and the
//     reduction applies alpha and beta (beta = 0 does not read the result, as
//     in BLAS).
//
// MFMA tiling: 256 threads = 2 x 2 waves, each wave 2 x 2 blocks of
// v_mfma_f32_32x32x16 (64 x 64), workgroup tile 128 x 128, 32 K per step.
// Operands are staged through LDS as packed 16-bit values (hi and lo planes).
// Accumulator layout of v_mfma_f32_32x32x16: lane l, element e of a block is
// row 8 (e / 4) + 4 (l / 32) + e % 4 and column l % 32.

#include <algorithm>
#include <cmath>

#include "compas/compas.h"
#include "compas/device/common.h"
#include "compas/device/mfma_tile.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

constexpr int GM_BM = 128, GM_BN = 128, GM_BK = 32, GM_LD = GM_BK + 8;  // LD: 16-bit elements per LDS row
constexpr int GV_BM = 64, GV_BN = 64, GV_BK = 16;

// Operands of one launch: up to two (lhs, rhs) terms summed along a virtual K
// axis of nterms * Kp (Kp = K rounded up to the kernel's K step, so that no
// step straddles two terms); the rhs of term 1 is multiplied by sign1.
template<typename IN>
struct GemmArgs {
    const IN* L[2];
    const IN* R[2];
    float sign1;
    int M, N, K, Kp, nterms;
    int64_t ldl, ldr;
    int k_chunk;  // virtual K per workgroup along z (a multiple of the K step)
    float* partial;  // [splits][M][N]
};

__device__ inline float to_float(float x) {
    return x;
}
__device__ inline float to_float(__bf16 x) {
    return float(x);
}

// n consecutive values of one row starting at local column k (< K), zero
// past K or for a row outside the operand (row == nullptr)
template<typename IN, int NV>
__device__ inline void load_row(const IN* row, int k, int K, bool vec, float scale, float (&v)[NV]) {
    if (row == nullptr) {
#pragma unroll
        for (int i = 0; i < NV; i++) v[i] = 0.0f;
        return;
    }
    if constexpr (std::is_same<IN, float>::value) {
        if (vec && k + NV <= K) {
#pragma unroll
            for (int i = 0; i < NV / 4; i++) {
                const float4 q = *reinterpret_cast<const float4*>(row + k + 4 * i);
                v[4 * i] = q.x * scale;
                v[4 * i + 1] = q.y * scale;
                v[4 * i + 2] = q.z * scale;
                v[4 * i + 3] = q.w * scale;
            }
            return;
        }
    }
#pragma unroll
    for (int i = 0; i < NV; i++) v[i] = (k + i < K) ? to_float(row[k + i]) * scale : 0.0f;
}

// The 16 bf16 bits of the row slice directly (BF16 operands: no split)
__device__ inline void load_row_bits(const __bf16* row, int k, int K, bool vec, bool negate, uint32_t (&w)[8]) {
    const uint32_t flip = negate ? 0x80008000u : 0u;
    if (row != nullptr && vec && k + 16 <= K) {
        const uint4 a = *reinterpret_cast<const uint4*>(row + k);
        const uint4 b = *reinterpret_cast<const uint4*>(row + k + 8);
        w[0] = a.x ^ flip, w[1] = a.y ^ flip, w[2] = a.z ^ flip, w[3] = a.w ^ flip;
        w[4] = b.x ^ flip, w[5] = b.y ^ flip, w[6] = b.z ^ flip, w[7] = b.w ^ flip;
        return;
    }
    const uint16_t* r16 = reinterpret_cast<const uint16_t*>(row);
    const uint16_t f16 = negate ? 0x8000u : 0u;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const uint32_t lo = (row != nullptr && k + 2 * i < K) ? uint32_t(r16[k + 2 * i] ^ f16) : 0u;
        const uint32_t hi = (row != nullptr && k + 2 * i + 1 < K) ? uint32_t(r16[k + 2 * i + 1] ^ f16) : 0u;
        w[i] = lo | (hi << 16);
    }
}

// P: mfma::P_BF16X3 (hi + lo) or P_BF16X1 (hi only); IN: float or __bf16
template<int P, typename IN>
__global__ void __launch_bounds__(256) gemm_mfma(GemmArgs<IN> a) {
    using mfma::bf16x8;
    using mfma::f32x16;
    constexpr bool LO = mfma::has_lo(P) && std::is_same<IN, float>::value;
    __shared__ __attribute__((aligned(16))) __bf16 a_hi[GM_BM * GM_LD];
    __shared__ __attribute__((aligned(16))) __bf16 b_hi[GM_BN * GM_LD];
    __shared__ __attribute__((aligned(16))) __bf16 a_lo[LO ? GM_BM * GM_LD : 8];
    __shared__ __attribute__((aligned(16))) __bf16 b_lo[LO ? GM_BN * GM_LD : 8];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.y * GM_BM, n0 = blockIdx.x * GM_BN;
    const int kb = blockIdx.z * a.k_chunk;
    const int ke = min(a.nterms * a.Kp, kb + a.k_chunk);

    // staging: thread t stages 16 consecutive K values of row t / 2 of A and of B
    const int srow = tid >> 1, skp = (tid & 1) * 16;
    const int am = m0 + srow, bn = n0 + srow;
    const bool vec_l = std::is_same<IN, float>::value ? (a.ldl % 4 == 0) : (a.ldl % 8 == 0);
    const bool vec_r = std::is_same<IN, float>::value ? (a.ldr % 4 == 0) : (a.ldr % 8 == 0);

    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[i][j][e] = 0.0f;

    for (int k0 = kb; k0 < ke; k0 += GM_BK) {
        const int term = k0 / a.Kp, kl = k0 - term * a.Kp + skp;
        const IN* lrow = am < a.M ? a.L[term] + int64_t(am) * a.ldl : nullptr;
        const IN* rrow = bn < a.N ? a.R[term] + int64_t(bn) * a.ldr : nullptr;
        const bool neg = term == 1 && a.sign1 < 0.0f;
        uint32_t ah[8], al[8], bh[8], bl[8];
        if constexpr (std::is_same<IN, float>::value) {
            float va[16], vb[16];
            load_row<IN, 16>(lrow, kl, a.K, vec_l, 1.0f, va);
            load_row<IN, 16>(rrow, kl, a.K, vec_r, neg ? -1.0f : 1.0f, vb);
#pragma unroll
            for (int i = 0; i < 8; i++) {
                mfma::split_pair<P>(va[2 * i], va[2 * i + 1], ah[i], al[i]);
                mfma::split_pair<P>(vb[2 * i], vb[2 * i + 1], bh[i], bl[i]);
            }
        } else {
            load_row_bits(lrow, kl, a.K, vec_l, false, ah);
            load_row_bits(rrow, kl, a.K, vec_r, neg, bh);
        }
        __syncthreads();  // the previous step's fragment reads are done
        {
            uint4* pa = reinterpret_cast<uint4*>(a_hi + srow * GM_LD + skp);
            uint4* pb = reinterpret_cast<uint4*>(b_hi + srow * GM_LD + skp);
            pa[0] = make_uint4(ah[0], ah[1], ah[2], ah[3]);
            pa[1] = make_uint4(ah[4], ah[5], ah[6], ah[7]);
            pb[0] = make_uint4(bh[0], bh[1], bh[2], bh[3]);
            pb[1] = make_uint4(bh[4], bh[5], bh[6], bh[7]);
            if constexpr (LO) {
                uint4* qa = reinterpret_cast<uint4*>(a_lo + srow * GM_LD + skp);
                uint4* qb = reinterpret_cast<uint4*>(b_lo + srow * GM_LD + skp);
                qa[0] = make_uint4(al[0], al[1], al[2], al[3]);
                qa[1] = make_uint4(al[4], al[5], al[6], al[7]);
                qb[0] = make_uint4(bl[0], bl[1], bl[2], bl[3]);
                qb[1] = make_uint4(bl[4], bl[5], bl[6], bl[7]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GM_BK / 16; kk++) {
            const int koff = kk * 16 + (lane >> 5) * 8;
            bf16x8 fa[2], fb[2], ga[2], gb[2];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const int r = wm * 64 + i * 32 + (lane & 31);
                fa[i] = *reinterpret_cast<const bf16x8*>(a_hi + r * GM_LD + koff);
                if constexpr (LO) ga[i] = *reinterpret_cast<const bf16x8*>(a_lo + r * GM_LD + koff);
            }
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int c = wn * 64 + j * 32 + (lane & 31);
                fb[j] = *reinterpret_cast<const bf16x8*>(b_hi + c * GM_LD + koff);
                if constexpr (LO) gb[j] = *reinterpret_cast<const bf16x8*>(b_lo + c * GM_LD + koff);
            }
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
            if constexpr (LO) {
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int j = 0; j < 2; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], gb[j], acc[i][j], 0, 0, 0);
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int j = 0; j < 2; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ga[i], fb[j], acc[i][j], 0, 0, 0);
            }
        }
    }

    float* out = a.partial + int64_t(blockIdx.z) * a.M * a.N;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const int m = m0 + wm * 64 + i * 32 + 8 * (e / 4) + 4 * (lane >> 5) + (e % 4);
                const int n = n0 + wn * 64 + j * 32 + (lane & 31);
                if (m < a.M && n < a.N) out[int64_t(m) * a.N + n] = acc[i][j][e];
            }
}

// FP32 VALU form (Pedantic): 64 x 64 tile, 4 x 4 outputs per thread, 16 K per step
__global__ void __launch_bounds__(256) gemm_fp32(GemmArgs<float> a) {
    __shared__ float Ls[GV_BK][GV_BM + 1];
    __shared__ float Rs[GV_BK][GV_BN + 1];
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int m0 = blockIdx.y * GV_BM, n0 = blockIdx.x * GV_BN;
    const int kb = blockIdx.z * a.k_chunk;
    const int ke = min(a.nterms * a.Kp, kb + a.k_chunk);
    const int srow = tid >> 2, skp = (tid & 3) * 4;
    const int am = m0 + srow, bn = n0 + srow;
    const bool vec_l = a.ldl % 4 == 0, vec_r = a.ldr % 4 == 0;

    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0.0f;

    for (int k0 = kb; k0 < ke; k0 += GV_BK) {
        const int term = k0 / a.Kp, kl = k0 - term * a.Kp + skp;
        const float* lrow = am < a.M ? a.L[term] + int64_t(am) * a.ldl : nullptr;
        const float* rrow = bn < a.N ? a.R[term] + int64_t(bn) * a.ldr : nullptr;
        float vl[4], vr[4];
        load_row<float, 4>(lrow, kl, a.K, vec_l, 1.0f, vl);
        load_row<float, 4>(rrow, kl, a.K, vec_r, (term == 1 && a.sign1 < 0.0f) ? -1.0f : 1.0f, vr);
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; i++) {
            Ls[skp + i][srow] = vl[i];
            Rs[skp + i][srow] = vr[i];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GV_BK; kk++) {
            float l[4], r[4];
#pragma unroll
            for (int i = 0; i < 4; i++) l[i] = Ls[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; j++) r[j] = Rs[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = fmaf(l[i], r[j], acc[i][j]);
        }
    }
    float* out = a.partial + int64_t(blockIdx.z) * a.M * a.N;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int m = m0 + ty + 16 * i, n = n0 + tx + 16 * j;
            if (m < a.M && n < a.N) out[int64_t(m) * a.N + n] = acc[i][j];
        }
}

// result[m][n] = alpha * sum_s partial[s][m][n] + beta * result[m][n] (FP64 sum)
__global__ void __launch_bounds__(256) gemm_reduce(
    const float* __restrict__ partial, int splits, int64_t mn, float alpha, float beta, float* result) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= mn) return;
    double s = 0.0;
    for (int p = 0; p < splits; p++) s += double(partial[int64_t(p) * mn + i]);
    double r = double(alpha) * s;
    if (beta != 0.0f) r += double(beta) * double(result[i]);
    result[i] = float(r);
}

template<typename OUT>
__global__ void complex_to_planar(int n, int m, const cfloat* __restrict__ in, OUT* __restrict__ out) {
    const int64_t total = int64_t(n) * m;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
        const cfloat v = in[i];
        out[i] = OUT(v.re);
        out[total + i] = OUT(v.im);
    }
}

__global__ void planar_to_complex(int n, int m, const float* __restrict__ in, cfloat* __restrict__ out) {
    const int64_t total = int64_t(n) * m;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x)
        out[i] = {in[i], in[total + i]};
}

}  // namespace kernels

namespace {

int device_cus(int device) {
    static int cached[256] = {0};
    if (device >= 0 && device < 256 && cached[device]) return cached[device];
    hipDeviceProp_t prop;
    COMPAS_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    if (device >= 0 && device < 256) cached[device] = prop.multiProcessorCount;
    return prop.multiProcessorCount;
}

enum class GemmKernel { FP32, BF16X3, BF16X1 };

// One launch (plus its FP64 reduction) of result = alpha * sum_t L_t R_t^T
// (R_1 times sign1) + beta * result; result is [M][N] contiguous.
template<typename IN>
void gemm_launch(
    const CompasContext& ctx, GemmKernel kind, const IN* const* L, const IN* const* R, int nterms, float sign1,
    int M, int N, int K, int64_t ldl, int64_t ldr, float alpha, float beta, float* result) {
    if (M == 0 || N == 0) return;
    ctx.activate();
    const bool valu = kind == GemmKernel::FP32;
    const int BM = valu ? kernels::GV_BM : kernels::GM_BM;
    const int BN = valu ? kernels::GV_BN : kernels::GM_BN;
    const int BK = valu ? kernels::GV_BK : kernels::GM_BK;
    const int Kp = int(dev::div_ceil(std::max(K, 1), BK)) * BK;
    const int steps = int(dev::div_ceil(int64_t(nterms) * Kp, BK));
    const int tiles = int(dev::div_ceil(M, BM) * dev::div_ceil(N, BN));
    // about two workgroups per CU, at least 8 K steps per workgroup
    int splits = std::max(1, int(dev::div_ceil(2 * device_cus(ctx.device()), tiles)));
    splits = std::min(splits, std::max(1, steps / 8));
    const int k_chunk = int(dev::div_ceil(steps, splits)) * BK;
    splits = int(dev::div_ceil(int64_t(nterms) * Kp, k_chunk));

    auto partial = ctx.empty<float, 1>({int64_t(splits) * M * N});
    kernels::GemmArgs<IN> a;
    for (int t = 0; t < 2; t++) {
        a.L[t] = L[std::min(t, nterms - 1)];
        a.R[t] = R[std::min(t, nterms - 1)];
    }
    a.sign1 = sign1;
    a.M = M, a.N = N, a.K = K, a.Kp = Kp, a.nterms = nterms;
    a.ldl = ldl, a.ldr = ldr;
    a.k_chunk = k_chunk;
    a.partial = partial.data();
    const dim3 grid(unsigned(dev::div_ceil(N, BN)), unsigned(dev::div_ceil(M, BM)), unsigned(splits));
    {
        KernelScope ks(ctx, valu ? "gemm_fp32" : "gemm_mfma");
        if constexpr (std::is_same<IN, float>::value) {
            if (kind == GemmKernel::FP32) {
                kernels::gemm_fp32<<<grid, 256, 0, ctx.stream()>>>(a);
            } else if (kind == GemmKernel::BF16X3) {
                kernels::gemm_mfma<mfma::P_BF16X3, float><<<grid, 256, 0, ctx.stream()>>>(a);
            } else {
                kernels::gemm_mfma<mfma::P_BF16X1, float><<<grid, 256, 0, ctx.stream()>>>(a);
            }
        } else {
            kernels::gemm_mfma<mfma::P_BF16X1, __bf16><<<grid, 256, 0, ctx.stream()>>>(a);
        }
        check_launch("gemm");
    }
    {
        KernelScope ks(ctx, "gemm_reduce");
        const int64_t mn = int64_t(M) * N;
        kernels::gemm_reduce<<<unsigned(dev::div_ceil(mn, 256)), 256, 0, ctx.stream()>>>(
            partial.data(), splits, mn, alpha, beta, result);
        check_launch("gemm_reduce");
    }
}

GemmKernel float_kernel(GemmComputeMethod kind) {
    switch (kind) {
        case GemmComputeMethod::Pedantic:
            return GemmKernel::FP32;
        case GemmComputeMethod::Regular:
            return GemmKernel::BF16X3;
        case GemmComputeMethod::Fast:
            return GemmKernel::BF16X1;
    }
    COMPAS_ERROR("invalid value for `GemmComputeMethod`");
}

void check_int_range(int64_t v) {
    if (v > int64_t(1) << 30) COMPAS_ERROR("GEMM dimension too large");
}

template<typename IN>
void gemm_impl(
    const CompasContext& ctx, Array<float, 2> result, Array<IN, 2> lhs, Array<IN, 2> rhs, float alpha, float beta,
    GemmKernel kind) {
    const int64_t m = result.size(0), n = result.size(1), k = lhs.size(1);
    COMPAS_CHECK(lhs.size(0) == m);  // gemm.cu:23-28
    COMPAS_CHECK(rhs.size(0) == n);
    COMPAS_CHECK(rhs.size(1) == k);
    check_int_range(m), check_int_range(n), check_int_range(k);
    const IN* L[1] = {lhs.data()};
    const IN* R[1] = {rhs.data()};
    gemm_launch<IN>(ctx, kind, L, R, 1, 1.0f, int(m), int(n), int(k), k, k, alpha, beta, result.data());
}

template<typename IN>
void complex_gemm_impl(
    const CompasContext& ctx, Array<float, 3> result, Array<IN, 3> lhs, Array<IN, 3> rhs, float alpha, float beta,
    GemmKernel kind) {
    COMPAS_CHECK(result.size(0) == 2);  // gemm.cu:155-157
    COMPAS_CHECK(lhs.size(0) == 2);
    COMPAS_CHECK(rhs.size(0) == 2);
    const int64_t m = result.size(1), n = result.size(2), k = lhs.size(2);
    COMPAS_CHECK(lhs.size(1) == m);
    COMPAS_CHECK(rhs.size(1) == n);
    COMPAS_CHECK(rhs.size(2) == k);
    check_int_range(m), check_int_range(n), check_int_range(k);
    const IN *lr = lhs.data(), *li = lhs.data() + m * k;
    const IN *rr = rhs.data(), *ri = rhs.data() + n * k;
    float* re = result.data();
    float* im = result.data() + m * n;
    {  // Re = Lr Rr^T - Li Ri^T
        const IN* L[2] = {lr, li};
        const IN* R[2] = {rr, ri};
        gemm_launch<IN>(ctx, kind, L, R, 2, -1.0f, int(m), int(n), int(k), k, k, alpha, beta, re);
    }
    {  // Im = Lr Ri^T + Li Rr^T
        const IN* L[2] = {lr, li};
        const IN* R[2] = {ri, rr};
        gemm_launch<IN>(ctx, kind, L, R, 2, 1.0f, int(m), int(n), int(k), k, k, alpha, beta, im);
    }
}

}  // namespace

void compute_gemm(
    const CompasContext& ctx, Array<float, 2> result, Array<float, 2> lhs, Array<float, 2> rhs, float alpha,
    float beta, GemmComputeMethod kind) {
    gemm_impl<float>(ctx, result, lhs, rhs, alpha, beta, float_kernel(kind));
}

void compute_gemm(
    const CompasContext& ctx, Array<float, 2> result, Array<bfloat16_t, 2> lhs, Array<bfloat16_t, 2> rhs,
    float alpha, float beta, GemmComputeMethod kind) {
    (void)float_kernel(kind);  // validates the method; BF16 products are exact in every method
    gemm_impl<bfloat16_t>(ctx, result, lhs, rhs, alpha, beta, GemmKernel::BF16X1);
}

void compute_complex_gemm(
    const CompasContext& ctx, Array<float, 3> result, Array<float, 3> lhs, Array<float, 3> rhs, float alpha,
    float beta, GemmComputeMethod kind) {
    complex_gemm_impl<float>(ctx, result, lhs, rhs, alpha, beta, float_kernel(kind));
}

void compute_complex_gemm(
    const CompasContext& ctx, Array<float, 3> result, Array<bfloat16_t, 3> lhs, Array<bfloat16_t, 3> rhs,
    float alpha, float beta, GemmComputeMethod kind) {
    (void)float_kernel(kind);
    complex_gemm_impl<bfloat16_t>(ctx, result, lhs, rhs, alpha, beta, GemmKernel::BF16X1);
}

template<typename OUT>
static void to_planar(const CompasContext& ctx, Array<OUT, 3> output, Array<cfloat, 2> input) {
    const int64_t n = input.size(0), m = input.size(1);  // gemm_kernels.cuh:9-19, gemm.cu:197-220
    COMPAS_CHECK(output.size(0) == 2);
    COMPAS_CHECK(output.size(1) == n);
    COMPAS_CHECK(output.size(2) == m);
    check_int_range(n), check_int_range(m);
    if (n * m == 0) return;
    ctx.activate();
    KernelScope ks(ctx, "complex_to_planar");
    const unsigned blocks = unsigned(std::min<int64_t>(4096, dev::div_ceil(n * m, 256)));
    kernels::complex_to_planar<OUT><<<blocks, 256, 0, ctx.stream()>>>(int(n), int(m), input.data(), output.data());
    check_launch("complex_to_planar");
}

void convert_complex_to_planar(const CompasContext& ctx, Array<float, 3> output, Array<cfloat, 2> input) {
    to_planar<float>(ctx, output, input);
}

void convert_complex_to_planar(const CompasContext& ctx, Array<bfloat16_t, 3> output, Array<cfloat, 2> input) {
    to_planar<bfloat16_t>(ctx, output, input);
}

void convert_planar_to_complex(const CompasContext& ctx, Array<cfloat, 2> output, Array<float, 3> input) {
    const int64_t n = output.size(0), m = output.size(1);  // gemm_kernels.cuh:21-32, gemm.cu:236-254
    COMPAS_CHECK(input.size(0) == 2);
    COMPAS_CHECK(input.size(1) == n);
    COMPAS_CHECK(input.size(2) == m);
    check_int_range(n), check_int_range(m);
    if (n * m == 0) return;
    ctx.activate();
    KernelScope ks(ctx, "planar_to_complex");
    const unsigned blocks = unsigned(std::min<int64_t>(4096, dev::div_ceil(n * m, 256)));
    kernels::planar_to_complex<<<blocks, 256, 0, ctx.stream()>>>(int(n), int(m), input.data(), output.data());
    check_launch("planar_to_complex");
}

}  // namespace compas
