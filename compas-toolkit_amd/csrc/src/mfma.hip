// MFMA complex contraction on gfx950: C[m][n] = sum_k A[m][k] B[n][k].
//
// Complex-to-real mapping.  With A stored interleaved (re, im) along K, row
// m of A is a real vector of length 2K.  For each complex column n the real
// GEMM uses two B rows built while staging into LDS:
//   row 2n   = (b_re, -b_im) pairs  ->  D[m][2n]   = Re(sum a b)
//   row 2n+1 = (b_im,  b_re) pairs  ->  D[m][2n+1] = Im(sum a b)
// so one real GEMM (M x 2N x 2K) produces interleaved complex output.
//
// Precision.  Every FP32 operand x is split into hi = round(x) and
// lo = round(x - hi) in BF16 (or FP16 after a per-row power-of-two scale that
// keeps both parts in the normal range); the product uses
// hi*hi + hi*lo + lo*hi with FP32 accumulation in the MFMA, which recovers
// ~16 (BF16) or ~22 (FP16) bits per product.  BF16x1 uses hi*hi only.
//
// Tiling.  256 threads = 4 waves (2 x 2); a workgroup owns 128 rows x 128
// real columns (64 complex) and walks a K slice in steps of 32 complex
// elements.  Each wave computes 64 x 64 with 2 x 2 v_mfma_f32_32x32x16
// blocks.  Workgroups are numbered so that the ones sharing A rows run on the
// same XCD (consecutive hardware ids are spread over the 8 XCDs), which lets
// the XCD's L2 serve the re-reads of A across column tiles.  K slices are
// reduced in FP64 by contraction_reduce.

#include <cmath>

#include "compas/device/common.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int MF_BM = 128;  // A rows per workgroup
constexpr int MF_BNC = 64;  // complex B columns per workgroup (128 real)
constexpr int MF_BKC = 32;  // complex K per step (64 real)
constexpr int MF_LD = 64 + 8;  // LDS row stride in 16-bit elements (padding against bank conflicts)

enum { P_BF16X3 = 0, P_FP16X3 = 1, P_BF16X1 = 2 };

template<int P>
struct Half16;
template<>
struct Half16<P_BF16X3> {
    using T = __bf16;
    using V = bf16x8;
};
template<>
struct Half16<P_BF16X1> {
    using T = __bf16;
    using V = bf16x8;
};
template<>
struct Half16<P_FP16X3> {
    using T = _Float16;
    using V = f16x8;
};

template<int P>
__device__ inline void split_hilo(float x, typename Half16<P>::T& hi, typename Half16<P>::T& lo) {
    using T = typename Half16<P>::T;
    hi = T(x);
    if constexpr (P != P_BF16X1) lo = T(x - float(hi));
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// (a, b) -> packed 16-bit pair hi = round(a, b) and lo = round((a, b) - hi),
// first element in the low half (one v_cvt_pk_* per pair)
template<int P>
__device__ inline void split_pair(float a, float b, uint32_t& hi, uint32_t& lo) {
    const f32x2 x = {a, b};
    if constexpr (P == P_FP16X3) {
        const f16x2 h = __builtin_convertvector(x, f16x2);
        const f32x2 r = x - __builtin_convertvector(h, f32x2);
        hi = __builtin_bit_cast(uint32_t, h);
        lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, f16x2));
    } else {
        const bf16x2 h = __builtin_convertvector(x, bf16x2);
        hi = __builtin_bit_cast(uint32_t, h);
        if constexpr (P != P_BF16X1) {
            // bf16 -> f32 is a 16-bit shift of the bit pattern
            const f32x2 hf = {__builtin_bit_cast(float, hi << 16), __builtin_bit_cast(float, hi & 0xFFFF0000u)};
            lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(x - hf, bf16x2));
        } else {
            lo = 0u;
        }
    }
}

template<int P>
__device__ inline f32x16 mfma(typename Half16<P>::V a, typename Half16<P>::V b, f32x16 c) {
    if constexpr (P == P_FP16X3) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    } else {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
}

// Logical tile index from the hardware workgroup id: hardware ids b, b+8, ...
// land on the same XCD, so give each XCD a contiguous range of logical ids.
__device__ inline int xcd_logical_id(int b, int total) {
    const int per = (total + 7) / 8;
    const int xcd = b % 8, slot = b / 8;
    const int full = total % 8 == 0 ? 8 : total % 8;  // XCDs that receive `per` ids
    int base = xcd < full ? xcd * per : full * per + (xcd - full) * (per - 1);
    return base + slot;
}

template<int P>
__global__ void __launch_bounds__(256) contraction_mfma(
    const cfloat* __restrict__ A,
    const cfloat* __restrict__ B,
    int M, int N, int K, int64_t lda, int64_t ldb,
    int k_chunk, int tiles_m, int tiles_n, int splits,
    const float* __restrict__ row_scale_a,  // FP16: per-row power-of-two scales (else nullptr)
    const float* __restrict__ row_scale_b,
    float2* __restrict__ partial,
    cfloat* __restrict__ direct, int nsub, int64_t out_outer) {
    using T = typename Half16<P>::T;
    using V = typename Half16<P>::V;
    constexpr bool LO = (P != P_BF16X1);

    __shared__ __attribute__((aligned(16))) T sA_hi[MF_BM * MF_LD];
    __shared__ __attribute__((aligned(16))) T sB_hi[2 * MF_BNC * MF_LD];
    __shared__ __attribute__((aligned(16))) T sA_lo[LO ? MF_BM * MF_LD : 8];
    __shared__ __attribute__((aligned(16))) T sB_lo[LO ? 2 * MF_BNC * MF_LD : 8];

    const int total = tiles_m * tiles_n * splits;
    const int lid = xcd_logical_id(blockIdx.x, total);
    if (lid >= total) return;
    const int tn = lid % tiles_n;
    const int tm = (lid / tiles_n) % tiles_m;
    const int split = lid / (tiles_n * tiles_m);
    const int m0 = tm * MF_BM, n0 = tn * MF_BNC;
    const int kb = split * k_chunk;
    const int ke = min(K, kb + k_chunk);

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;

    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[i][j][e] = 0.0f;

    // staging assignment: A -> row tid/2, 16 complex; B -> row tid/4, 8 complex
    const int a_row = tid >> 1, a_k = (tid & 1) * 16;
    const int b_row = tid >> 2, b_k = (tid & 3) * 8;
    const int am = m0 + a_row, bn = n0 + b_row;
    const float sa = (row_scale_a && am < M) ? row_scale_a[am] : 1.0f;
    const float sb = (row_scale_b && bn < N) ? row_scale_b[bn] : 1.0f;
    const cfloat* a_src = A + int64_t(am) * lda;
    const cfloat* b_src = B + int64_t(bn) * ldb;
    // 16-byte vector loads need 16-byte aligned rows
    const bool vec_a = ((reinterpret_cast<uintptr_t>(A) & 15) == 0) && (lda % 2 == 0);
    const bool vec_b = ((reinterpret_cast<uintptr_t>(B) & 15) == 0) && (ldb % 2 == 0);

    // register prefetch of one K step: 16 complex of A, 8 complex of B
    cfloat ra[16], rb[8];
    auto gload = [&](int k0) {
        const int ka = k0 + a_k;
        if (vec_a && am < M && ka + 16 <= ke) {
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const float4 v = *reinterpret_cast<const float4*>(a_src + ka + 2 * i);
                ra[2 * i] = {v.x, v.y};
                ra[2 * i + 1] = {v.z, v.w};
            }
        } else {
#pragma unroll
            for (int i = 0; i < 16; i++) ra[i] = (am < M && ka + i < ke) ? a_src[ka + i] : cfloat {0, 0};
        }
        const int kb_ = k0 + b_k;
        if (vec_b && bn < N && kb_ + 8 <= ke) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float4 v = *reinterpret_cast<const float4*>(b_src + kb_ + 2 * i);
                rb[2 * i] = {v.x, v.y};
                rb[2 * i + 1] = {v.z, v.w};
            }
        } else {
#pragma unroll
            for (int i = 0; i < 8; i++) rb[i] = (bn < N && kb_ + i < ke) ? b_src[kb_ + i] : cfloat {0, 0};
        }
    };

    if (kb < ke) gload(kb);
    for (int k0 = kb; k0 < ke; k0 += MF_BKC) {
        // ---- stage A: (re, im) pairs packed hi / lo, 16-byte LDS stores
        {
            uint32_t h[16], l[16];
#pragma unroll
            for (int i = 0; i < 16; i++) split_pair<P>(ra[i].re * sa, ra[i].im * sa, h[i], l[i]);
            uint4* dh = reinterpret_cast<uint4*>(sA_hi + a_row * MF_LD + 2 * a_k);
#pragma unroll
            for (int i = 0; i < 4; i++) dh[i] = make_uint4(h[4 * i], h[4 * i + 1], h[4 * i + 2], h[4 * i + 3]);
            if constexpr (LO) {
                uint4* dl = reinterpret_cast<uint4*>(sA_lo + a_row * MF_LD + 2 * a_k);
#pragma unroll
                for (int i = 0; i < 4; i++) dl[i] = make_uint4(l[4 * i], l[4 * i + 1], l[4 * i + 2], l[4 * i + 3]);
            }
        }
        // ---- stage B: conj row (re, -im) and swap row (im, re)
        {
            uint32_t ch[8], sh[8], cl[8], sl[8];
#pragma unroll
            for (int i = 0; i < 8; i++) {
                uint32_t h, l;
                split_pair<P>(rb[i].re * sb, rb[i].im * sb, h, l);
                ch[i] = h ^ 0x80000000u;  // negate the imaginary half
                sh[i] = __builtin_amdgcn_alignbit(h, h, 16);  // swap halves
                cl[i] = l ^ 0x80000000u;
                sl[i] = __builtin_amdgcn_alignbit(l, l, 16);
            }
            uint4* dch = reinterpret_cast<uint4*>(sB_hi + (2 * b_row) * MF_LD + 2 * b_k);
            uint4* dsh = reinterpret_cast<uint4*>(sB_hi + (2 * b_row + 1) * MF_LD + 2 * b_k);
#pragma unroll
            for (int i = 0; i < 2; i++) {
                dch[i] = make_uint4(ch[4 * i], ch[4 * i + 1], ch[4 * i + 2], ch[4 * i + 3]);
                dsh[i] = make_uint4(sh[4 * i], sh[4 * i + 1], sh[4 * i + 2], sh[4 * i + 3]);
            }
            if constexpr (LO) {
                uint4* dcl = reinterpret_cast<uint4*>(sB_lo + (2 * b_row) * MF_LD + 2 * b_k);
                uint4* dsl = reinterpret_cast<uint4*>(sB_lo + (2 * b_row + 1) * MF_LD + 2 * b_k);
#pragma unroll
                for (int i = 0; i < 2; i++) {
                    dcl[i] = make_uint4(cl[4 * i], cl[4 * i + 1], cl[4 * i + 2], cl[4 * i + 3]);
                    dsl[i] = make_uint4(sl[4 * i], sl[4 * i + 1], sl[4 * i + 2], sl[4 * i + 3]);
                }
            }
        }
        __syncthreads();
        if (k0 + MF_BKC < ke) gload(k0 + MF_BKC);  // in flight during the MFMAs below

#pragma unroll
        for (int kk = 0; kk < 4; kk++) {
            const int koff = kk * 16 + (lane >> 5) * 8;
            V ah[2], al[2], bh[2], bl[2];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const int r = wm * 64 + i * 32 + (lane & 31);
                ah[i] = *reinterpret_cast<const V*>(sA_hi + r * MF_LD + koff);
                if constexpr (LO) al[i] = *reinterpret_cast<const V*>(sA_lo + r * MF_LD + koff);
            }
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int c = wn * 64 + j * 32 + (lane & 31);
                bh[j] = *reinterpret_cast<const V*>(sB_hi + c * MF_LD + koff);
                if constexpr (LO) bl[j] = *reinterpret_cast<const V*>(sB_lo + c * MF_LD + koff);
            }
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    acc[i][j] = mfma<P>(ah[i], bh[j], acc[i][j]);
                    if constexpr (LO) {
                        acc[i][j] = mfma<P>(ah[i], bl[j], acc[i][j]);
                        acc[i][j] = mfma<P>(al[i], bh[j], acc[i][j]);
                    }
                }
        }
        __syncthreads();
    }

    // ---- epilogue: D[row][realcol], realcol = 2n + {re, im}
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const int row = wm * 64 + i * 32 + (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3);
                const int rc = wn * 64 + j * 32 + (lane & 31);
                const int m = m0 + row, n = n0 + (rc >> 1), part = rc & 1;
                if (m >= M || n >= N) continue;
                float v = acc[i][j][e];
                if (row_scale_a) v /= row_scale_a[m] * row_scale_b[n];
                if (direct) {
                    const int64_t o = int64_t(n / nsub) * out_outer + int64_t(m) * nsub + (n % nsub);
                    reinterpret_cast<float*>(direct)[2 * o + part] = v;
                } else {
                    reinterpret_cast<float*>(partial)[2 * ((int64_t(split) * M + m) * N + n) + part] = v;
                }
            }
}

// Per-row power of two s with max_k |x[row][k]| * s ~ 2^14 (FP16 operands).
__global__ void __launch_bounds__(256) row_pow2_scale(const cfloat* __restrict__ X, int rows, int K, int64_t ld,
                                                      float* __restrict__ scale) {
    __shared__ float red[256];
    const int r = blockIdx.x;
    float mx = 0.0f;
    for (int k = threadIdx.x; k < K; k += 256) {
        const cfloat v = X[int64_t(r) * ld + k];
        mx = fmaxf(mx, fmaxf(fabsf(v.re), fabsf(v.im)));
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float m = red[0];
        int e = 0;
        if (m > 0.0f && isfinite(m)) {
            frexpf(m, &e);  // m = f * 2^e, f in [0.5, 1)
        }
        scale[r] = ldexpf(1.0f, 14 - e);
    }
}

__global__ void contraction_reduce_mf(
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

static int cu_count(int device) {
    static int cached[64] = {0};
    if (device >= 0 && device < 64 && cached[device]) return cached[device];
    hipDeviceProp_t prop;
    COMPAS_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    if (device >= 0 && device < 64) cached[device] = prop.multiProcessorCount;
    return prop.multiProcessorCount;
}

template<int P>
static void launch_mfma(const CompasContext& ctx, const cfloat* A, const cfloat* B, cfloat* out, ContractionShape s,
                        const char* name) {
    using namespace kernels;
    const int tiles_m = int(dev::div_ceil(s.M, MF_BM));
    const int tiles_n = int(dev::div_ceil(s.N, MF_BNC));
    const int tiles = tiles_m * tiles_n;
    // enough workgroups for ~2 per CU, K slices of at least 8 steps
    const int target = 2 * cu_count(ctx.device());
    int splits = std::max(1, (target + tiles - 1) / tiles);
    splits = std::min(splits, std::max(1, s.K / (8 * MF_BKC)));
    int k_chunk = int(dev::div_ceil(dev::div_ceil(s.K, splits), MF_BKC)) * MF_BKC;
    splits = std::max(1, int(dev::div_ceil(s.K, k_chunk)));
    const int total = tiles * splits;

    Array<float, 1> scale_a, scale_b;
    const float* sa = nullptr;
    const float* sb = nullptr;
    if (P == P_FP16X3) {
        scale_a = ctx.empty<float, 1>({s.M});
        scale_b = ctx.empty<float, 1>({s.N});
        KernelScope ks(ctx, "row_pow2_scale");
        row_pow2_scale<<<s.M, 256, 0, ctx.stream()>>>(A, s.M, s.K, s.lda, scale_a.data());
        row_pow2_scale<<<s.N, 256, 0, ctx.stream()>>>(B, s.N, s.K, s.ldb, scale_b.data());
        check_launch("row_pow2_scale");
        sa = scale_a.data();
        sb = scale_b.data();
    }

    const bool direct = (splits == 1 && !s.accumulate);
    Array<float2, 1> partial;
    if (!direct) partial = ctx.empty<float2, 1>({int64_t(splits) * s.M * s.N});
    {
        KernelScope ks(ctx, name);
        contraction_mfma<P><<<total, 256, 0, ctx.stream()>>>(
            A, B, s.M, s.N, s.K, s.lda, s.ldb, k_chunk, tiles_m, tiles_n, splits, sa, sb,
            direct ? nullptr : partial.data(), direct ? out : nullptr, s.nsub, s.out_outer);
        check_launch(name);
    }
    if (!direct) {
        KernelScope ks(ctx, "contraction_reduce");
        const int64_t n = int64_t(s.M) * s.N;
        contraction_reduce_mf<<<dev::div_ceil(n, 256), 256, 0, ctx.stream()>>>(
            partial.data(), splits, s.M, s.N, out, s.nsub, s.out_outer, s.accumulate);
        check_launch("contraction_reduce");
    }
}

void complex_contraction_mfma(
    const CompasContext& ctx,
    const cfloat* A,
    const cfloat* B,
    cfloat* out,
    ContractionShape s,
    ContractionPrecision precision,
    bool* handled) {
    *handled = true;
    switch (precision) {
        case ContractionPrecision::BF16x3:
            launch_mfma<kernels::P_BF16X3>(ctx, A, B, out, s, "contraction_mfma");
            return;
        case ContractionPrecision::FP16x3:
            launch_mfma<kernels::P_FP16X3>(ctx, A, B, out, s, "contraction_mfma");
            return;
        case ContractionPrecision::BF16x1:
            launch_mfma<kernels::P_BF16X1>(ctx, A, B, out, s, "contraction_mfma");
            return;
        default:
            *handled = false;
    }
}

}  // namespace compas
