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
// ~16 (BF16) or ~22 (FP16) bits per product.  BF16x1 and FP16x1 use hi*hi
// only.
//
// Tiling.  A workgroup of WM x WN waves owns 64 WM rows x 64 WN real columns
// (32 WN complex) and walks a K slice in steps of BKC complex elements; each
// wave computes 64 x 64 with 2 x 2 v_mfma_f32_32x32x16 blocks.  The next K
// step is prefetched into registers while the MFMAs of the current one run.
// Operands are converted to packed 16-bit pairs on the way into LDS (one
// v_cvt_pk per pair, 128-bit stores).  Workgroups are numbered so that those
// sharing A rows run on the same XCD (consecutive hardware ids are spread
// over the 8 XCDs), so the XCD's L2 serves the re-reads of A across column
// tiles.  K slices are reduced in FP64.
//
// Variants (WM, WN, BKC) are timed on the live shape by a small autotuner on
// first use and the fastest is cached per shape.

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

#include "compas/device/common.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

enum { P_BF16X3 = 0, P_FP16X3 = 1, P_BF16X1 = 2, P_FP16X1 = 3, NUM_P = 4 };

constexpr bool is_fp16(int P) {
    return P == P_FP16X3 || P == P_FP16X1;
}
constexpr bool has_lo(int P) {
    return P == P_BF16X3 || P == P_FP16X3;
}

template<int P, bool F = is_fp16(P)>
struct Half16 {
    using T = __bf16;
    using V = bf16x8;
};
template<int P>
struct Half16<P, true> {
    using T = _Float16;
    using V = f16x8;
};

// (a, b) -> packed 16-bit pair hi = round(a, b) and lo = round((a, b) - hi),
// first element in the low half (one v_cvt_pk_* per pair)
template<int P>
__device__ inline void split_pair(float a, float b, uint32_t& hi, uint32_t& lo) {
    const f32x2 x = {a, b};
    if constexpr (is_fp16(P)) {
        const f16x2 h = __builtin_convertvector(x, f16x2);
        hi = __builtin_bit_cast(uint32_t, h);
        if constexpr (has_lo(P)) {
            const f32x2 r = x - __builtin_convertvector(h, f32x2);
            lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, f16x2));
        } else {
            lo = 0u;
        }
    } else {
        const bf16x2 h = __builtin_convertvector(x, bf16x2);
        hi = __builtin_bit_cast(uint32_t, h);
        if constexpr (has_lo(P)) {
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
    if constexpr (is_fp16(P)) {
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
    const int base = xcd < full ? xcd * per : full * per + (xcd - full) * (per - 1);
    return base + slot;
}

template<int P, int WM, int WN, int BKC>
struct MfmaTile {
    static constexpr int THREADS = 64 * WM * WN;
    static constexpr int BM = 64 * WM;  // A rows
    static constexpr int BNR = 64 * WN;  // real B rows (2 per complex column)
    static constexpr int BNC = 32 * WN;  // complex B columns
    static constexpr int LD = 2 * BKC + 8;  // LDS row stride, 16-bit elements (+16 bytes of padding)
    static constexpr bool LO = has_lo(P);
    static constexpr int A_PER = BKC / WN;  // complex per thread per step (contiguous in K)
    static constexpr int A_TPR = WN;  // threads per A row
    static constexpr int B_PER = BKC / (2 * WM);
    static constexpr int B_TPR = 2 * WM;
    static constexpr int LDS_BYTES = (BM + BNR) * LD * 2 * (LO ? 2 : 1);
    static_assert(A_PER >= 4 && A_PER % 4 == 0, "A staging granularity");
    static_assert(B_PER >= 2 && B_PER % 2 == 0, "B staging granularity");
};

template<int N>
__device__ inline void store_words(uint32_t* dst, const uint32_t* w) {
    if constexpr (N % 4 == 0) {
#pragma unroll
        for (int i = 0; i < N / 4; i++)
            reinterpret_cast<uint4*>(dst)[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < N / 2; i++) reinterpret_cast<uint2*>(dst)[i] = make_uint2(w[2 * i], w[2 * i + 1]);
    }
}

template<int P, int WM, int WN, int BKC>
__global__ void __launch_bounds__(64 * WM * WN) contraction_mfma(
    const cfloat* __restrict__ A,
    const cfloat* __restrict__ B,
    int M, int N, int K, int64_t lda, int64_t ldb,
    int k_chunk, int tiles_m, int tiles_n, int splits,
    const float* __restrict__ row_scale_a,  // FP16: per-row power-of-two scales (else nullptr)
    const float* __restrict__ row_scale_b,
    float2* __restrict__ partial,
    cfloat* __restrict__ direct, int nsub, int64_t out_outer) {
    using Tl = MfmaTile<P, WM, WN, BKC>;
    using T = typename Half16<P>::T;
    using V = typename Half16<P>::V;
    constexpr bool LO = Tl::LO;
    constexpr int LD = Tl::LD;

    __shared__ __attribute__((aligned(16))) T sA_hi[Tl::BM * LD];
    __shared__ __attribute__((aligned(16))) T sB_hi[Tl::BNR * LD];
    __shared__ __attribute__((aligned(16))) T sA_lo[LO ? Tl::BM * LD : 8];
    __shared__ __attribute__((aligned(16))) T sB_lo[LO ? Tl::BNR * LD : 8];

    const int total = tiles_m * tiles_n * splits;
    const int lid = xcd_logical_id(blockIdx.x, total);
    if (lid >= total) return;
    const int tn = lid % tiles_n;
    const int tm = (lid / tiles_n) % tiles_m;
    const int ks = lid / (tiles_n * tiles_m);
    const int m0 = tm * Tl::BM, n0 = tn * Tl::BNC;
    const int kb = ks * k_chunk;
    const int ke = min(K, kb + k_chunk);

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WN, wn = wave % WN;

    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[i][j][e] = 0.0f;

    const int a_row = tid / Tl::A_TPR, a_k = (tid % Tl::A_TPR) * Tl::A_PER;
    const int b_row = tid / Tl::B_TPR, b_k = (tid % Tl::B_TPR) * Tl::B_PER;
    const int am = m0 + a_row, bn = n0 + b_row;
    const float sa = (row_scale_a && am < M) ? row_scale_a[am] : 1.0f;
    const float sb = (row_scale_b && bn < N) ? row_scale_b[bn] : 1.0f;
    const cfloat* a_src = A + int64_t(am) * lda;
    const cfloat* b_src = B + int64_t(bn) * ldb;
    // 16-byte vector loads need 16-byte aligned rows
    const bool vec_a = ((reinterpret_cast<uintptr_t>(A) & 15) == 0) && (lda % 2 == 0);
    const bool vec_b = ((reinterpret_cast<uintptr_t>(B) & 15) == 0) && (ldb % 2 == 0);

    // register prefetch of one K step
    cfloat ra[Tl::A_PER], rb[Tl::B_PER];
    auto gload = [&](int k0) {
        const int ka = k0 + a_k;
        if (vec_a && am < M && ka + Tl::A_PER <= ke) {
#pragma unroll
            for (int i = 0; i < Tl::A_PER / 2; i++) {
                const float4 v = *reinterpret_cast<const float4*>(a_src + ka + 2 * i);
                ra[2 * i] = {v.x, v.y};
                ra[2 * i + 1] = {v.z, v.w};
            }
        } else {
#pragma unroll
            for (int i = 0; i < Tl::A_PER; i++) ra[i] = (am < M && ka + i < ke) ? a_src[ka + i] : cfloat {0, 0};
        }
        const int kb_ = k0 + b_k;
        if (vec_b && bn < N && kb_ + Tl::B_PER <= ke) {
#pragma unroll
            for (int i = 0; i < Tl::B_PER / 2; i++) {
                const float4 v = *reinterpret_cast<const float4*>(b_src + kb_ + 2 * i);
                rb[2 * i] = {v.x, v.y};
                rb[2 * i + 1] = {v.z, v.w};
            }
        } else {
#pragma unroll
            for (int i = 0; i < Tl::B_PER; i++) rb[i] = (bn < N && kb_ + i < ke) ? b_src[kb_ + i] : cfloat {0, 0};
        }
    };

    if (kb < ke) gload(kb);
    for (int k0 = kb; k0 < ke; k0 += BKC) {
        // ---- stage A: (re, im) pairs packed hi / lo
        {
            uint32_t h[Tl::A_PER], l[Tl::A_PER];
#pragma unroll
            for (int i = 0; i < Tl::A_PER; i++) split_pair<P>(ra[i].re * sa, ra[i].im * sa, h[i], l[i]);
            store_words<Tl::A_PER>(reinterpret_cast<uint32_t*>(sA_hi + a_row * LD + 2 * a_k), h);
            if constexpr (LO) store_words<Tl::A_PER>(reinterpret_cast<uint32_t*>(sA_lo + a_row * LD + 2 * a_k), l);
        }
        // ---- stage B: conj row (re, -im) and swap row (im, re)
        {
            uint32_t ch[Tl::B_PER], sh[Tl::B_PER], cl[Tl::B_PER], sl[Tl::B_PER];
#pragma unroll
            for (int i = 0; i < Tl::B_PER; i++) {
                uint32_t h, l;
                split_pair<P>(rb[i].re * sb, rb[i].im * sb, h, l);
                ch[i] = h ^ 0x80000000u;  // negate the imaginary half
                sh[i] = __builtin_amdgcn_alignbit(h, h, 16);  // swap halves
                cl[i] = l ^ 0x80000000u;
                sl[i] = __builtin_amdgcn_alignbit(l, l, 16);
            }
            store_words<Tl::B_PER>(reinterpret_cast<uint32_t*>(sB_hi + (2 * b_row) * LD + 2 * b_k), ch);
            store_words<Tl::B_PER>(reinterpret_cast<uint32_t*>(sB_hi + (2 * b_row + 1) * LD + 2 * b_k), sh);
            if constexpr (LO) {
                store_words<Tl::B_PER>(reinterpret_cast<uint32_t*>(sB_lo + (2 * b_row) * LD + 2 * b_k), cl);
                store_words<Tl::B_PER>(reinterpret_cast<uint32_t*>(sB_lo + (2 * b_row + 1) * LD + 2 * b_k), sl);
            }
        }
        __syncthreads();
        if (k0 + BKC < ke) gload(k0 + BKC);  // in flight during the MFMAs below

#pragma unroll
        for (int kk = 0; kk < BKC / 8; kk++) {
            const int koff = kk * 16 + (lane >> 5) * 8;
            V ah[2], al[2], bh[2], bl[2];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const int r = wm * 64 + i * 32 + (lane & 31);
                ah[i] = *reinterpret_cast<const V*>(sA_hi + r * LD + koff);
                if constexpr (LO) al[i] = *reinterpret_cast<const V*>(sA_lo + r * LD + koff);
            }
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int c = wn * 64 + j * 32 + (lane & 31);
                bh[j] = *reinterpret_cast<const V*>(sB_hi + c * LD + koff);
                if constexpr (LO) bl[j] = *reinterpret_cast<const V*>(sB_lo + c * LD + koff);
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
                    reinterpret_cast<float*>(partial)[2 * ((int64_t(ks) * M + m) * N + n) + part] = v;
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
        if (m > 0.0f && isfinite(m)) frexpf(m, &e);  // m = f * 2^e, f in [0.5, 1)
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

namespace {

using namespace kernels;

int cu_count(int device) {
    static int cached[64] = {0};
    if (device >= 0 && device < 64 && cached[device]) return cached[device];
    hipDeviceProp_t prop;
    COMPAS_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    if (device >= 0 && device < 64) cached[device] = prop.multiProcessorCount;
    return prop.multiProcessorCount;
}

struct Plan {
    int tiles_m, tiles_n, splits, k_chunk, threads;
};

// Workgroups of one variant that fit on a CU at once (LDS and registers).
template<int P, int WM, int WN, int BKC>
int resident_per_cu() {
    static const int cached = [] {
        int n = 0;
        COMPAS_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n, contraction_mfma<P, WM, WN, BKC>, MfmaTile<P, WM, WN, BKC>::THREADS, 0));
        return std::max(1, n);
    }();
    return cached;
}

// Split-K plan.  All workgroups do the same amount of work, so the launch
// takes ceil(workgroups / resident slots) workgroup lifetimes: a grid just
// above a multiple of the slots leaves a nearly empty last round.  The split
// count minimises rounds x K-chunk (the critical path), with a small charge
// per split for the partial-sum traffic.
template<int P, int WM, int WN, int BKC>
Plan make_plan(const CompasContext& ctx, const ContractionShape& s) {
    using Tl = MfmaTile<P, WM, WN, BKC>;
    Plan p;
    p.threads = Tl::THREADS;
    p.tiles_m = int(dev::div_ceil(s.M, Tl::BM));
    p.tiles_n = int(dev::div_ceil(s.N, Tl::BNC));
    const int64_t tiles = int64_t(p.tiles_m) * p.tiles_n;
    const int64_t slots = int64_t(resident_per_cu<P, WM, WN, BKC>()) * cu_count(ctx.device());
    const int max_splits = std::max(1, std::min(64, s.K / (8 * BKC)));
    double best_cost = 1e300;
    p.splits = 1;
    p.k_chunk = std::max(BKC, int(dev::div_ceil(s.K, BKC)) * BKC);
    for (int want = 1; want <= max_splits; want++) {
        const int kc = std::max(BKC, int(dev::div_ceil(dev::div_ceil(s.K, want), BKC)) * BKC);
        const int splits = std::max(1, int(dev::div_ceil(s.K, kc)));
        const int64_t rounds = dev::div_ceil(tiles * splits, slots);
        const double cost = double(rounds) * kc * (1.0 + 0.01 * (splits - 1));
        if (cost < best_cost * (1.0 - 1e-9)) {
            best_cost = cost;
            p.splits = splits;
            p.k_chunk = kc;
        }
    }
    return p;
}

template<int P, int WM, int WN, int BKC>
void launch_variant(const CompasContext& ctx, const cfloat* A, const cfloat* B, cfloat* out,
                    const ContractionShape& s, const float* sa, const float* sb) {
    const Plan p = make_plan<P, WM, WN, BKC>(ctx, s);
    const int total = p.tiles_m * p.tiles_n * p.splits;
    const bool direct = (p.splits == 1 && !s.accumulate);
    Array<float2, 1> partial;
    if (!direct) partial = ctx.empty<float2, 1>({int64_t(p.splits) * s.M * s.N});
    {
        KernelScope ks(ctx, "contraction_mfma");
        contraction_mfma<P, WM, WN, BKC><<<total, p.threads, 0, ctx.stream()>>>(
            A, B, s.M, s.N, s.K, s.lda, s.ldb, p.k_chunk, p.tiles_m, p.tiles_n, p.splits, sa, sb,
            direct ? nullptr : partial.data(), direct ? out : nullptr, s.nsub, s.out_outer);
        check_launch("contraction_mfma");
    }
    if (!direct) {
        KernelScope ks(ctx, "contraction_reduce");
        const int64_t n = int64_t(s.M) * s.N;
        contraction_reduce_mf<<<dev::div_ceil(n, 256), 256, 0, ctx.stream()>>>(
            partial.data(), p.splits, s.M, s.N, out, s.nsub, s.out_outer, s.accumulate);
        check_launch("contraction_reduce");
    }
}

using VariantFn = void (*)(const CompasContext&, const cfloat*, const cfloat*, cfloat*, const ContractionShape&,
                           const float*, const float*);

struct Variant {
    const char* name;
    VariantFn fn[NUM_P];  // indexed by P
};

#define COMPAS_MFMA_VARIANT(WM, WN, BKC)                                                          \
    {#WM "x" #WN "x" #BKC,                                                                        \
     {&launch_variant<P_BF16X3, WM, WN, BKC>, &launch_variant<P_FP16X3, WM, WN, BKC>,             \
      &launch_variant<P_BF16X1, WM, WN, BKC>, &launch_variant<P_FP16X1, WM, WN, BKC>}}

const Variant VARIANTS[] = {
    COMPAS_MFMA_VARIANT(2, 2, 32),
    COMPAS_MFMA_VARIANT(4, 2, 32),
    COMPAS_MFMA_VARIANT(2, 4, 32),
    COMPAS_MFMA_VARIANT(4, 2, 16),
    COMPAS_MFMA_VARIANT(2, 4, 16),
    COMPAS_MFMA_VARIANT(4, 4, 16),
};
constexpr int NUM_VARIANTS = int(sizeof(VARIANTS) / sizeof(VARIANTS[0]));
constexpr int DEFAULT_VARIANT = 4;  // 2x4x16: fastest at 1120 x 896 x 50176 on MI355X

std::mutex g_tune_mutex;
std::map<std::tuple<int, int, int, int64_t, int64_t, int, int, bool>, int> g_tuned;
int g_forced_variant = -1;

bool autotune_enabled() {
    const char* e = std::getenv("COMPAS_AUTOTUNE");
    return !(e && e[0] == '0');
}

int pick_variant(const CompasContext& ctx, const cfloat* A, const cfloat* B, const ContractionShape& s, int P,
                 const float* sa, const float* sb) {
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        if (g_forced_variant >= 0) return g_forced_variant;
    }
    const auto key = std::make_tuple(ctx.device(), s.M, s.N, s.lda, s.ldb, s.K, P, s.accumulate);
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) return it->second;
    }
    if (const char* e = std::getenv("COMPAS_CONTRACTION_VARIANT")) {
        const int v = std::atoi(e);
        if (v >= 0 && v < NUM_VARIANTS) return v;
    }
    if (!autotune_enabled() || int64_t(s.M) * s.N * s.K < (int64_t(1) << 24)) return DEFAULT_VARIANT;
    // time every variant into a scratch output (the caller's output is not touched)
    auto scratch = ctx.empty<cfloat, 1>({int64_t(s.M) * s.N});
    ContractionShape t = s;
    t.nsub = s.N;
    t.out_outer = 0;
    t.accumulate = false;
    hipEvent_t e0, e1;
    COMPAS_HIP_CHECK(hipEventCreate(&e0));
    COMPAS_HIP_CHECK(hipEventCreate(&e1));
    int best = DEFAULT_VARIANT;
    float best_ms = 1e30f;
    for (int v = 0; v < NUM_VARIANTS; v++) {
        VARIANTS[v].fn[P](ctx, A, B, scratch.data(), t, sa, sb);  // warm-up
        COMPAS_HIP_CHECK(hipEventRecord(e0, ctx.stream()));
        for (int r = 0; r < 2; r++) VARIANTS[v].fn[P](ctx, A, B, scratch.data(), t, sa, sb);
        COMPAS_HIP_CHECK(hipEventRecord(e1, ctx.stream()));
        COMPAS_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        COMPAS_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_ms) {
            best_ms = ms;
            best = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tuned[key] = best;
    return best;
}

}  // namespace

void set_contraction_variant(int v) {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_forced_variant = (v >= 0 && v < NUM_VARIANTS) ? v : -1;
}

int contraction_variant_count() {
    return NUM_VARIANTS;
}

std::string contraction_tuning_report() {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    std::string out;
    for (auto& kv : g_tuned) {
        char line[256];
        static const char* const names[NUM_P] = {"bf16x3", "fp16x3", "bf16x1", "fp16x1"};
        std::snprintf(line, sizeof(line), "M=%d N=%d K=%d prec=%s -> %s\n", std::get<1>(kv.first),
                      std::get<2>(kv.first), std::get<5>(kv.first), names[std::get<6>(kv.first)],
                      VARIANTS[kv.second].name);
        out += line;
    }
    return out;
}

void complex_contraction_mfma(
    const CompasContext& ctx,
    const cfloat* A,
    const cfloat* B,
    cfloat* out,
    ContractionShape s,
    ContractionPrecision precision,
    bool* handled) {
    int P;
    switch (precision) {
        case ContractionPrecision::BF16x3: P = P_BF16X3; break;
        case ContractionPrecision::FP16x3: P = P_FP16X3; break;
        case ContractionPrecision::BF16x1: P = P_BF16X1; break;
        case ContractionPrecision::FP16x1: P = P_FP16X1; break;
        default: *handled = false; return;
    }
    *handled = true;
    Array<float, 1> scale_a, scale_b;
    const float* sa = nullptr;
    const float* sb = nullptr;
    if (P == P_FP16X3 || P == P_FP16X1) {
        scale_a = ctx.empty<float, 1>({s.M});
        scale_b = ctx.empty<float, 1>({s.N});
        KernelScope ks(ctx, "row_pow2_scale");
        row_pow2_scale<<<s.M, 256, 0, ctx.stream()>>>(A, s.M, s.K, s.lda, scale_a.data());
        row_pow2_scale<<<s.N, 256, 0, ctx.stream()>>>(B, s.N, s.K, s.ldb, scale_b.data());
        check_launch("row_pow2_scale");
        sa = scale_a.data();
        sb = scale_b.data();
    }
    const int v = pick_variant(ctx, A, B, s, P, sa, sb);
    VARIANTS[v].fn[P](ctx, A, B, out, s, sa, sb);
}

}  // namespace compas
