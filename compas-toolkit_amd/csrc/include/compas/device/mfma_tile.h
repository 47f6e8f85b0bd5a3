#pragma once
// MFMA building blocks of the complex contractions on gfx950 (wave64).
//
// Complex-to-real mapping.  With A stored interleaved (re, im) along K, row
// m of A is a real vector of length 2K.  For each complex column n the real
// GEMM uses two B rows
//   row 2n   = (b_re, -b_im) pairs  ->  D[m][2n]   = Re(sum a b)
//   row 2n+1 = (b_im,  b_re) pairs  ->  D[m][2n+1] = Im(sum a b)
// so one real GEMM (M x 2N x 2K) produces interleaved complex output.  LDS
// holds each complex column once, as (b_re, b_im) pairs; the lanes feeding
// real column 2n and 2n+1 read the same LDS row (a broadcast) and form their
// row in registers with one v_perm_b32 (half swap) and one v_xor_b32 (sign
// of the imaginary half) per packed word.
//
// Precision.  Every FP32 operand x is split into hi = round(x) and
// lo = round(x - hi) in BF16 (or FP16 after a per-row power-of-two scale that
// keeps both parts in the normal range); the product uses
// hi*hi + hi*lo + lo*hi with FP32 accumulation in the MFMA, which recovers
// ~16 (BF16) or ~22 (FP16) bits per product.  BF16x1 and FP16x1 use hi*hi
// only.
//
// Tiling.  A workgroup of WM x WN waves owns 64 WM rows x 64 WN real columns
// (32 WN complex) and walks a K range in steps of BKC complex elements; each
// wave computes 64 x 64 with 2 x 2 v_mfma_f32_32x32x16 blocks.  The next K
// step is prefetched into registers while the MFMAs of the current one run.
// Operands are converted to packed 16-bit pairs on the way into LDS (one
// v_cvt_pk per pair, 128-bit stores).
//
// Accumulator layout of v_mfma_f32_32x32x16 (block (i, j) of a wave, lane l,
// element e): row 32 i + 8 (e / 4) + 4 (l / 32) + e % 4, real column
// 32 j + l % 32.

#include <cstdint>

#include "compas/device/common.h"

namespace compas {
namespace mfma {

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

// split_pair of a complex value scaled by s; the per-row power-of-two scales
// exist only for the FP16 precisions (BF16 has FP32's exponent range), so
// the BF16 forms skip the two multiplies
template<int P>
__device__ inline void split_scaled(cfloat v, float s, uint32_t& hi, uint32_t& lo) {
    if constexpr (is_fp16(P)) {
        split_pair<P>(v.re * s, v.im * s, hi, lo);
    } else {
        split_pair<P>(v.re, v.im, hi, lo);
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

// WM x WN waves; each wave owns TM x TN 32 x 32 output tiles (TM row tiles
// of A, TN tiles of 16 complex B columns); BKC complex K per step.
template<int P, int WM, int WN, int BKC, int TM = 2, int TN = 2>
struct MfmaTile {
    static constexpr int THREADS = 64 * WM * WN;
    static constexpr int BM = 32 * TM * WM;  // A rows
    static constexpr int BNC = 16 * TN * WN;  // complex B columns
    static constexpr int LD = 2 * BKC;  // LDS row stride, 16-bit elements (no padding: rows are swizzled)
    static constexpr bool LO = has_lo(P);
    static constexpr int A_TPR = THREADS / BM;  // threads per A row
    static constexpr int A_PER = BKC / A_TPR;  // complex per thread per step (contiguous in K)
    static constexpr int B_TPR = THREADS / BNC;
    static constexpr int B_PER = BKC / B_TPR;
    static constexpr int LDS_BYTES = (BM + BNC) * LD * 2 * (LO ? 2 : 1);
    static_assert(THREADS % BM == 0 && THREADS % BNC == 0, "staging rows");
    static_assert(A_PER >= 4 && A_PER % 4 == 0, "A staging granularity");
    static_assert(B_PER >= 2 && B_PER % 2 == 0, "B staging granularity");
    static_assert(BKC == 16 || BKC == 32, "row swizzle");

    // Offset, in 32-bit words, of word w (the packed pair of complex K index
    // w) of LDS row r.  The 16-byte chunks of a row are permuted by the Gray
    // code of the row's 128-byte line, so that a fragment read (one chunk of
    // each of 8 or 16 consecutive rows) and a staging write (whole chunks of
    // 2 or 4 consecutive rows) fall on distinct banks, whether the LDS serves
    // 128 or 256 bytes per cycle.
    __device__ static int word(int r, int w) {
        const int line = r / (32 / BKC);
        const int f = (line ^ (line >> 1)) & (BKC / 4 - 1);
        return r * BKC + (((w >> 2) ^ f) << 2) + (w & 3);
    }
};

// LDS of one workgroup: packed hi (and lo) parts of the A and B tiles
template<int P, int WM, int WN, int BKC, int TM = 2, int TN = 2>
struct MfmaSmem {
    using Tl = MfmaTile<P, WM, WN, BKC, TM, TN>;
    using T = typename Half16<P>::T;
    __attribute__((aligned(16))) T a_hi[Tl::BM * Tl::LD];
    __attribute__((aligned(16))) T b_hi[Tl::BNC * Tl::LD];  // one row per complex column
    __attribute__((aligned(16))) T a_lo[Tl::LO ? Tl::BM * Tl::LD : 8];
    __attribute__((aligned(16))) T b_lo[Tl::LO ? Tl::BNC * Tl::LD : 8];
};

// Stores the N packed words w[] of LDS row r from word w0 on (whole chunks
// when N is a multiple of 4, else pairs within one chunk).
template<typename Tl, int N>
__device__ inline void store_row(void* plane, int r, int w0, const uint32_t* w) {
    uint32_t* p = reinterpret_cast<uint32_t*>(plane);
    if constexpr (N % 4 == 0) {
#pragma unroll
        for (int i = 0; i < N / 4; i++)
            *reinterpret_cast<uint4*>(p + Tl::word(r, w0 + 4 * i)) = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < N / 2; i++) *reinterpret_cast<uint2*>(p + Tl::word(r, w0 + 2 * i)) = make_uint2(w[2 * i], w[2 * i + 1]);
    }
}

// A 16-byte fragment of LDS row r at 16-bit offset k (a multiple of 8)
template<typename Tl, typename V, typename T>
__device__ inline V load_frag(const T* plane, int r, int k) {
    return *reinterpret_cast<const V*>(reinterpret_cast<const uint32_t*>(plane) + Tl::word(r, k >> 1));
}

// The B operand row of a lane from the packed (re, im) words of its complex
// column: half swap by byte permutation, then the sign flip.
template<typename V>
__device__ inline V b_row_form_t(uint4 w, uint32_t sel, uint32_t flip) {
    w.x = __builtin_amdgcn_perm(w.x, w.x, sel) ^ flip;
    w.y = __builtin_amdgcn_perm(w.y, w.y, sel) ^ flip;
    w.z = __builtin_amdgcn_perm(w.z, w.z, sel) ^ flip;
    w.w = __builtin_amdgcn_perm(w.w, w.w, sel) ^ flip;
    return __builtin_bit_cast(V, w);
}

// Row of the A tile that thread `tid` stages, and row of the B tile (complex
// column).  Each thread stages one A row and one B row for the whole K range.
template<int P, int WM, int WN, int BKC, int TM = 2, int TN = 2>
__device__ inline int staged_a_row(int tid) {
    return tid / MfmaTile<P, WM, WN, BKC, TM, TN>::A_TPR;
}
template<int P, int WM, int WN, int BKC, int TM = 2, int TN = 2>
__device__ inline int staged_b_row(int tid) {
    return tid / MfmaTile<P, WM, WN, BKC, TM, TN>::B_TPR;
}

// The MFMAs of one staged K step: each wave multiplies its TM row tiles of
// A by its TN column tiles of B out of LDS.
template<int P, int WM, int WN, int BKC, int TM, int TN>
__device__ __forceinline__ void mfma_tile_step(f32x16 (&acc)[TM][TN], const MfmaSmem<P, WM, WN, BKC, TM, TN>& sm,
                                               int lane, int wm, int wn, uint32_t sel, uint32_t flip) {
    using Tl = MfmaTile<P, WM, WN, BKC, TM, TN>;
    using V = typename Half16<P>::V;
    constexpr bool LO = Tl::LO;
#pragma unroll
    for (int kk = 0; kk < BKC / 8; kk++) {
        const int koff = kk * 16 + (lane >> 5) * 8;
        V ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
        for (int i = 0; i < TM; i++) {
            const int r = wm * (32 * TM) + i * 32 + (lane & 31);
            ah[i] = load_frag<Tl, V>(sm.a_hi, r, koff);
            if constexpr (LO) al[i] = load_frag<Tl, V>(sm.a_lo, r, koff);
        }
#pragma unroll
        for (int j = 0; j < TN; j++) {
            const int c = wn * (16 * TN) + j * 16 + (lane & 31) / 2;  // complex column of this lane
            bh[j] = b_row_form_t<V>(load_frag<Tl, uint4>(sm.b_hi, c, koff), sel, flip);
            if constexpr (LO) bl[j] = b_row_form_t<V>(load_frag<Tl, uint4>(sm.b_lo, c, koff), sel, flip);
        }
        // hi*hi, then hi*lo, then lo*hi over the accumulators: each
        // accumulator's products stay in that order, and TM x TN independent
        // MFMAs separate two that chain on one accumulator
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int j = 0; j < TN; j++) acc[i][j] = mfma<P>(ah[i], bh[j], acc[i][j]);
        if constexpr (LO) {
#pragma unroll
            for (int i = 0; i < TM; i++)
#pragma unroll
                for (int j = 0; j < TN; j++) acc[i][j] = mfma<P>(ah[i], bl[j], acc[i][j]);
#pragma unroll
            for (int i = 0; i < TM; i++)
#pragma unroll
                for (int j = 0; j < TN; j++) acc[i][j] = mfma<P>(al[i], bh[j], acc[i][j]);
        }
    }
}

// acc += A_tile[.., kb:ke] * B_tile[.., kb:ke]^T for the workgroup's tile.
// a_src / b_src: start of the calling thread's staged A row / B row (complex,
// K-contiguous), nullptr for a row outside the operand (staged as zeros);
// sa / sb scale the row's values before the split; vec_*: the row may be read
// with 16-byte loads.  All threads of the workgroup must call this together.
template<int P, int WM, int WN, int BKC, int TM = 2, int TN = 2>
__device__ inline void mfma_accumulate(
    f32x16 (&acc)[TM][TN], MfmaSmem<P, WM, WN, BKC, TM, TN>& sm,
    const cfloat* __restrict__ a_src, float sa, bool vec_a,
    const cfloat* __restrict__ b_src, float sb, bool vec_b,
    int kb, int ke) {
    using Tl = MfmaTile<P, WM, WN, BKC, TM, TN>;
    constexpr bool LO = Tl::LO;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const int a_row = tid / Tl::A_TPR, a_k = (tid % Tl::A_TPR) * Tl::A_PER;
    const int b_row = tid / Tl::B_TPR, b_k = (tid % Tl::B_TPR) * Tl::B_PER;
    // even lanes feed Re columns (b_re, -b_im), odd lanes Im columns (b_im, b_re)
    const uint32_t sel = (lane & 1) ? 0x01000302u : 0x03020100u;
    const uint32_t flip = (lane & 1) ? 0u : 0x80000000u;

    // register prefetch of one K step
    cfloat ra[Tl::A_PER], rb[Tl::B_PER];
    auto gload = [&](int k0) {
        const int ka = k0 + a_k;
        if (vec_a && a_src && ka + Tl::A_PER <= ke) {
#pragma unroll
            for (int i = 0; i < Tl::A_PER / 2; i++) {
                const float4 v = *reinterpret_cast<const float4*>(a_src + ka + 2 * i);
                ra[2 * i] = {v.x, v.y};
                ra[2 * i + 1] = {v.z, v.w};
            }
        } else {
#pragma unroll
            for (int i = 0; i < Tl::A_PER; i++) ra[i] = (a_src && ka + i < ke) ? a_src[ka + i] : cfloat {0, 0};
        }
        const int kb_ = k0 + b_k;
        if (vec_b && b_src && kb_ + Tl::B_PER <= ke) {
#pragma unroll
            for (int i = 0; i < Tl::B_PER / 2; i++) {
                const float4 v = *reinterpret_cast<const float4*>(b_src + kb_ + 2 * i);
                rb[2 * i] = {v.x, v.y};
                rb[2 * i + 1] = {v.z, v.w};
            }
        } else {
#pragma unroll
            for (int i = 0; i < Tl::B_PER; i++) rb[i] = (b_src && kb_ + i < ke) ? b_src[kb_ + i] : cfloat {0, 0};
        }
    };

    if (kb < ke) gload(kb);
    for (int k0 = kb; k0 < ke; k0 += BKC) {
        // ---- stage A: (re, im) pairs packed hi / lo
        {
            uint32_t h[Tl::A_PER], l[Tl::A_PER];
#pragma unroll
            for (int i = 0; i < Tl::A_PER; i++) split_scaled<P>(ra[i], sa, h[i], l[i]);
            store_row<Tl, Tl::A_PER>(sm.a_hi, a_row, a_k, h);
            if constexpr (LO) store_row<Tl, Tl::A_PER>(sm.a_lo, a_row, a_k, l);
        }
        // ---- stage B: (re, im) pairs, one row per complex column
        {
            uint32_t h[Tl::B_PER], l[Tl::B_PER];
#pragma unroll
            for (int i = 0; i < Tl::B_PER; i++) split_scaled<P>(rb[i], sb, h[i], l[i]);
            store_row<Tl, Tl::B_PER>(sm.b_hi, b_row, b_k, h);
            if constexpr (LO) store_row<Tl, Tl::B_PER>(sm.b_lo, b_row, b_k, l);
        }
        __syncthreads();
        if (k0 + BKC < ke) gload(k0 + BKC);  // in flight during the MFMAs below
        mfma_tile_step<P, WM, WN, BKC, TM, TN>(acc, sm, lane, wm, wn, sel, flip);
        __syncthreads();
    }
}

// The same product with two LDS buffers: the split and store of the next K
// step go to the other buffer while this step's MFMAs read the current one,
// so each step needs one barrier instead of two and the staging overlaps the
// matrix pipe.  Twice the LDS of mfma_accumulate.
template<int P, int WM, int WN, int BKC, int TM = 2, int TN = 2>
__device__ inline void mfma_accumulate_db(
    f32x16 (&acc)[TM][TN], MfmaSmem<P, WM, WN, BKC, TM, TN> (&sm)[2],
    const cfloat* __restrict__ a_src, float sa, bool vec_a,
    const cfloat* __restrict__ b_src, float sb, bool vec_b,
    int kb, int ke) {
    using Tl = MfmaTile<P, WM, WN, BKC, TM, TN>;
    constexpr bool LO = Tl::LO;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const int a_row = tid / Tl::A_TPR, a_k = (tid % Tl::A_TPR) * Tl::A_PER;
    const int b_row = tid / Tl::B_TPR, b_k = (tid % Tl::B_TPR) * Tl::B_PER;
    const uint32_t sel = (lane & 1) ? 0x01000302u : 0x03020100u;
    const uint32_t flip = (lane & 1) ? 0u : 0x80000000u;

    cfloat ra[Tl::A_PER], rb[Tl::B_PER];
    auto gload = [&](int k0) {
        const int ka = k0 + a_k;
        if (vec_a && a_src && ka + Tl::A_PER <= ke) {
#pragma unroll
            for (int i = 0; i < Tl::A_PER / 2; i++) {
                const float4 v = *reinterpret_cast<const float4*>(a_src + ka + 2 * i);
                ra[2 * i] = {v.x, v.y};
                ra[2 * i + 1] = {v.z, v.w};
            }
        } else {
#pragma unroll
            for (int i = 0; i < Tl::A_PER; i++) ra[i] = (a_src && ka + i < ke) ? a_src[ka + i] : cfloat {0, 0};
        }
        const int kb_ = k0 + b_k;
        if (vec_b && b_src && kb_ + Tl::B_PER <= ke) {
#pragma unroll
            for (int i = 0; i < Tl::B_PER / 2; i++) {
                const float4 v = *reinterpret_cast<const float4*>(b_src + kb_ + 2 * i);
                rb[2 * i] = {v.x, v.y};
                rb[2 * i + 1] = {v.z, v.w};
            }
        } else {
#pragma unroll
            for (int i = 0; i < Tl::B_PER; i++) rb[i] = (b_src && kb_ + i < ke) ? b_src[kb_ + i] : cfloat {0, 0};
        }
    };
    auto stage = [&](MfmaSmem<P, WM, WN, BKC, TM, TN>& buf) {
        uint32_t h[Tl::A_PER], l[Tl::A_PER];
#pragma unroll
        for (int i = 0; i < Tl::A_PER; i++) split_scaled<P>(ra[i], sa, h[i], l[i]);
        store_row<Tl, Tl::A_PER>(buf.a_hi, a_row, a_k, h);
        if constexpr (LO) store_row<Tl, Tl::A_PER>(buf.a_lo, a_row, a_k, l);
        uint32_t hb[Tl::B_PER], lb[Tl::B_PER];
#pragma unroll
        for (int i = 0; i < Tl::B_PER; i++) split_scaled<P>(rb[i], sb, hb[i], lb[i]);
        store_row<Tl, Tl::B_PER>(buf.b_hi, b_row, b_k, hb);
        if constexpr (LO) store_row<Tl, Tl::B_PER>(buf.b_lo, b_row, b_k, lb);
    };
    auto compute = [&](const MfmaSmem<P, WM, WN, BKC, TM, TN>& buf) {
        mfma_tile_step<P, WM, WN, BKC, TM, TN>(acc, buf, lane, wm, wn, sel, flip);
    };

    if (kb >= ke) return;  // uniform over the workgroup
    gload(kb);
    stage(sm[0]);
    __syncthreads();
    int cur = 0;
    for (int k0 = kb; k0 < ke; k0 += BKC) {
        const bool more = k0 + BKC < ke;  // uniform over the workgroup
        if (more) gload(k0 + BKC);  // in flight during the MFMAs below
        compute(sm[cur]);
        if (more) stage(sm[cur ^ 1]);  // the other buffer: last read before the previous barrier
        __syncthreads();
        cur ^= 1;
    }
}

// A B operand generated while staging instead of read from memory: complex
// column n = c ns + s (coil c, sample s) is, for voxel v,
//   B[n][v] = C[c][v] T[s / 4][v] q[v]^(s % 4),
// with T the column factor of every fourth sample and q the factor between
// consecutive samples (the Cartesian coil-weighted sample factors).  ns is a
// multiple of 4, so the four columns a thread generates share c and s / 4.
struct GenB {
    const cfloat* C;  // [nc][nv]
    const cfloat* T;  // [ns / 4][nv]
    const cfloat* q;  // [nv]
    int ns, nv, N;  // N = nc ns complex columns
};

// mfma_accumulate with the B tile generated: thread t stages voxel t % BKC of
// the K step for the four consecutive columns 4 gc .. 4 gc + 3 of the tile,
// gc = t / BKC (three loads, four complex multiplies), where the loaded form
// stages one column for BKC / B_TPR voxels.  Column 4 gc + i goes to LDS row
// gc + G i (G = THREADS / BKC), so that one store instruction covers
// consecutive rows: the caller's epilogue maps LDS row rho back to column
// 4 (rho % G) + rho / G.  The A staging and the MFMA loop are those of
// mfma_accumulate.  n0: first complex column of the tile.  FULL: every row
// and column of the tile exists, the K range is whole steps and the A rows
// take 16-byte loads, so the staging loads carry no bounds tests.
template<int P, int WM, int WN, int BKC, int TM = 2, int TN = 2, bool FULL = false>
__device__ inline void mfma_accumulate_genb(
    f32x16 (&acc)[TM][TN], MfmaSmem<P, WM, WN, BKC, TM, TN>& sm,
    const cfloat* __restrict__ a_src, float sa, bool vec_a,
    const GenB& g, int n0, int kb, int ke) {
    using Tl = MfmaTile<P, WM, WN, BKC, TM, TN>;
    constexpr bool LO = Tl::LO;
    constexpr int G = Tl::THREADS / BKC;  // column groups of four
    static_assert(4 * G == Tl::BNC, "four generated columns per thread");
    static_assert(!is_fp16(P), "generated operands carry no per-row scales");

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const int a_row = tid / Tl::A_TPR, a_k = (tid % Tl::A_TPR) * Tl::A_PER;
    const uint32_t sel = (lane & 1) ? 0x01000302u : 0x03020100u;
    const uint32_t flip = (lane & 1) ? 0u : 0x80000000u;

    // generated columns of this thread: 4 gc .. 4 gc + 3 of the tile, voxel gk of each K step
    const int gk = tid % BKC, gc = tid / BKC;
    const int n = n0 + 4 * gc;
    const bool col_ok = n < g.N;
    const int coil = col_ok ? n / g.ns : 0, quad = col_ok ? (n % g.ns) / 4 : 0;
    const cfloat* __restrict__ gC = g.C + int64_t(coil) * g.nv;
    const cfloat* __restrict__ gT = g.T + int64_t(quad) * g.nv;

    cfloat ra[Tl::A_PER];
    cfloat rC, rT, rq;
    auto gload = [&](int k0) {
        const int ka = k0 + a_k;
        if (FULL || (vec_a && a_src && ka + Tl::A_PER <= ke)) {
#pragma unroll
            for (int i = 0; i < Tl::A_PER / 2; i++) {
                const float4 v = *reinterpret_cast<const float4*>(a_src + ka + 2 * i);
                ra[2 * i] = {v.x, v.y};
                ra[2 * i + 1] = {v.z, v.w};
            }
        } else {
#pragma unroll
            for (int i = 0; i < Tl::A_PER; i++) ra[i] = (a_src && ka + i < ke) ? a_src[ka + i] : cfloat {0, 0};
        }
        const int v = k0 + gk;
        if (FULL || (col_ok && v < ke)) {
            rC = gC[v];
            rT = gT[v];
            rq = g.q[v];
        } else {
            rC = {0, 0};
            rT = {0, 0};
            rq = {0, 0};
        }
    };

    if (kb < ke) gload(kb);
    for (int k0 = kb; k0 < ke; k0 += BKC) {
        {
            uint32_t h[Tl::A_PER], l[Tl::A_PER];
#pragma unroll
            for (int i = 0; i < Tl::A_PER; i++) split_scaled<P>(ra[i], sa, h[i], l[i]);
            store_row<Tl, Tl::A_PER>(sm.a_hi, a_row, a_k, h);
            if constexpr (LO) store_row<Tl, Tl::A_PER>(sm.a_lo, a_row, a_k, l);
        }
        {
            cfloat w = dev::cmul(rC, rT);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                uint32_t h, l;
                split_pair<P>(w.re, w.im, h, l);
                const int row = gc + G * i;  // column 4 gc + i
                reinterpret_cast<uint32_t*>(sm.b_hi)[Tl::word(row, gk)] = h;
                if constexpr (LO) reinterpret_cast<uint32_t*>(sm.b_lo)[Tl::word(row, gk)] = l;
                if (i < 3) w = dev::cmul(w, rq);
            }
        }
        __syncthreads();
        if (k0 + BKC < ke) gload(k0 + BKC);  // in flight during the MFMAs below
        mfma_tile_step<P, WM, WN, BKC, TM, TN>(acc, sm, lane, wm, wn, sel, flip);
        __syncthreads();
    }
}

template<int TM, int TN>
__device__ inline void zero(f32x16 (&acc)[TM][TN]) {
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[i][j][e] = 0.0f;
}

}  // namespace mfma
}  // namespace compas
