// MFMA complex contraction on gfx950: C[m][n] = sum_k A[m][k] B[n][k].
//
// The tile machinery (complex-to-real mapping, hi/lo operand split, LDS
// staging, MFMA main loop) lives in compas/device/mfma_tile.h.  This file
// holds the plain contraction kernel, its split-K plan and reduction, and a
// small autotuner: the tile variants (WM, WN, BKC) are timed on the live shape
// on first use and the fastest is cached per shape (runtimes that opt out,
// such as the request server's slots, run the default tile on new shapes).  Workgroups are numbered
// so that those sharing A rows run on the same XCD (consecutive hardware ids
// are spread over the 8 XCDs), so the XCD's L2 serves the re-reads of A
// across column tiles.  K slices are reduced in FP64.

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "compas/device/common.h"
#include "compas/device/mfma_tile.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

using namespace mfma;

// Epilogue of a workgroup's tile: D[row][realcol], realcol = 2n + {re, im},
// to the output (direct, column n at (n / nsub) out_outer + m nsub + n % nsub)
// or to split ks of the partial sums.  Scales: FP16 per-row scales, else
// nullptr.  GPERM > 0: tile column rho holds column 4 (rho % GPERM) + rho /
// GPERM (the generated operand's store order).
template<int WM, int WN, int GPERM = 0, int TM, int TN>
__device__ inline void store_tile(const f32x16 (&acc)[TM][TN], int m0, int n0, int ks, int M, int N,
                                       const float* __restrict__ row_scale_a, const float* __restrict__ row_scale_b,
                                       float2* __restrict__ partial, cfloat* __restrict__ direct, int nsub,
                                       int64_t out_outer) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave / WN, wn = wave % WN;
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const int row = wm * (32 * TM) + i * 32 + (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3);
                const int rc = wn * (32 * TN) + j * 32 + (lane & 31);
                const int rho = rc >> 1, part = rc & 1;
                const int m = m0 + row, n = n0 + (GPERM ? 4 * (rho % GPERM) + rho / GPERM : rho);
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

template<int P, int WM, int WN, int BKC, bool DB, int TM, int TN>
__global__ void __launch_bounds__(64 * WM * WN) contraction_mfma(
    const cfloat* __restrict__ A,
    const cfloat* __restrict__ B,
    int M, int N, int K, int64_t lda, int64_t ldb,
    int k_chunk, int tiles_m, int tiles_n, int splits,
    const float* __restrict__ row_scale_a,  // FP16: per-row power-of-two scales (else nullptr)
    const float* __restrict__ row_scale_b,
    float2* __restrict__ partial,
    cfloat* __restrict__ direct, int nsub, int64_t out_outer) {
    using Tl = MfmaTile<P, WM, WN, BKC, TM, TN>;
    __shared__ MfmaSmem<P, WM, WN, BKC, TM, TN> sm[DB ? 2 : 1];

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
    f32x16 acc[TM][TN];
    zero(acc);

    const int am = m0 + staged_a_row<P, WM, WN, BKC, TM, TN>(tid);
    const int bn = n0 + staged_b_row<P, WM, WN, BKC, TM, TN>(tid);
    const float sa = (row_scale_a && am < M) ? row_scale_a[am] : 1.0f;
    const float sb = (row_scale_b && bn < N) ? row_scale_b[bn] : 1.0f;
    // 16-byte vector loads need 16-byte aligned rows
    const bool vec_a = ((reinterpret_cast<uintptr_t>(A) & 15) == 0) && (lda % 2 == 0);
    const bool vec_b = ((reinterpret_cast<uintptr_t>(B) & 15) == 0) && (ldb % 2 == 0);
    if constexpr (DB) {
        mfma_accumulate_db<P, WM, WN, BKC, TM, TN>(acc, sm, am < M ? A + int64_t(am) * lda : nullptr, sa, vec_a,
                                           bn < N ? B + int64_t(bn) * ldb : nullptr, sb, vec_b, kb, ke);
    } else {
        mfma_accumulate<P, WM, WN, BKC, TM, TN>(acc, sm[0], am < M ? A + int64_t(am) * lda : nullptr, sa, vec_a,
                                        bn < N ? B + int64_t(bn) * ldb : nullptr, sb, vec_b, kb, ke);
    }

    store_tile<WM, WN>(acc, m0, n0, ks, M, N, row_scale_a, row_scale_b, partial, direct, nsub, out_outer);
}

// One output tile of contraction_mfma_genb over K range [kb, ke); FULL as
// in mfma_accumulate_genb.
template<int P, int WM, int WN, bool FULL>
__device__ __forceinline__ void genb_tile(
    MfmaSmem<P, WM, WN, 16>& sm, const cfloat* __restrict__ A, const GenB& g, int M, int64_t lda, bool vec_a,
    int m0, int n0, int ks, int kb, int ke, float2* __restrict__ partial, cfloat* __restrict__ direct, int nsub,
    int64_t out_outer) {
    constexpr int BKC = 16;
    using Tl = MfmaTile<P, WM, WN, BKC>;
    f32x16 acc[2][2];
    zero(acc);
    const int am = m0 + staged_a_row<P, WM, WN, BKC>(threadIdx.x);
    mfma_accumulate_genb<P, WM, WN, BKC, 2, 2, FULL>(acc, sm, am < M ? A + int64_t(am) * lda : nullptr, 1.0f, vec_a,
                                                     g, n0, kb, ke);
    store_tile<WM, WN, Tl::THREADS / BKC>(acc, m0, n0, ks, M, g.N, nullptr, nullptr, partial, direct, nsub,
                                          out_outer);
}

// C = A B^T with B generated in staging (GenB); WM x WN waves, 16 complex K
// per step, BF16 precisions only.  Split-K and the output forms are those of
// contraction_mfma.  A workgroup whose tile lies inside the output, whose K
// range is whole steps and whose A rows take 16-byte loads runs the staging
// without bounds tests.
template<int P, int WM, int WN>
__global__ void __launch_bounds__(64 * WM * WN) contraction_mfma_genb(
    const cfloat* __restrict__ A, GenB g, int M, int K, int64_t lda,
    int k_chunk, int tiles_m, int tiles_n, int splits,
    float2* __restrict__ partial, cfloat* __restrict__ direct, int nsub, int64_t out_outer) {
    constexpr int BKC = 16;
    using Tl = MfmaTile<P, WM, WN, BKC>;
    __shared__ MfmaSmem<P, WM, WN, BKC> sm;

    const int total = tiles_m * tiles_n * splits;
    const int lid = xcd_logical_id(blockIdx.x, total);
    if (lid >= total) return;
    const int tn = lid % tiles_n;
    const int tm = (lid / tiles_n) % tiles_m;
    const int ks = lid / (tiles_n * tiles_m);
    const int m0 = tm * Tl::BM, n0 = tn * Tl::BNC;
    const int kb = ks * k_chunk;
    const int ke = min(K, kb + k_chunk);

    const bool vec_a = ((reinterpret_cast<uintptr_t>(A) & 15) == 0) && (lda % 2 == 0);
    const bool full = vec_a && m0 + Tl::BM <= M && n0 + Tl::BNC <= g.N && (ke - kb) % BKC == 0;
    if (full) {
        genb_tile<P, WM, WN, true>(sm, A, g, M, lda, vec_a, m0, n0, ks, kb, ke, partial, direct, nsub, out_outer);
    } else {
        genb_tile<P, WM, WN, false>(sm, A, g, M, lda, vec_a, m0, n0, ks, kb, ke, partial, direct, nsub, out_outer);
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
template<int P, int WM, int WN, int BKC, bool DB, int TM, int TN>
int resident_per_cu() {
    static const int cached = [] {
        int n = 0;
        COMPAS_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n, contraction_mfma<P, WM, WN, BKC, DB, TM, TN>, MfmaTile<P, WM, WN, BKC, TM, TN>::THREADS, 0));
        return std::max(1, n);
    }();
    return cached;
}

// Split-K plan.  All workgroups do the same amount of work, so the launch
// takes ceil(workgroups / resident slots) workgroup lifetimes: a grid just
// above a multiple of the slots leaves a nearly empty last round.  The split
// count minimises rounds x K-chunk (the critical path), with a small charge
// per split for the partial-sum traffic.
Plan make_plan_for(const ContractionShape& s, int BM, int BNC, int BKC, int threads, int64_t slots) {
    Plan p;
    p.threads = threads;
    p.tiles_m = int(dev::div_ceil(s.M, BM));
    p.tiles_n = int(dev::div_ceil(s.N, BNC));
    const int64_t tiles = int64_t(p.tiles_m) * p.tiles_n;
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

template<int P, int WM, int WN, int BKC, bool DB, int TM, int TN>
Plan make_plan(const CompasContext& ctx, const ContractionShape& s) {
    using Tl = MfmaTile<P, WM, WN, BKC, TM, TN>;
    const int64_t slots = int64_t(resident_per_cu<P, WM, WN, BKC, DB, TM, TN>()) * cu_count(ctx.device());
    return make_plan_for(s, Tl::BM, Tl::BNC, BKC, Tl::THREADS, slots);
}

template<int P, int WM, int WN, int BKC, bool DB, int TM, int TN>
void launch_variant(const CompasContext& ctx, const cfloat* A, const cfloat* B, cfloat* out,
                    const ContractionShape& s, const float* sa, const float* sb) {
    const Plan p = make_plan<P, WM, WN, BKC, DB, TM, TN>(ctx, s);
    const int total = p.tiles_m * p.tiles_n * p.splits;
    const bool direct = (p.splits == 1 && !s.accumulate);
    Array<float2, 1> partial;
    if (!direct) partial = ctx.empty<float2, 1>({int64_t(p.splits) * s.M * s.N});
    {
        KernelScope ks(ctx, "contraction_mfma");
        contraction_mfma<P, WM, WN, BKC, DB, TM, TN><<<total, p.threads, 0, ctx.stream()>>>(
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

#define COMPAS_MFMA_VARIANT(WM, WN, BKC, DB, TM, TN, NAME)                                                 \
    {NAME,                                                                                                    \
     {&launch_variant<P_BF16X3, WM, WN, BKC, DB, TM, TN>, &launch_variant<P_FP16X3, WM, WN, BKC, DB, TM, TN>, \
      &launch_variant<P_BF16X1, WM, WN, BKC, DB, TM, TN>, &launch_variant<P_FP16X1, WM, WN, BKC, DB, TM, TN>}}

// WM x WN waves of 64 x 64 outputs, BKC complex K per step; "db": two LDS
// buffers, one barrier per step.  (Tiles of 4 x 2 and 2 x 4 sub-tiles per
// wave, MfmaTile's TM / TN, measured 1.25-2.10 ms against 1.13 ms for 2x4x16
// at 1120 x 896 x 50176, and are not offered.)
const Variant VARIANTS[] = {
    COMPAS_MFMA_VARIANT(2, 2, 32, false, 2, 2, "2x2x32"),
    COMPAS_MFMA_VARIANT(4, 2, 32, false, 2, 2, "4x2x32"),
    COMPAS_MFMA_VARIANT(2, 4, 32, false, 2, 2, "2x4x32"),
    COMPAS_MFMA_VARIANT(4, 2, 16, false, 2, 2, "4x2x16"),
    COMPAS_MFMA_VARIANT(2, 4, 16, false, 2, 2, "2x4x16"),
    COMPAS_MFMA_VARIANT(4, 4, 16, false, 2, 2, "4x4x16"),
    COMPAS_MFMA_VARIANT(2, 4, 16, true, 2, 2, "2x4x16db"),
};
constexpr int NUM_VARIANTS = int(sizeof(VARIANTS) / sizeof(VARIANTS[0]));
constexpr int DEFAULT_VARIANT = 4;  // 2x4x16: fastest at 1120 x 896 x 50176 on MI355X

std::mutex g_tune_mutex;
std::map<std::tuple<int, int, int, int64_t, int64_t, int, int, bool>, int> g_tuned;
std::map<std::tuple<int, int, int, int64_t, int64_t, int, int, bool>, std::vector<float>> g_trial_ms;  // per variant
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
    // runtimes that do not time trials (the request server's slots, whose
    // request sizes vary per request) run the default tile for a new shape
    if (!ctx.runtime()->autotune()) return DEFAULT_VARIANT;
    // timing trials synchronise the host: not inside a graph capture (the
    // default tile is captured; a shape run once eagerly first uses its tuned one)
    if (ctx.runtime()->capturing()) return DEFAULT_VARIANT;
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
    std::vector<float> trial(NUM_VARIANTS, 0.0f);
    for (int v = 0; v < NUM_VARIANTS; v++) {
        VARIANTS[v].fn[P](ctx, A, B, scratch.data(), t, sa, sb);  // warm-up
        COMPAS_HIP_CHECK(hipEventRecord(e0, ctx.stream()));
        for (int r = 0; r < 2; r++) VARIANTS[v].fn[P](ctx, A, B, scratch.data(), t, sa, sb);
        COMPAS_HIP_CHECK(hipEventRecord(e1, ctx.stream()));
        COMPAS_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        COMPAS_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        trial[v] = ms / 2;
        if (ms < best_ms) {
            best_ms = ms;
            best = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tuned[key] = best;
    g_trial_ms[key] = trial;
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
        std::snprintf(line, sizeof(line), "M=%d N=%d K=%d prec=%s -> %s", std::get<1>(kv.first),
                      std::get<2>(kv.first), std::get<5>(kv.first), names[std::get<6>(kv.first)],
                      VARIANTS[kv.second].name);
        out += line;
        auto t = g_trial_ms.find(kv.first);
        if (t != g_trial_ms.end()) {  // the trial time of every variant, ms
            out += " (";
            for (int v = 0; v < NUM_VARIANTS; v++) {
                std::snprintf(line, sizeof(line), "%s%s %.3f", v ? ", " : "", VARIANTS[v].name, t->second[v]);
                out += line;
            }
            out += ")";
        }
        out += "\n";
    }
    return out;
}

void row_pow2_scales(const CompasContext& ctx, const cfloat* X, int rows, int K, int64_t ld, float* scale) {
    if (rows <= 0) return;
    KernelScope ks(ctx, "row_pow2_scale");
    row_pow2_scale<<<rows, 256, 0, ctx.stream()>>>(X, rows, K, ld, scale);
    check_launch("row_pow2_scale");
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
        row_pow2_scales(ctx, A, s.M, s.K, s.lda, scale_a.data());
        row_pow2_scales(ctx, B, s.N, s.K, s.ldb, scale_b.data());
        sa = scale_a.data();
        sb = scale_b.data();
    }
    const int v = pick_variant(ctx, A, B, s, P, sa, sb);
    VARIANTS[v].fn[P](ctx, A, B, out, s, sa, sb);
}

namespace {

// tile of the generated-operand contraction (waves along M x along N)
constexpr int GENB_WM = 2, GENB_WN = 4;

template<int P>
void launch_genb(const CompasContext& ctx, const cfloat* A, const GenB& g, cfloat* out, const ContractionShape& s) {
    using Tl = MfmaTile<P, GENB_WM, GENB_WN, 16>;
    static const int resident = [] {
        int n = 0;
        COMPAS_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n, contraction_mfma_genb<P, GENB_WM, GENB_WN>, Tl::THREADS, 0));
        return std::max(1, n);
    }();
    const Plan p = make_plan_for(s, Tl::BM, Tl::BNC, 16, Tl::THREADS, int64_t(resident) * cu_count(ctx.device()));
    const int total = p.tiles_m * p.tiles_n * p.splits;
    const bool direct = (p.splits == 1 && !s.accumulate);
    Array<float2, 1> partial;
    if (!direct) partial = ctx.empty<float2, 1>({int64_t(p.splits) * s.M * s.N});
    {
        KernelScope ks(ctx, "contraction_mfma");
        contraction_mfma_genb<P, GENB_WM, GENB_WN><<<total, p.threads, 0, ctx.stream()>>>(
            A, g, s.M, s.K, s.lda, p.k_chunk, p.tiles_m, p.tiles_n, p.splits,
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

}  // namespace

bool complex_contraction_generated(
    const CompasContext& ctx, const cfloat* A, const cfloat* C, const cfloat* T, const cfloat* q, int ns,
    cfloat* out, ContractionShape s, ContractionPrecision precision) {
    precision = effective_precision(precision);
    if (precision != ContractionPrecision::BF16x3 && precision != ContractionPrecision::BF16x1) return false;
    COMPAS_CHECK(ns > 0 && ns % 4 == 0 && s.N % ns == 0);
    COMPAS_CHECK(s.M >= 0 && s.K >= 0 && s.nsub > 0);
    if (s.M == 0 || s.N == 0) return true;
    const GenB g {C, T, q, ns, s.K, s.N};
    if (precision == ContractionPrecision::BF16x3) {
        launch_genb<P_BF16X3>(ctx, A, g, out, s);
    } else {
        launch_genb<P_BF16X1>(ctx, A, g, out, s);
    }
    return true;
}

}  // namespace compas
