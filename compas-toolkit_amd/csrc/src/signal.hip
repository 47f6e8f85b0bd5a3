// magnetization_to_signal and phase_encoding on gfx950.
//
// Semantics: reference src/trajectories/signal.cu:13-356,
// signal_kernels.cuh:12-345, include/compas/trajectories/cartesian_view.cuh:28-63,
// spiral_view.cuh:13-73, src/trajectories/phase_encoding_kernels.cuh:10-25.
//
// Cartesian:  S[c][r][s] = sum_v m[r][v] * W[c*ns + s][v] with
//   W = rho_v * C[c][v] * exp((ns/2 - s) dt R2) * exp(i phi),
//   phi = s*Theta_v - ns/2 (dt 2 pi B0 + dk_re x)  (readout factor folded in),
//   Theta_v = dk_re x + dk_im y + dt 2 pi B0,
// so the raw echoes are the A operand of one contraction (M = nr, N = nc*ns,
// K = nv) and the coil dimension is folded into N.  phi is formed in FP64 and
// reduced to [-pi, pi) before the float sincos.  For the BF16 precisions W is not stored:
// the contraction generates it while staging from W[c ns + 4a + b][v] =
// C[c][v] T[a][v] q[v]^b (T: the coil-free weight of every fourth sample, q
// = exp(-dt R2) exp(i Theta_v)), which replaces nc ns rows of W in HBM by
// ns / 4 + 1 rows of factors.
//
// Spiral:  S[c][r][s] = sum_v A[r][v] C[c][v] E_rv^s with per-readout
// exponents, evaluated by a direct kernel: FP64-anchored powers E^s0 every 8
// samples, complex recurrence in between.

#include <cstdlib>
#include <string>

#include "compas/device/common.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

__global__ void signal_cartesian_weights(
    dev::TissueView tissue, const cfloat* __restrict__ coils, int nc, int ns,
    float dt, float dk_re, float dk_im, cfloat* __restrict__ W) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    const int nv = tissue.nvoxels;
    if (v >= nv) return;
    const float R2 = 1.0f / tissue.T2(v);
    const double B0 = tissue.B0(v), x = tissue.x(v), y = tissue.y(v);
    const double off = double(dt) * dev::TWO_PI * B0;
    const double theta = double(dk_re) * x + double(dk_im) * y + off;
    const double phi = double(s) * theta - 0.5 * double(ns) * (off + double(dk_re) * x);
    const float mag = expf((0.5f * float(ns) - float(s)) * dt * R2);
    const cfloat e = dev::cmul(dev::polar_f(mag, dev::reduce_phase(phi)), tissue.rho(v));
    for (int c = 0; c < nc; c++) {
        W[(size_t(c) * ns + s) * nv + v] = dev::cmul(e, coils[size_t(c) * nv + v]);
    }
}

// The factors of W in the form the contraction generates it from
// (W[c ns + s][v] = C[c][v] T[s / 4][v] q[v]^(s % 4)): rows a < ns / 4 of T
// are the coil-free weight at sample 4a, row ns / 4 is the ratio q of two
// consecutive samples, exp(-dt R2) exp(i Theta_v).
__global__ void signal_cartesian_anchors(
    dev::TissueView tissue, int ns, float dt, float dk_re, float dk_im, cfloat* __restrict__ T) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = blockIdx.y;
    const int nv = tissue.nvoxels;
    if (v >= nv) return;
    const float R2 = 1.0f / tissue.T2(v);
    const double B0 = tissue.B0(v), x = tissue.x(v), y = tissue.y(v);
    const double off = double(dt) * dev::TWO_PI * B0;
    const double theta = double(dk_re) * x + double(dk_im) * y + off;
    if (a == ns / 4) {
        T[size_t(a) * nv + v] = dev::polar_f(expf(-dt * R2), dev::reduce_phase(theta));
        return;
    }
    const int s = 4 * a;
    const double phi = double(s) * theta - 0.5 * double(ns) * (off + double(dk_re) * x);
    const float mag = expf((0.5f * float(ns) - float(s)) * dt * R2);
    T[size_t(a) * nv + v] = dev::cmul(dev::polar_f(mag, dev::reduce_phase(phi)), tissue.rho(v));
}

// A[r][v] = m[r][v] * rewind * prephaser * rho (spiral_view.cuh:13-57, signal_kernels.cuh:12-30)
__global__ void signal_spiral_factors(
    dev::TissueView tissue, const cfloat* __restrict__ echos, const cfloat* __restrict__ k_start,
    int nr, int ns, float dt, cfloat* __restrict__ A) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    const int nv = tissue.nvoxels;
    if (v >= nv) return;
    const float R2 = 1.0f / tissue.T2(v);
    const double B0 = tissue.B0(v), x = tissue.x(v);
    const double half = 0.5 * double(ns) * double(dt);
    const double phi = -half * dev::TWO_PI * B0 + double(k_start[r].re) * x;
    const float mag = expf(float(half) * R2);
    const cfloat f = dev::cmul(dev::polar_f(mag, dev::reduce_phase(phi)), tissue.rho(v));
    A[size_t(r) * nv + v] = dev::cmul(echos[size_t(r) * nv + v], f);
}

constexpr int SP_COILS = 4;  // coils per block

// Per-voxel inputs of the spiral sum, loaded one voxel ahead of their use.
template<bool TWO>
struct SpiralInputs {
    float T2, x, y, B0;
    cfloat A, Ab;
    cfloat C[SP_COILS];
};

template<bool TWO>
__device__ inline SpiralInputs<TWO> spiral_inputs(
    const dev::TissueView& t, const cfloat* __restrict__ A, const cfloat* __restrict__ Ab,
    const cfloat* __restrict__ coils, int c0, int ncl, size_t row, int v) {
    SpiralInputs<TWO> in;
    const int nv = t.nvoxels;
    in.T2 = t.T2(v);
    in.x = t.x(v);
    in.y = t.y(v);
    in.B0 = t.B0(v);
    in.A = A[row + v];
    in.Ab = TWO ? Ab[row + v] : cfloat {0, 0};
#pragma unroll
    for (int c = 0; c < SP_COILS; c++) in.C[c] = (c < ncl) ? coils[size_t(c0 + c) * nv + v] : cfloat {0, 0};
    return in;
}

// out[c][r][s] = sum_v C[c][v] (A[r][v] + (s - ns/2) Ab[r][v]) E_rv^s  (Ab only if TWO)
// grid: (nr, ceil(ns / (8 SPS)), ceil(nc / SP_COILS)); block 256 = 4 waves.
// lane % 32 strides over voxels; (wave, lane / 32) selects SPS consecutive
// samples.  E^s0 is evaluated from an FP64 phase at the first sample of the
// thread, then by the complex recurrence.
template<bool TWO, int SPS>
__global__ void __launch_bounds__(256) signal_spiral_direct(
    dev::TissueView tissue, const cfloat* __restrict__ A, const cfloat* __restrict__ Ab,
    const cfloat* __restrict__ coils, int nc,
    const cfloat* __restrict__ delta_k, int nr, int ns, float dt, cfloat* __restrict__ signal) {
    const int r = blockIdx.x;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int vl = lane & 31;
    const int sg = wave * 2 + (lane >> 5);
    const int s0 = blockIdx.y * 8 * SPS + sg * SPS;
    const int c0 = blockIdx.z * SP_COILS;
    const int ncl = min(SP_COILS, nc - c0);
    const int nv = tissue.nvoxels;
    const size_t row = size_t(r) * nv;
    const cfloat dk = delta_k[r];
    const double two_pi_dt = double(dt) * dev::TWO_PI;
    const float half = 0.5f * float(ns);

    float acc_re[SP_COILS][SPS], acc_im[SP_COILS][SPS];
#pragma unroll
    for (int c = 0; c < SP_COILS; c++)
#pragma unroll
        for (int i = 0; i < SPS; i++) {
            acc_re[c][i] = 0.0f;
            acc_im[c][i] = 0.0f;
        }

    if (s0 < ns && vl < nv) {
        SpiralInputs<TWO> next = spiral_inputs<TWO>(tissue, A, Ab, coils, c0, ncl, row, vl);
        for (int v = vl; v < nv; v += 32) {
            const SpiralInputs<TWO> in = next;
            if (v + 32 < nv) next = spiral_inputs<TWO>(tissue, A, Ab, coils, c0, ncl, row, v + 32);
            const float R2 = 1.0f / in.T2;
            const double theta = double(dk.re) * double(in.x) + double(dk.im) * double(in.y)
                + two_pi_dt * double(in.B0);
            const float decay = -dt * R2;
            const cfloat step = dev::polar_reduced(expf(decay), theta);
            const cfloat anchor = dev::polar_reduced(expf(float(s0) * decay), double(s0) * theta);
            cfloat cur = dev::cmul(anchor, in.A);
            cfloat cur_b {0, 0};
            if constexpr (TWO) cur_b = dev::cmul(anchor, in.Ab);
#pragma unroll
            for (int i = 0; i < SPS; i++) {
                cfloat amp = cur;
                if constexpr (TWO) {
                    const float sp = float(s0 + i) - half;
                    amp = {fmaf(sp, cur_b.re, cur.re), fmaf(sp, cur_b.im, cur.im)};
                    cur_b = dev::cmul(cur_b, step);
                }
#pragma unroll
                for (int c = 0; c < SP_COILS; c++) {
                    acc_re[c][i] = fmaf(amp.re, in.C[c].re, fmaf(-amp.im, in.C[c].im, acc_re[c][i]));
                    acc_im[c][i] = fmaf(amp.re, in.C[c].im, fmaf(amp.im, in.C[c].re, acc_im[c][i]));
                }
                cur = dev::cmul(cur, step);
            }
        }
    }
    // reduce over the 32 voxel lanes of each half-wave
#pragma unroll
    for (int c = 0; c < SP_COILS; c++)
#pragma unroll
        for (int i = 0; i < SPS; i++) {
            float re = acc_re[c][i], im = acc_im[c][i];
#pragma unroll
            for (int off = 16; off > 0; off >>= 1) {
                re += __shfl_xor(re, off, 32);
                im += __shfl_xor(im, off, 32);
            }
            acc_re[c][i] = re;
            acc_im[c][i] = im;
        }
    if (vl == 0 && s0 < ns) {
#pragma unroll
        for (int c = 0; c < SP_COILS; c++) {
            if (c >= ncl) continue;
#pragma unroll
            for (int i = 0; i < SPS; i++) {
                if (s0 + i < ns) signal[(size_t(c0 + c) * nr + r) * ns + s0 + i] = {acc_re[c][i], acc_im[c][i]};
            }
        }
    }
}

__global__ void phase_encoding_kernel(
    dev::TissueView tissue, const cfloat* __restrict__ echos, const cfloat* __restrict__ k_start, int nr,
    cfloat* __restrict__ out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    const int nv = tissue.nvoxels;
    if (v >= nv) return;
    const float ky = k_start[r].im;
    const double phi = double(tissue.y(v)) * double(ky);
    out[size_t(r) * nv + v] = dev::cmul(echos[size_t(r) * nv + v], dev::polar_f(1.0f, dev::reduce_phase(phi)));
}

}  // namespace kernels

static void check_signal_shapes(const Array<cfloat, 2>& echos, const TissueParameters& p, const Trajectory& t,
                                const Array<cfloat, 2>& coils) {
    COMPAS_CHECK(echos.size(0) == t.nreadouts);
    COMPAS_CHECK(echos.size(1) == p.nvoxels);
    COMPAS_CHECK(coils.size(1) == p.nvoxels);
    COMPAS_CHECK(t.samples_per_readout >= 0);
}

// The Cartesian weights are generated inside the contraction unless
// COMPAS_CARTESIAN_WEIGHTS=stored asks for the stored form (for comparisons).
static bool generated_weights() {
    static const bool on = [] {
        const char* e = std::getenv("COMPAS_CARTESIAN_WEIGHTS");
        return !(e && std::string(e) == "stored");
    }();
    return on;
}

static Array<cfloat, 3> signal_cartesian(
    const CompasContext& ctx, const Array<cfloat, 2>& echos, const TissueParameters& p,
    const CartesianTrajectory& t, const Array<cfloat, 2>& coils, SimulateSignalMethod method) {
    const int nc = int(coils.size(0)), nr = t.nreadouts, ns = t.samples_per_readout, nv = p.nvoxels;
    auto signal = ctx.empty<cfloat, 3>({nc, nr, ns});
    if (signal.numel() == 0) return signal;
    if (nv == 0) {
        ctx.fill_zero(signal.data(), signal.nbytes());
        return signal;
    }
    ContractionShape s {nr, nc * ns, nv, nv, nv, ns, int64_t(nr) * ns, false};
    const ContractionPrecision prec = effective_precision(signal_precision(method));
    if ((prec == ContractionPrecision::BF16x3 || prec == ContractionPrecision::BF16x1) && ns % 4 == 0 &&
        generated_weights()) {
        // W generated in the contraction's staging from ns / 4 + 1 rows of
        // factors instead of written to HBM (nc ns rows)
        auto T = ctx.empty<cfloat, 2>({ns / 4 + 1, nv});
        {
            KernelScope ks(ctx, "signal_cartesian_anchors");
            dim3 grid(dev::div_ceil(nv, 256), ns / 4 + 1);
            kernels::signal_cartesian_anchors<<<grid, 256, 0, ctx.stream()>>>(
                dev::tissue_view(p), ns, t.delta_t, t.delta_k.re, t.delta_k.im, T.data());
            check_launch("signal_cartesian_anchors");
        }
        const bool done = complex_contraction_generated(ctx, echos.data(), coils.data(), T.data(),
                                                        T.data() + size_t(ns / 4) * nv, ns, signal.data(), s, prec);
        COMPAS_CHECK(done);
        return signal;
    }
    auto W = ctx.empty<cfloat, 2>({int64_t(nc) * ns, nv});
    {
        KernelScope ks(ctx, "signal_cartesian_weights");
        dim3 grid(dev::div_ceil(nv, 256), ns);
        kernels::signal_cartesian_weights<<<grid, 256, 0, ctx.stream()>>>(
            dev::tissue_view(p), coils.data(), nc, ns, t.delta_t, t.delta_k.re, t.delta_k.im, W.data());
        check_launch("signal_cartesian_weights");
    }
    complex_contraction(ctx, echos.data(), W.data(), signal.data(), s, signal_precision(method));
    return signal;
}

void spiral_sum(
    const CompasContext& ctx, const TissueParameters& p, const cfloat* A, const cfloat* Ab, const cfloat* coils,
    int nc, const SpiralTrajectory& t, ContractionPrecision precision, cfloat* out) {
    const int nr = t.nreadouts, ns = t.samples_per_readout;
    if (nc == 0 || nr == 0 || ns == 0) return;
    if (spiral_sum_mfma(ctx, p, A, Ab, coils, nc, t, precision, out)) return;
    // 8 samples per thread: measured fastest at 4 coils x 1120 x 512 over 50176
    // voxels (24.7 ms, against 35.1 ms for 4 and 30.8 ms for 16, which needs
    // 256 VGPRs and leaves one wave per SIMD)
    constexpr int SPS = 8;
    KernelScope ks(ctx, "signal_spiral_direct");
    dim3 grid(nr, dev::div_ceil(ns, 8 * SPS), dev::div_ceil(nc, kernels::SP_COILS));
    const auto tv = dev::tissue_view(p);
    if (Ab) {
        kernels::signal_spiral_direct<true, SPS><<<grid, 256, 0, ctx.stream()>>>(
            tv, A, Ab, coils, nc, t.delta_k.data(), nr, ns, t.delta_t, out);
    } else {
        kernels::signal_spiral_direct<false, SPS><<<grid, 256, 0, ctx.stream()>>>(
            tv, A, nullptr, coils, nc, t.delta_k.data(), nr, ns, t.delta_t, out);
    }
    check_launch("signal_spiral_direct");
}

static Array<cfloat, 3> signal_spiral(
    const CompasContext& ctx, const Array<cfloat, 2>& echos, const TissueParameters& p,
    const SpiralTrajectory& t, const Array<cfloat, 2>& coils, SimulateSignalMethod method) {
    const int nc = int(coils.size(0)), nr = t.nreadouts, ns = t.samples_per_readout, nv = p.nvoxels;
    COMPAS_CHECK(t.k_start.size(0) == nr && t.delta_k.size(0) == nr);
    auto signal = ctx.empty<cfloat, 3>({nc, nr, ns});
    if (signal.numel() == 0) return signal;
    if (nv == 0) {
        ctx.fill_zero(signal.data(), signal.nbytes());
        return signal;
    }
    auto A = ctx.empty<cfloat, 2>({nr, nv});
    {
        KernelScope ks(ctx, "signal_spiral_factors");
        dim3 grid(dev::div_ceil(nv, 256), nr);
        kernels::signal_spiral_factors<<<grid, 256, 0, ctx.stream()>>>(
            dev::tissue_view(p), echos.data(), t.k_start.data(), nr, ns, t.delta_t, A.data());
        check_launch("signal_spiral_factors");
    }
    spiral_sum(ctx, p, A.data(), nullptr, coils.data(), nc, t, effective_precision(signal_precision(method)),
               signal.data());
    return signal;
}

Array<cfloat, 3> magnetization_to_signal(
    const CompasContext& ctx,
    Array<cfloat, 2> echos,
    TissueParameters parameters,
    const Trajectory& trajectory,
    Array<cfloat, 2> coils,
    SimulateSignalMethod method) {
    ctx.activate();
    check_signal_shapes(echos, parameters, trajectory, coils);
    if (const auto* cart = dynamic_cast<const CartesianTrajectory*>(&trajectory)) {
        (void)signal_precision(method);  // validates the enumerator
        return signal_cartesian(ctx, echos, parameters, *cart, coils, method);
    } else if (const auto* sp = dynamic_cast<const SpiralTrajectory*>(&trajectory)) {
        return signal_spiral(ctx, echos, parameters, *sp, coils, method);
    }
    COMPAS_ERROR("invalid trajectory type");
}

Array<cfloat, 2> phase_encoding(
    const CompasContext& ctx,
    const Array<cfloat, 2>& echos,
    const TissueParameters& parameters,
    const CartesianTrajectory& trajectory) {
    ctx.activate();
    const int nr = trajectory.nreadouts, nv = parameters.nvoxels;
    COMPAS_CHECK(echos.size(0) == nr && echos.size(1) == nv);
    COMPAS_CHECK(trajectory.k_start.size(0) == nr);
    auto out = ctx.empty<cfloat, 2>({nr, nv});
    if (out.numel() == 0) return out;
    KernelScope ks(ctx, "phase_encoding");
    dim3 grid(dev::div_ceil(nv, 256), nr);
    kernels::phase_encoding_kernel<<<grid, 256, 0, ctx.stream()>>>(
        dev::tissue_view(parameters), echos.data(), trajectory.k_start.data(), nr, out.data());
    check_launch("phase_encoding");
    return out;
}

}  // namespace compas
