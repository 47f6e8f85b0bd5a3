// Jacobian products on the spiral trajectory with six parameters
// (T1, T2, B1, B0, rho_x, rho_y).
//
// The reference's Jacobian covers the Cartesian trajectory and four parameters
// (src/jacobian/product.cu, hermitian.cu).  These overloads extend it to the
// spiral signal of spiral_view.cuh:13-73 with the same conventions
// (product_kernels.cuh:56-60: the T1 and T2 rows are scaled by T1 and T2, the
// rho rows multiply m and i m).  With the spiral sample factor
//   Phi[r][s][v] = exp((s - ns/2) dt (-R2 + i 2 pi B0)
//                      + i (Re k0_r x + s (Re dk_r x + Im dk_r y)))
// the signal is S = sum_v C rho m Phi, dPhi/dT2 = (s - ns/2) dt R2^2 Phi and
// dPhi/dB0 = (s - ns/2) i 2 pi dt Phi, so
//   J.v   = sum_v C (a + (s - ns/2) b) Phi
//           a = rho (v0 T1 dm_T1 + v1 T2 dm_T2 + v2 dm_B1 + v3 dm_B0) + (v4 + i v5) m
//           b = rho m dt (v1 T2 R2^2 + i 2 pi v3)
// runs on the spiral sum kernel of the signal with a second amplitude, and
// J^H.y is its exact adjoint: per voxel, coil and readout
//   U = sum_s y conj(Phi),  U' = sum_s (s - ns/2) y conj(Phi)
// reduced over readouts against conj(m) and conj(dm_p).
//
// The spiral sample factor differs per readout, so neither product is one
// GEMM over all readouts.  The FP32 forms here run on VALU with FP64 phase
// anchors; with a mixed-precision method both run as one GEMM per readout
// with a generated operand on MFMA (signal_spiral_mfma.hip for J.v through
// the spiral sum, jacobian_spiral_mfma.hip for J^H.y).

#include "compas/device/common.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

// readout factor Phi[r][0][v] = exp(ns/2 dt R2) exp(i phi0), phi0 = -ns/2 dt 2 pi B0 + Re k0_r x
__device__ inline cfloat spiral_readout_factor(const dev::TissueView& t, int v, cfloat k0, int ns, float dt) {
    const float R2 = 1.0f / t.T2(v);
    const double half = 0.5 * double(ns) * double(dt);
    const double phi = -half * dev::TWO_PI * double(t.B0(v)) + double(k0.re) * double(t.x(v));
    return dev::polar_f(expf(float(half) * R2), dev::reduce_phase(phi));
}

// J.v amplitudes times the readout factor: Aa = a Phi0, Ab = b Phi0 ([nr][nv])
__global__ void spiral_jv_amplitudes(
    dev::TissueView t, const cfloat* __restrict__ m, const cfloat* __restrict__ dT1,
    const cfloat* __restrict__ dT2, const cfloat* __restrict__ dB1, const cfloat* __restrict__ dB0,
    const cfloat* __restrict__ vec, const cfloat* __restrict__ k_start, int nr, int ns, float dt,
    cfloat* __restrict__ Aa, cfloat* __restrict__ Ab) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    const int nv = t.nvoxels;
    if (v >= nv) return;
    const size_t i = size_t(r) * nv + v;
    const cfloat rho = t.rho(v), me = m[i];
    const float T1 = t.T1(v), T2 = t.T2(v), R2 = 1.0f / T2;
    const cfloat v0 = vec[v], v1 = vec[size_t(nv) + v], v2 = vec[2 * size_t(nv) + v],
                 v3 = vec[3 * size_t(nv) + v], v4 = vec[4 * size_t(nv) + v], v5 = vec[5 * size_t(nv) + v];
    cfloat a = dev::cmul(v0, dT1[i]) * T1;
    a = a + dev::cmul(v1, dT2[i]) * T2;
    a = a + dev::cmul(v2, dB1[i]);
    a = a + dev::cmul(v3, dB0[i]);
    a = dev::cmul(a, rho);
    a = a + dev::cmul(v4 + dev::cmul(v5, cfloat {0.0f, 1.0f}), me);
    cfloat b = v1 * (T2 * R2 * R2) + dev::cmul(v3, cfloat {0.0f, dev::TWO_PI_F});
    b = dev::cmul(dev::cmul(b, rho), me) * dt;
    const cfloat f = spiral_readout_factor(t, v, k_start[r], ns, dt);
    Aa[i] = dev::cmul(a, f);
    Ab[i] = dev::cmul(b, f);
}

constexpr int JHS_THREADS = 256;  // voxels per workgroup (one per lane)
constexpr int JHS_ANCHOR = 32;  // samples between FP64-anchored evaluations of Phi
constexpr int JHS_BLOCK = 8;  // samples per unrolled block (y of the block loaded up front)

struct SpiralVoxel {
    float R2, decay, T1, T2;
    double B0, x, y, half;
    cfloat rho;
};

// Adds the coils [cg, cg + CG) of one voxel to out[6]: per readout r in
// [r0, r1), U = sum_s y conj(Phi) and U' = sum_s (s - ns/2) y conj(Phi), then
//   out += conj(rho T1 C) sum_r conj(dm_T1) U, conj(rho T2 C) sum_r
//   (conj(dm_T2) U + dt R2^2 conj(m) U'), conj(rho C) sum_r conj(dm_B1) U,
//   conj(rho C) sum_r (conj(dm_B0) U - i 2 pi dt conj(m) U'), conj(C) sum_r conj(m) U.
// y is wave-uniform: the loads of a sample block become scalar loads issued
// before the block's arithmetic.
template<int CG>
__device__ inline void spiral_hermitian_coils(
    const SpiralVoxel& sv, int vv, int nv, const cfloat* __restrict__ m, const cfloat* __restrict__ dT1,
    const cfloat* __restrict__ dT2, const cfloat* __restrict__ dB1, const cfloat* __restrict__ dB0,
    const cfloat* __restrict__ coils, int cg, const cfloat* __restrict__ y, const cfloat* __restrict__ k_start,
    const cfloat* __restrict__ delta_k, int nr, int ns, float dt, int r0, int r1, cfloat (&out)[6]) {
    cfloat g1[CG], g2[CG], gb1[CG], gb0[CG], gu[CG];
#pragma unroll
    for (int c = 0; c < CG; c++) g1[c] = g2[c] = gb1[c] = gb0[c] = gu[c] = {0, 0};
    const size_t coil_stride = size_t(nr) * ns;

    for (int r = r0; r < r1; r++) {
        const cfloat dk = delta_k[r], k0 = k_start[r];
        const double theta = double(dk.re) * sv.x + double(dk.im) * sv.y + double(dt) * dev::TWO_PI * sv.B0;
        const double phi0 = -sv.half * dev::TWO_PI * sv.B0 + double(k0.re) * sv.x;
        const cfloat step = dev::polar_f(expf(sv.decay), dev::reduce_phase(theta));
        cfloat U[CG], Up[CG];
#pragma unroll
        for (int c = 0; c < CG; c++) U[c] = Up[c] = {0, 0};
        const cfloat* yr = y + (size_t(cg) * nr + r) * ns;
        for (int sa = 0; sa < ns; sa += JHS_ANCHOR) {
            // Phi at the anchor sample: magnitude exp((ns/2 - s) dt R2), phase phi0 + s theta
            cfloat ph = dev::polar_f(expf((0.5f * float(ns) - float(sa)) * dt * sv.R2),
                                     dev::reduce_phase(phi0 + double(sa) * theta));
            const int se = min(ns, sa + JHS_ANCHOR);
            int s = sa;
            for (; s + JHS_BLOCK <= se; s += JHS_BLOCK) {
                cfloat yb[CG][JHS_BLOCK];
#pragma unroll
                for (int c = 0; c < CG; c++)
#pragma unroll
                    for (int j = 0; j < JHS_BLOCK; j++) yb[c][j] = yr[size_t(c) * coil_stride + s + j];
#pragma unroll
                for (int j = 0; j < JHS_BLOCK; j++) {
                    const float sp = float(s + j) - 0.5f * float(ns);
#pragma unroll
                    for (int c = 0; c < CG; c++) {
                        const cfloat yv = yb[c][j];
                        const cfloat q {fmaf(yv.re, ph.re, yv.im * ph.im), fmaf(yv.im, ph.re, -yv.re * ph.im)};
                        U[c].re += q.re;
                        U[c].im += q.im;
                        Up[c].re = fmaf(sp, q.re, Up[c].re);
                        Up[c].im = fmaf(sp, q.im, Up[c].im);
                    }
                    ph = dev::cmul(ph, step);
                }
            }
            for (; s < se; s++) {
                const float sp = float(s) - 0.5f * float(ns);
#pragma unroll
                for (int c = 0; c < CG; c++) {
                    const cfloat yv = yr[size_t(c) * coil_stride + s];
                    const cfloat q {fmaf(yv.re, ph.re, yv.im * ph.im), fmaf(yv.im, ph.re, -yv.re * ph.im)};
                    U[c].re += q.re;
                    U[c].im += q.im;
                    Up[c].re = fmaf(sp, q.re, Up[c].re);
                    Up[c].im = fmaf(sp, q.im, Up[c].im);
                }
                ph = dev::cmul(ph, step);
            }
        }
        const size_t i = size_t(r) * nv + vv;
        const cfloat me = m[i], a1 = dT1[i], a2 = dT2[i], ab1 = dB1[i], ab0 = dB0[i];
#pragma unroll
        for (int c = 0; c < CG; c++) {
            g1[c] = dev::cfma_conj(U[c], a1, g1[c]);
            g2[c] = dev::cfma_conj(U[c], a2, g2[c]);
            g2[c] = dev::cfma_conj(Up[c] * (dt * sv.R2 * sv.R2), me, g2[c]);
            gb1[c] = dev::cfma_conj(U[c], ab1, gb1[c]);
            gb0[c] = dev::cfma_conj(U[c], ab0, gb0[c]);
            gb0[c] = dev::cfma_conj(dev::cmul(Up[c], cfloat {0.0f, -dev::TWO_PI_F * dt}), me, gb0[c]);
            gu[c] = dev::cfma_conj(U[c], me, gu[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < CG; c++) {
        const cfloat C = coils[size_t(cg + c) * nv + vv];
        const cfloat rc = dev::cmul(sv.rho, C);
        out[0] = out[0] + dev::cmul(conj(rc * sv.T1), g1[c]);
        out[1] = out[1] + dev::cmul(conj(rc * sv.T2), g2[c]);
        out[2] = out[2] + dev::cmul(conj(rc), gb1[c]);
        out[3] = out[3] + dev::cmul(conj(rc), gb0[c]);
        out[4] = out[4] + dev::cmul(conj(C), gu[c]);
    }
}

// Partial J^H.y over readouts [r0, r1) for one voxel per lane:
// partial[chunk][p][v], p over the six parameters.  Coils run in groups of
// four, then one group of the remaining one to three.
__global__ void __launch_bounds__(JHS_THREADS) spiral_hermitian(
    dev::TissueView t, const cfloat* __restrict__ m, const cfloat* __restrict__ dT1,
    const cfloat* __restrict__ dT2, const cfloat* __restrict__ dB1, const cfloat* __restrict__ dB0,
    const cfloat* __restrict__ coils, int nc, const cfloat* __restrict__ y, const cfloat* __restrict__ k_start,
    const cfloat* __restrict__ delta_k, int nr, int ns, float dt, int r_chunk, cfloat* __restrict__ partial) {
    const int v = blockIdx.x * JHS_THREADS + threadIdx.x;
    const int chunk = blockIdx.y;
    const int nv = t.nvoxels;
    const int r0 = chunk * r_chunk, r1 = min(nr, r0 + r_chunk);
    const bool ok = v < nv;
    const int vv = ok ? v : 0;  // out-of-range lanes compute on voxel 0 and store nothing

    SpiralVoxel sv;
    sv.T1 = t.T1(vv);
    sv.T2 = t.T2(vv);
    sv.R2 = 1.0f / sv.T2;
    sv.decay = -dt * sv.R2;
    sv.B0 = t.B0(vv);
    sv.x = t.x(vv);
    sv.y = t.y(vv);
    sv.half = 0.5 * double(ns) * double(dt);
    sv.rho = t.rho(vv);

    cfloat out[6];
#pragma unroll
    for (int p = 0; p < 6; p++) out[p] = {0, 0};
    int cg = 0;
    for (; cg + 4 <= nc; cg += 4)
        spiral_hermitian_coils<4>(sv, vv, nv, m, dT1, dT2, dB1, dB0, coils, cg, y, k_start, delta_k, nr, ns, dt,
                                  r0, r1, out);
    switch (nc - cg) {
        case 3:
            spiral_hermitian_coils<3>(sv, vv, nv, m, dT1, dT2, dB1, dB0, coils, cg, y, k_start, delta_k, nr, ns, dt,
                                      r0, r1, out);
            break;
        case 2:
            spiral_hermitian_coils<2>(sv, vv, nv, m, dT1, dT2, dB1, dB0, coils, cg, y, k_start, delta_k, nr, ns, dt,
                                      r0, r1, out);
            break;
        case 1:
            spiral_hermitian_coils<1>(sv, vv, nv, m, dT1, dT2, dB1, dB0, coils, cg, y, k_start, delta_k, nr, ns, dt,
                                      r0, r1, out);
            break;
        default:
            break;
    }
    out[5] = dev::cmul(out[4], cfloat {0.0f, -1.0f});
    if (ok) {
#pragma unroll
        for (int p = 0; p < 6; p++) partial[(size_t(chunk) * 6 + p) * nv + v] = out[p];
    }
}

// JHy[p][v] = sum over readout chunks (FP64)
__global__ void spiral_hermitian_reduce(const cfloat* __restrict__ partial, int chunks, int nv, cfloat* __restrict__ JHy) {
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= 6 * int64_t(nv)) return;
    double re = 0, im = 0;
    for (int k = 0; k < chunks; k++) {
        const cfloat p = partial[int64_t(k) * 6 * nv + idx];
        re += double(p.re);
        im += double(p.im);
    }
    JHy[idx] = {float(re), float(im)};
}

}  // namespace kernels

void spiral_hermitian_reduce_partials(const CompasContext& ctx, const cfloat* partial, int chunks, int nv, cfloat* JHy) {
    KernelScope ks(ctx, "spiral_hermitian_reduce");
    const int64_t n = 6 * int64_t(nv);
    kernels::spiral_hermitian_reduce<<<dev::div_ceil(n, 256), 256, 0, ctx.stream()>>>(partial, chunks, nv, JHy);
    check_launch("spiral_hermitian_reduce");
}

namespace {

void check_spiral_jacobian_inputs(
    const Array<cfloat, 2>& echos, const Array<cfloat, 2>& dT1, const Array<cfloat, 2>& dT2,
    const Array<cfloat, 2>& dB1, const Array<cfloat, 2>& dB0, const TissueParameters& p, const SpiralTrajectory& t,
    const Array<cfloat, 2>& coils) {
    const int nr = t.nreadouts, nv = p.nvoxels;
    for (const Array<cfloat, 2>* a : {&echos, &dT1, &dT2, &dB1, &dB0}) {
        COMPAS_CHECK(a->size(0) == nr);
        COMPAS_CHECK(a->size(1) == nv);
    }
    COMPAS_CHECK(coils.size(1) == nv);
    COMPAS_CHECK(t.k_start.size(0) == nr && t.delta_k.size(0) == nr);
}

}  // namespace

Array<cfloat, 3> compute_jacobian(
    const CompasContext& ctx,
    Array<cfloat, 2> echos,
    Array<cfloat, 2> dT1,
    Array<cfloat, 2> dT2,
    Array<cfloat, 2> dB1,
    Array<cfloat, 2> dB0,
    TissueParameters p,
    SpiralTrajectory traj,
    Array<cfloat, 2> coils,
    Array<cfloat, 2> vector,
    JacobianComputeMethod kind) {
    ctx.activate();
    const ContractionPrecision precision = effective_precision(jacobian_precision(kind));
    check_spiral_jacobian_inputs(echos, dT1, dT2, dB1, dB0, p, traj, coils);
    const int nr = traj.nreadouts, ns = traj.samples_per_readout, nv = p.nvoxels, nc = int(coils.size(0));
    COMPAS_CHECK(vector.size(0) == 6);
    COMPAS_CHECK(vector.size(1) == nv);
    auto Jv = ctx.empty<cfloat, 3>({nc, nr, ns});
    if (Jv.numel() == 0) return Jv;
    if (nv == 0) {
        ctx.fill_zero(Jv.data(), Jv.nbytes());
        return Jv;
    }
    auto Aa = ctx.empty<cfloat, 2>({nr, nv});
    auto Ab = ctx.empty<cfloat, 2>({nr, nv});
    {
        KernelScope ks(ctx, "spiral_jv_amplitudes");
        dim3 grid(dev::div_ceil(nv, 256), nr);
        kernels::spiral_jv_amplitudes<<<grid, 256, 0, ctx.stream()>>>(
            dev::tissue_view(p), echos.data(), dT1.data(), dT2.data(), dB1.data(), dB0.data(), vector.data(),
            traj.k_start.data(), nr, ns, traj.delta_t, Aa.data(), Ab.data());
        check_launch("spiral_jv_amplitudes");
    }
    spiral_sum(ctx, p, Aa.data(), Ab.data(), coils.data(), nc, traj, precision, Jv.data());
    return Jv;
}

Array<cfloat, 2> compute_jacobian_hermitian(
    const CompasContext& ctx,
    Array<cfloat, 2> echos,
    Array<cfloat, 2> dT1,
    Array<cfloat, 2> dT2,
    Array<cfloat, 2> dB1,
    Array<cfloat, 2> dB0,
    TissueParameters p,
    SpiralTrajectory traj,
    Array<cfloat, 2> coils,
    Array<cfloat, 3> vector,
    JacobianComputeMethod kind) {
    ctx.activate();
    const ContractionPrecision precision = effective_precision(jacobian_precision(kind));
    check_spiral_jacobian_inputs(echos, dT1, dT2, dB1, dB0, p, traj, coils);
    const int nr = traj.nreadouts, ns = traj.samples_per_readout, nv = p.nvoxels, nc = int(coils.size(0));
    COMPAS_CHECK(vector.size(0) == nc);
    COMPAS_CHECK(vector.size(1) == nr);
    COMPAS_CHECK(vector.size(2) == ns);
    auto JHy = ctx.empty<cfloat, 2>({6, nv});
    if (JHy.numel() == 0) return JHy;
    if (nc == 0 || nr == 0 || ns == 0) {
        ctx.fill_zero(JHy.data(), JHy.nbytes());
        return JHy;
    }
    // mixed precision: the sample sums on MFMA (up to eight coils)
    if (spiral_hermitian_mfma(ctx, p, echos.data(), dT1.data(), dT2.data(), dB1.data(), dB0.data(), coils.data(), nc,
                              vector.data(), traj, precision, JHy.data()))
        return JHy;
    // FP32 on VALU; readout chunks so that the grid fills the device: about 8 workgroups per CU
    const int vblocks = int(dev::div_ceil(nv, kernels::JHS_THREADS));
    const int want_chunks = std::max(1, std::min(nr, int(dev::div_ceil(8 * 256, vblocks))));
    const int r_chunk = int(dev::div_ceil(nr, want_chunks));
    const int chunks = int(dev::div_ceil(nr, r_chunk));
    auto partial = ctx.empty<cfloat, 1>({int64_t(chunks) * 6 * nv});
    {
        KernelScope ks(ctx, "spiral_hermitian");
        dim3 grid(vblocks, chunks);
        kernels::spiral_hermitian<<<grid, kernels::JHS_THREADS, 0, ctx.stream()>>>(
            dev::tissue_view(p), echos.data(), dT1.data(), dT2.data(), dB1.data(), dB0.data(), coils.data(), nc,
            vector.data(), traj.k_start.data(), traj.delta_k.data(), nr, ns, traj.delta_t, r_chunk, partial.data());
        check_launch("spiral_hermitian");
    }
    spiral_hermitian_reduce_partials(ctx, partial.data(), chunks, nv, JHy.data());
    return JHy;
}

}  // namespace compas
