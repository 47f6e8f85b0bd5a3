// Jacobian products J.v and J^H.y (Cartesian, parameters T1, T2, rho_x, rho_y).
//
// Semantics: reference src/jacobian/product_kernels.cuh:7-67 (J.v) and
// hermitian_kernels.cuh:30-92 (J^H.y), with the absolute sample decay
// E[s][v] = exp((s - ns/2)(-dt R2 + i Theta_v)), dE/dT2 = (s - ns/2) dt R2^2 E
// of cartesian_view.cuh:65-92.
//
// J.v is one contraction with K = 2 nv:
//   Jv[c][r][s] = sum_v adj_phase[r][v] C[c][v] E[s][v] + adj_decay[r][v] C[c][v] dE[s][v]
// with the adjoint sources of product_kernels.cuh:112-139 as the A operand
// ([nr][2nv]) and the coil-weighted decays as the B operand ([nc*ns][2nv]).
//
// J^H.y: on MFMA precisions the contraction of y against conj(E) / conj(dE)
// and the readout reduction with m, dm/dT1, dm/dT2 are one fused kernel
// (hermitian_mfma.hip).  The FP32 path contracts per coil (M = nr, N = 2 nv,
// K = ns) and reduces over readouts per voxel in a second pass
// (hermitian_kernels.cuh:314-360).

#include "compas/device/common.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

// E and dE/dT2 at (s, v) with the phase formed in FP64
__device__ inline void abs_decay(
    const dev::TissueView& t, int v, int s, int ns, float dt, float dk_re, float dk_im, cfloat* E, cfloat* dE) {
    const float R2 = 1.0f / t.T2(v);
    const double sp = double(s) - 0.5 * double(ns);
    const double theta = double(dk_re) * double(t.x(v)) + double(dk_im) * double(t.y(v))
        + double(dt) * dev::TWO_PI * double(t.B0(v));
    const float mag = expf(float(sp) * (-dt * R2));
    *E = dev::polar_f(mag, dev::reduce_phase(sp * theta));
    *dE = (*E) * (float(sp) * dt * R2 * R2);
}

// A operand of J.v: [nr][2nv] = [adj_phase | adj_decay]
__global__ void jacobian_adjoint_sources(
    dev::TissueView t, const cfloat* __restrict__ m, const cfloat* __restrict__ d1, const cfloat* __restrict__ d2,
    const cfloat* __restrict__ vec, int nr, cfloat* __restrict__ out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    const int nv = t.nvoxels;
    if (v >= nv) return;
    const size_t i = size_t(r) * nv + v;
    const cfloat me = m[i], a1 = d1[i], a2 = d2[i];
    const cfloat rho = t.rho(v);
    const cfloat v0 = vec[v], v1 = vec[nv + v], v2 = vec[2 * size_t(nv) + v], v3 = vec[3 * size_t(nv) + v];
    const cfloat t1r = rho * t.T1(v), t2r = rho * t.T2(v);
    cfloat ph = dev::cmul(dev::cmul(v0, t1r), a1);
    ph = ph + dev::cmul(dev::cmul(v1, t2r), a2);
    ph = ph + dev::cmul(v2, me);
    ph = ph + dev::cmul(dev::cmul(v3, cfloat {0.0f, 1.0f}), me);
    const cfloat dc = dev::cmul(dev::cmul(v1, t2r), me);
    out[size_t(r) * 2 * nv + v] = ph;
    out[size_t(r) * 2 * nv + nv + v] = dc;
}

// B operand of J.v: [nc*ns][2nv] = [C E | C dE]
__global__ void jacobian_coil_decay(
    dev::TissueView t, const cfloat* __restrict__ coils, int nc, int ns, float dt, float dk_re, float dk_im,
    cfloat* __restrict__ out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    const int nv = t.nvoxels;
    if (v >= nv) return;
    cfloat E, dE;
    abs_decay(t, v, s, ns, dt, dk_re, dk_im, &E, &dE);
    for (int c = 0; c < nc; c++) {
        const cfloat C = coils[size_t(c) * nv + v];
        cfloat* row = out + (size_t(c) * ns + s) * 2 * nv;
        row[v] = dev::cmul(C, E);
        row[nv + v] = dev::cmul(C, dE);
    }
}

// B operand of J^H.y: [2nv][ns] = [conj(E)^T ; conj(dE)^T]
__global__ void hermitian_decay(
    dev::TissueView t, int ns, float dt, float dk_re, float dk_im, cfloat* __restrict__ out) {
    const int s = blockIdx.y * blockDim.x + threadIdx.x;
    const int v = blockIdx.x;
    const int nv = t.nvoxels;
    if (s >= ns) return;
    cfloat E, dE;
    abs_decay(t, v, s, ns, dt, dk_re, dk_im, &E, &dE);
    out[size_t(v) * ns + s] = conj(E);
    out[(size_t(nv) + v) * ns + s] = conj(dE);
}

// Per voxel: sum over readouts of the contracted y against conj(m), conj(dm);
// JHy[p][v] (+)= lin_p^* * (...)   (hermitian_kernels.cuh:314-360)
__global__ void hermitian_finalize(
    dev::TissueView t, const cfloat* __restrict__ Ev,  // [nr][2nv]
    const cfloat* __restrict__ m, const cfloat* __restrict__ d1, const cfloat* __restrict__ d2,
    const cfloat* __restrict__ coil, int nr, bool first, cfloat* __restrict__ JHy) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int nv = t.nvoxels;
    if (v >= nv) return;
    cfloat mHv {0, 0}, g0 {0, 0}, g1 {0, 0};
    for (int r = 0; r < nr; r++) {
        const size_t i = size_t(r) * nv + v;
        const cfloat e = Ev[size_t(r) * 2 * nv + v];
        const cfloat de = Ev[size_t(r) * 2 * nv + nv + v];
        const cfloat me = m[i];
        mHv = dev::cfma_conj(e, me, mHv);
        g0 = dev::cfma_conj(e, d1[i], g0);
        g1 = dev::cfma_conj(e, d2[i], g1);
        g1 = dev::cfma_conj(de, me, g1);
    }
    const cfloat C = coil[v];
    const cfloat rho = t.rho(v);
    const cfloat l0 = conj(dev::cmul(rho * t.T1(v), C));
    const cfloat l1 = conj(dev::cmul(rho * t.T2(v), C));
    const cfloat cc = conj(C);
    const cfloat o0 = dev::cmul(l0, g0);
    const cfloat o1 = dev::cmul(l1, g1);
    const cfloat o2 = dev::cmul(cc, mHv);
    const cfloat o3 = dev::cmul(o2, cfloat {0.0f, -1.0f});
    if (first) {
        JHy[v] = o0;
        JHy[size_t(nv) + v] = o1;
        JHy[2 * size_t(nv) + v] = o2;
        JHy[3 * size_t(nv) + v] = o3;
    } else {
        JHy[v] += o0;
        JHy[size_t(nv) + v] += o1;
        JHy[2 * size_t(nv) + v] += o2;
        JHy[3 * size_t(nv) + v] += o3;
    }
}

}  // namespace kernels

static void check_jacobian_inputs(
    const Array<cfloat, 2>& echos, const Array<cfloat, 2>& d1, const Array<cfloat, 2>& d2, int nr, int nv,
    const Array<cfloat, 2>& coils) {
    COMPAS_CHECK(echos.size(0) == nr);
    COMPAS_CHECK(echos.size(1) == nv);
    COMPAS_CHECK(d1.size(0) == nr);
    COMPAS_CHECK(d1.size(1) == nv);
    COMPAS_CHECK(d2.size(0) == nr);
    COMPAS_CHECK(d2.size(1) == nv);
    COMPAS_CHECK(coils.size(1) == nv);
}

Array<cfloat, 3> compute_jacobian(
    const CompasContext& ctx,
    Array<cfloat, 2> echos,
    Array<cfloat, 2> dT1,
    Array<cfloat, 2> dT2,
    TissueParameters p,
    CartesianTrajectory traj,
    Array<cfloat, 2> coils,
    Array<cfloat, 2> vector,
    JacobianComputeMethod kind) {
    ctx.activate();
    const int ns = traj.samples_per_readout, nr = traj.nreadouts, nv = p.nvoxels, nc = int(coils.size(0));
    check_jacobian_inputs(echos, dT1, dT2, nr, nv, coils);
    COMPAS_CHECK(vector.size(0) == 4);  // product.cu:312
    COMPAS_CHECK(vector.size(1) == nv);
    const ContractionPrecision prec = jacobian_precision(kind);

    auto Jv = ctx.empty<cfloat, 3>({nc, nr, ns});
    if (Jv.numel() == 0) return Jv;
    if (nv == 0) {
        ctx.fill_zero(Jv.data(), Jv.nbytes());
        return Jv;
    }
    auto Aop = ctx.empty<cfloat, 2>({nr, 2 * int64_t(nv)});
    auto Bop = ctx.empty<cfloat, 2>({int64_t(nc) * ns, 2 * int64_t(nv)});
    const auto tv = dev::tissue_view(p);
    {
        KernelScope ks(ctx, "jacobian_adjoint_sources");
        dim3 grid(dev::div_ceil(nv, 256), nr);
        kernels::jacobian_adjoint_sources<<<grid, 256, 0, ctx.stream()>>>(
            tv, echos.data(), dT1.data(), dT2.data(), vector.data(), nr, Aop.data());
        check_launch("jacobian_adjoint_sources");
    }
    {
        KernelScope ks(ctx, "jacobian_coil_decay");
        dim3 grid(dev::div_ceil(nv, 256), ns);
        kernels::jacobian_coil_decay<<<grid, 256, 0, ctx.stream()>>>(
            tv, coils.data(), nc, ns, traj.delta_t, traj.delta_k.re, traj.delta_k.im, Bop.data());
        check_launch("jacobian_coil_decay");
    }
    ContractionShape s {nr, nc * ns, 2 * nv, 2 * int64_t(nv), 2 * int64_t(nv), ns, int64_t(nr) * ns, false};
    complex_contraction(ctx, Aop.data(), Bop.data(), Jv.data(), s, prec);
    return Jv;
}

Array<cfloat, 2> compute_jacobian_hermitian(
    const CompasContext& ctx,
    Array<cfloat, 2> echos,
    Array<cfloat, 2> dT1,
    Array<cfloat, 2> dT2,
    TissueParameters p,
    CartesianTrajectory traj,
    Array<cfloat, 2> coils,
    Array<cfloat, 3> vector,
    JacobianComputeMethod kind) {
    ctx.activate();
    const int ns = traj.samples_per_readout, nr = traj.nreadouts, nv = p.nvoxels, nc = int(coils.size(0));
    check_jacobian_inputs(echos, dT1, dT2, nr, nv, coils);
    COMPAS_CHECK(vector.size(0) == nc);  // hermitian.cu:273-275
    COMPAS_CHECK(vector.size(1) == nr);
    COMPAS_CHECK(vector.size(2) == ns);
    const ContractionPrecision prec = jacobian_precision(kind);

    auto JHy = ctx.empty<cfloat, 2>({4, nv});
    if (JHy.numel() == 0) return JHy;
    if (nc == 0 || nr == 0 || ns == 0) {
        ctx.fill_zero(JHy.data(), JHy.nbytes());
        return JHy;
    }
    const auto tv = dev::tissue_view(p);
    auto EH = ctx.empty<cfloat, 2>({2 * int64_t(nv), ns});
    {
        KernelScope ks(ctx, "hermitian_decay");
        dim3 grid(nv, dev::div_ceil(ns, 64));
        kernels::hermitian_decay<<<grid, 64, 0, ctx.stream()>>>(
            tv, ns, traj.delta_t, traj.delta_k.re, traj.delta_k.im, EH.data());
        check_launch("hermitian_decay");
    }
    if (jacobian_hermitian_mfma(ctx, p, vector.data(), EH.data(), echos.data(), dT1.data(), dT2.data(),
                                coils.data(), nc, nr, ns, effective_precision(prec), JHy.data()))
        return JHy;
    auto Ev = ctx.empty<cfloat, 2>({nr, 2 * int64_t(nv)});
    for (int c = 0; c < nc; c++) {
        ContractionShape s {nr, 2 * nv, ns, ns, ns, 2 * nv, 0, false};
        complex_contraction(ctx, vector.data() + size_t(c) * nr * ns, EH.data(), Ev.data(), s, prec);
        KernelScope ks(ctx, "hermitian_finalize");
        kernels::hermitian_finalize<<<dev::div_ceil(nv, 256), 256, 0, ctx.stream()>>>(
            tv, Ev.data(), echos.data(), dT1.data(), dT2.data(), coils.data() + size_t(c) * nv, nr, c == 0,
            JHy.data());
        check_launch("hermitian_finalize");
    }
    return JHy;
}

}  // namespace compas
