// Element-wise kernels: zero fill, accumulation, finite-difference
// derivative (reference src/simulate/derivative.cu:11-94,
// derivative_kernels.cuh:6-35) and the residual of the reconstruction
// objective (src/utils/residual.cu:8-46, residual_kernels.cuh:8-64).

#include "compas/device/common.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

// Zero fill by a kernel (not hipMemsetAsync) so callers may capture it in graphs.
__global__ void fill_zero_u32(uint32_t* p, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        p[i] = 0u;
}

__global__ void add_into_kernel(cfloat* dst, const cfloat* src, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        dst[i] += src[i];
}

__global__ void add_difference_to_parameters(
    const float* __restrict__ in, float* __restrict__ out, int nv, int field, float delta) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    for (int f = 0; f < TissueParameterField::NUM_FIELDS; f++) {
        out[size_t(f) * nv + v] = in[size_t(f) * nv + v];
    }
    out[size_t(field) * nv + v] += delta;
}

// e1 <- (e1 - e0) / delta, in place
__global__ void finite_difference(const cfloat* __restrict__ e0, cfloat* e1, int64_t n, float inv_delta) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const cfloat a = e0[i], b = e1[i];
        e1[i] = {(b.re - a.re) * inv_delta, (b.im - a.im) * inv_delta};
    }
}

// diff = lhs - rhs; per-block partial sums of |diff|^2 (FP64)
__global__ void __launch_bounds__(256) residual_kernel(
    const cfloat* __restrict__ lhs, const cfloat* __restrict__ rhs, int64_t n, cfloat* __restrict__ diff,
    double* __restrict__ partial) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const cfloat d = {lhs[i].re - rhs[i].re, lhs[i].im - rhs[i].im};
        diff[i] = d;
        acc += double(d.re) * d.re + double(d.im) * d.im;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) sum_partials(const double* __restrict__ partial, int n, float* out) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) acc += partial[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = float(red[0]);
}

// Plain streaming copy (16-byte accesses): the bandwidth probe and the
// reference point that scales hardware traffic counters to bytes.
__global__ void __launch_bounds__(256) stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

}  // namespace kernels

float bandwidth_probe(const CompasContext& ctx, int64_t nbytes) {
    ctx.activate();
    const int64_t n = std::max<int64_t>(1, nbytes / 16);
    auto a = ctx.empty<uint4, 1>({n});
    auto b = ctx.empty<uint4, 1>({n});
    const unsigned blocks = unsigned(std::min<int64_t>((n + 255) / 256, 256 * 64));
    hipEvent_t e0, e1;
    COMPAS_HIP_CHECK(hipEventCreate(&e0));
    COMPAS_HIP_CHECK(hipEventCreate(&e1));
    kernels::stream_copy<<<blocks, 256, 0, ctx.stream()>>>(a.data(), b.data(), n);  // warm-up
    COMPAS_HIP_CHECK(hipEventRecord(e0, ctx.stream()));
    {
        KernelScope ks(ctx, "stream_copy");
        kernels::stream_copy<<<blocks, 256, 0, ctx.stream()>>>(a.data(), b.data(), n);
        check_launch("stream_copy");
    }
    COMPAS_HIP_CHECK(hipEventRecord(e1, ctx.stream()));
    COMPAS_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    COMPAS_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
}

static unsigned grid_for(int64_t n) {
    return unsigned(std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), 4096));
}

void CompasContext::fill_zero(void* ptr, size_t nbytes) const {
    if (nbytes == 0) return;
    COMPAS_CHECK(nbytes % 4 == 0);
    activate();
    const int64_t n = int64_t(nbytes / 4);
    KernelScope ks(*this, "fill_zero");
    kernels::fill_zero_u32<<<grid_for(n), 256, 0, stream()>>>(static_cast<uint32_t*>(ptr), n);
    check_launch("fill_zero");
}

void add_into(const CompasContext& ctx, cfloat* dst, const cfloat* src, int64_t n) {
    if (n == 0) return;
    KernelScope ks(ctx, "add_into");
    kernels::add_into_kernel<<<grid_for(n), 256, 0, ctx.stream()>>>(dst, src, n);
    check_launch("add_into");
}

TissueParameters perturbed_parameters(const CompasContext& ctx, const TissueParameters& p, int field, float delta) {
    ctx.activate();
    const int nv = p.nvoxels;
    COMPAS_CHECK(field >= 0 && field < TissueParameterField::NUM_FIELDS);  // derivative.cu:25
    auto data = ctx.empty<float, 2>({TissueParameterField::NUM_FIELDS, nv});
    if (nv > 0) {
        KernelScope ks(ctx, "add_difference_to_parameters");
        kernels::add_difference_to_parameters<<<dev::div_ceil(nv, 256), 256, 0, ctx.stream()>>>(
            p.data.data(), data.data(), nv, field, delta);
        check_launch("add_difference_to_parameters");
    }
    // the perturbed parameters keep the original presence flags
    // (derivative.cu:43-45): a field the map does not hold (B1, B0 or z absent)
    // keeps its default in the simulation, so its perturbation has no effect and
    // the difference is (sim(p) - echos) / delta, zero for echoes of the same map
    TissueParameters q = p;
    q.data = data;
    return q;
}

Array<cfloat, 2> finite_difference(
    const CompasContext& ctx, const Array<cfloat, 2>& echos, Array<cfloat, 2> next, float delta) {
    ctx.activate();
    COMPAS_CHECK(next.size(0) == echos.size(0) && next.size(1) == echos.size(1));
    COMPAS_CHECK(next.data() != echos.data() || next.numel() == 0);
    const int64_t n = next.numel();
    if (n > 0) {
        KernelScope ks(ctx, "finite_difference");
        kernels::finite_difference<<<grid_for(n), 256, 0, ctx.stream()>>>(echos.data(), next.data(), n,
                                                                           1.0f / delta);
        check_launch("finite_difference");
    }
    return next;
}

template<typename Sequence>
static Array<cfloat, 2> derivative_impl(
    const CompasContext& ctx, int field, Array<cfloat, 2> echos, TissueParameters p, Sequence seq, float delta) {
    COMPAS_CHECK(echos.size(1) == p.nvoxels);  // derivative.cu:24
    auto q = perturbed_parameters(ctx, p, field, delta);
    auto next = simulate_magnetization(ctx, q, seq);
    return finite_difference(ctx, echos, next, delta);
}

Array<cfloat, 2> simulate_magnetization_derivative(
    const CompasContext& ctx, int field, Array<cfloat, 2> echos, TissueParameters p, FISPSequence seq, float delta) {
    return derivative_impl(ctx, field, echos, p, seq, delta);
}

Array<cfloat, 2> simulate_magnetization_derivative(
    const CompasContext& ctx, int field, Array<cfloat, 2> echos, TissueParameters p, pSSFPSequence seq, float delta) {
    return derivative_impl(ctx, field, echos, p, seq, delta);
}

Array<cfloat, 3> compute_residual(
    const CompasContext& ctx, Array<cfloat, 3> lhs, Array<cfloat, 3> rhs, float* objective_out) {
    ctx.activate();
    COMPAS_CHECK(lhs.sizes() == rhs.sizes());  // residual.cu:14
    auto diff = ctx.empty<cfloat, 3>(lhs.sizes());
    const int64_t n = lhs.numel();
    const unsigned blocks = unsigned(std::min<int64_t>(1024, std::max<int64_t>(1, (n + 255) / 256)));
    auto partial = ctx.empty<double, 1>({int64_t(blocks)});
    auto objective = ctx.empty<float, 1>({1});
    {
        KernelScope ks(ctx, "residual");
        kernels::residual_kernel<<<blocks, 256, 0, ctx.stream()>>>(lhs.data(), rhs.data(), n, diff.data(), partial.data());
        check_launch("residual");
        kernels::sum_partials<<<1, 256, 0, ctx.stream()>>>(partial.data(), int(blocks), objective.data());
        check_launch("sum_partials");
    }
    if (objective_out) objective.copy_to(objective_out);
    return diff;
}

}  // namespace compas
