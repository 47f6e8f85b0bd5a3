#pragma once
// Internal building blocks shared by the compute entry points.

#include <algorithm>
#include <cstdint>

#include "compas/compas.h"

namespace compas {

namespace kernels {
struct FispArgs;
struct PssfpTableJob;
}  // namespace kernels

// dst[i] += src[i] for i < n (complex)
void add_into(const CompasContext& ctx, cfloat* dst, const cfloat* src, int64_t n);

// Operand precision of the complex contraction engine.
enum struct ContractionPrecision {
    FP32,  // FP32 VALU (no MFMA)
    BF16x3,  // BF16 hi+lo split, 3 MFMA products, FP32 accumulate
    FP16x3,  // FP16 hi+lo split with per-row power-of-two scaling, 3 MFMA products
    BF16x1,  // BF16 single product
    FP16x1,  // FP16 single product with per-row power-of-two scaling
};

// Process-wide override of the precision that the methods select (for
// side-by-side accuracy/speed studies).  Values follow ContractionPrecision's
// order; -1 clears the override.
void set_contraction_precision_override(int precision);

// Complex contraction C = A * B^T with both operands K-contiguous:
//   C[m][n] = sum_k A[m*lda + k] * B[n*ldb + k]      (complex FP32 operands)
// Output element (m, n) is written to
//   out[(n / nsub) * out_outer + m * nsub + (n % nsub)]
// which covers both plain row-major C (nsub = N) and coil-major signal layouts
// ([nc][nr][ns] with n = c*ns + s, nsub = ns, out_outer = M*ns).  Partial sums
// over K splits are reduced in FP64.  `accumulate` adds into `out`.
struct ContractionShape {
    int M, N, K;
    int64_t lda, ldb;
    int nsub;
    int64_t out_outer;
    bool accumulate;
};

void complex_contraction(
    const CompasContext& ctx,
    const cfloat* A,
    const cfloat* B,
    cfloat* out,
    ContractionShape shape,
    ContractionPrecision precision);

// complex_contraction with the B operand generated in the kernel's staging
// instead of read: B[c ns + s][v] = C[c][v] T[s / 4][v] q[v]^(s % 4) with C
// [N / ns][K], T [ns / 4][K], q [K] (ns a multiple of 4).  BF16 precisions
// only: returns false, doing nothing, for the others.
bool complex_contraction_generated(
    const CompasContext& ctx, const cfloat* A, const cfloat* C, const cfloat* T, const cfloat* q, int ns,
    cfloat* out, ContractionShape shape, ContractionPrecision precision);

// MFMA tile variants: force one (v < 0 restores autotuning), count them, and
// report the variant chosen per tuned shape.
void set_contraction_variant(int v);
int contraction_variant_count();
std::string contraction_tuning_report();

// Copies `nbytes` between two device buffers with a plain streaming kernel
// (after one warm-up) and returns the elapsed milliseconds.
float bandwidth_probe(const CompasContext& ctx, int64_t nbytes);

// Precision after the process-wide override (if one is set).
ContractionPrecision effective_precision(ContractionPrecision requested);

// Per-row power-of-two scales s with max_k |x[row][k]| s ~ 2^14 for FP16
// operands of `rows` rows of length K with row stride ld.
void row_pow2_scales(const CompasContext& ctx, const cfloat* X, int rows, int K, int64_t ld, float* scale);

// J^H.y with the contraction and readout reduction fused on MFMA.  y
// [nc][nr][ns]; EH [2 nv][ns] = conj(E) rows then conj(dE/dT2) rows; m, d1, d2
// [nr][nv]; coils [nc][nv]; writes JHy [4][nv].  Returns false (doing
// nothing) when `precision` is not an MFMA precision or nc < 3.
bool jacobian_hermitian_mfma(
    const CompasContext& ctx, const TissueParameters& p, const cfloat* y, const cfloat* EH, const cfloat* m,
    const cfloat* d1, const cfloat* d2, const cfloat* coils, int nc, int nr, int ns,
    ContractionPrecision precision, cfloat* JHy);

// Spiral sum out[c][r][s] = sum_v C[c][v] (A[r][v] + (s - ns/2) Ab[r][v]) E_rv^s
// with the per-readout sample exponent of the spiral trajectory (Ab may be
// nullptr).  A and Ab already carry the readout factor (rewind, prephaser).
void spiral_sum(
    const CompasContext& ctx, const TissueParameters& p, const cfloat* A, const cfloat* Ab, const cfloat* coils,
    int nc, const SpiralTrajectory& t, ContractionPrecision precision, cfloat* out);

// The same sums with the voxel contraction on MFMA (BF16 hi + lo, or hi only);
// false when the precision is FP32 or the coils exceed one 16-row group
// (eight with one amplitude, four with two): the caller runs the VALU form.
bool spiral_sum_mfma(
    const CompasContext& ctx, const TissueParameters& p, const cfloat* A, const cfloat* Ab, const cfloat* coils,
    int nc, const SpiralTrajectory& t, ContractionPrecision precision, cfloat* out);

// Per (slice, TR) RF table of the FISP kernels: (deg2rad |RF|, cos phi, sin phi, 0).
Array<float4, 1> fisp_rf_table(const CompasContext& ctx, const FISPSequence& seq);

// simulate_magnetization(FISP) into the caller's device buffer `echos`
// ([nTR * usf][nvoxels], `capacity` complex values) without allocating: the
// RF table goes to a static per-GPU device array of 65536 (slice, TR)
// entries.  Calls that overlap on one GPU race on that array.
void simulate_magnetization_into(
    const CompasContext& ctx, const TissueParameters& parameters, const FISPSequence& seq, cfloat* echos,
    int64_t capacity);

// pSSFP launch layout: ZPL slice positions per lane, G = ceil(nz / ZPL) wave
// groups per block of 64 voxels, at most per_launch groups per launch (the
// partial echoes of groups 1.. are held in one buffer of <= 8 GiB).
// waves_per_cu[zpl] is the resident waves per CU of the ZPL instance.
struct PssfpLayout {
    int zpl, groups, per_launch;
};
PssfpLayout pssfp_layout(const CompasContext& ctx, int nz, int nv, int nTR, const int* waves_per_cu, int max_zpl);

// J^H.y on the spiral trajectory with the sample sums on MFMA (BF16 hi + lo,
// or hi only for the single-product precisions); false when the precision is
// FP32 or there are more than eight coils (the caller runs the VALU form).
bool spiral_hermitian_mfma(
    const CompasContext& ctx, const TissueParameters& p, const cfloat* m, const cfloat* d1, const cfloat* d2,
    const cfloat* db1, const cfloat* db0, const cfloat* coils, int nc, const cfloat* y, const SpiralTrajectory& traj,
    ContractionPrecision precision, cfloat* JHy);

// JHy[p][v] = sum over `chunks` partial [6][nv] planes (FP64)
void spiral_hermitian_reduce_partials(const CompasContext& ctx, const cfloat* partial, int chunks, int nv, cfloat* JHy);

// device bytes of partial echoes one pSSFP simulation of this size allocates
int64_t pssfp_scratch_bytes(const CompasContext& ctx, int nz, int nv, int nTR);

// echos[i] += sum over count partial planes of n elements (plane j at
// partial + j * stride), in FP64, in order
void pssfp_sum_groups(
    const CompasContext& ctx, cfloat* echos, const cfloat* partial, int count, int64_t n, int64_t stride);

// Work-table launches of the request server: workgroup b runs items[b].
//   FISP: item (job, voxel block of 64), every job with max_state == spl,
//   fisp_table_supported(spl); one lane per voxel.
//   pSSFP: item (job, voxel block, group, -), slice positions per lane
//   zpl = pssfp_table_zpl(nz) of every job (a function of nz alone, so a
//   request's result does not depend on what else shares the launch).
bool fisp_table_supported(int max_state);
void launch_fisp_table(const CompasContext& ctx, int spl, const kernels::FispArgs* jobs, const int2* items, int n);
int pssfp_table_zpl(int nz);
void launch_pssfp_table(
    const CompasContext& ctx, int zpl, const kernels::PssfpTableJob* jobs, const int4* items, int n);

// Finite-difference pieces of simulate_magnetization_derivative
// (derivative.cu:11-94): the parameters with `field` moved by `delta` (the
// presence flags kept), and (next - echos) / delta, written over next (which
// is returned: the difference needs no memory beyond the perturbed echoes).
TissueParameters perturbed_parameters(const CompasContext& ctx, const TissueParameters& p, int field, float delta);
Array<cfloat, 2> finite_difference(
    const CompasContext& ctx, const Array<cfloat, 2>& echos, Array<cfloat, 2> next, float delta);

ContractionPrecision signal_precision(SimulateSignalMethod method);
ContractionPrecision jacobian_precision(JacobianComputeMethod method);

}  // namespace compas
