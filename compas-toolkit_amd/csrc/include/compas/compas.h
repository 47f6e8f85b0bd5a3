#pragma once
// compas:: C++ API of the simulation hot path.
//
// Names, argument order and meaning follow the reference's public headers:
//   include/compas/parameters/tissue.h:16-67, tissue_view.cuh:13-26
//   include/compas/sequences/fisp.h:10-60, pssfp.h:9-72
//   include/compas/trajectories/{trajectory,cartesian,spiral,signal,phase_encoding}.h
//   include/compas/simulate/{fisp,pssfp,derivative}.h
//   include/compas/jacobian/{product,hermitian}.h
//   include/compas/utils/residual.h
// Every compute function enqueues work on the context's stream and returns an
// Array whose contents are valid once the stream reaches that point
// (synchronize(), or any copy_to()).

#include <vector>

#include "compas/core/complex.h"
#include "compas/core/context.h"

namespace compas {

struct Object {
    virtual ~Object() = default;
};

// ---------------------------------------------------------------- tissue --

namespace TissueParameterField {
enum {
    T1 = 0,
    T2 = 1,
    B1 = 2,
    B0 = 3,
    RHO_X = 4,
    RHO_Y = 5,
    X = 6,
    Y = 7,
    Z = 8,
    NUM_FIELDS = 9,
};
}

// Structure-of-arrays [NUM_FIELDS][nvoxels] float32 in HBM: each field is a
// contiguous row, so kernels read every field with coalesced loads.
struct TissueParameters: Object {
    Array<float, 2> data;
    int nvoxels = 0;
    int chunk_size = 0;
    bool has_z = true;
    bool has_b0 = true;
    bool has_b1 = true;

    TissueParameters() = default;
    TissueParameters(Array<float, 2> d, bool has_z, bool has_b0, bool has_b1) :
        data(d),
        nvoxels(int(d.size(1))),
        chunk_size(int(d.size(1))),
        has_z(has_z),
        has_b0(has_b0),
        has_b1(has_b1) {}
};

TissueParameters make_tissue_parameters(
    const CompasContext& ctx,
    int num_voxels,
    int chunk_size,
    View<float> T1,
    View<float> T2,
    View<float> B1,
    View<float> B0,
    View<float> rho_x,
    View<float> rho_y,
    View<float> x,
    View<float> y,
    View<float> z = {});

// ------------------------------------------------------------- sequences --

struct FISPSequence: Object {
    Array<cfloat> RF_train;  // [nTR], |RF| = flip angle in degrees, arg = phase
    Array<cfloat, 2> sliceprofiles;  // [nslices][nTR]
    float TR = 0;
    float TE = 0;
    float TW = 0;
    int max_state = 0;
    float TI = 0;
    int undersampling_factor = 1;
    int repetitions = 1;
    bool inversion_prepulse = true;
    bool wait_spoiling = true;

    FISPSequence() = default;
    FISPSequence(
        Array<cfloat> RF,
        Array<cfloat, 2> profiles,
        float TR,
        float TE,
        float TW,
        int max_state,
        float TI,
        int undersampling_factor = 1,
        int repetitions = 1,
        bool inversion_prepulse = true,
        bool wait_spoiling = true) :
        RF_train(RF),
        sliceprofiles(profiles),
        TR(TR),
        TE(TE),
        TW(TW),
        max_state(max_state),
        TI(TI),
        undersampling_factor(undersampling_factor),
        repetitions(repetitions),
        inversion_prepulse(inversion_prepulse),
        wait_spoiling(wait_spoiling) {}
};

FISPSequence make_fisp_sequence(
    const CompasContext& context,
    View<cfloat> RF_train,
    View<cfloat, 2> sliceprofiles,
    float TR,
    float TE,
    int max_state,
    float TI);

struct RepetitionData {
    float ex;
    float inv;
    float pr;
};

struct pSSFPSequence: Object {
    Array<cfloat> RF_train;  // [nTR]
    float TR = 0;  // stored, not used by the simulation (as in the reference)
    int nRF = 0;
    int nTR = 0;
    Array<cfloat> gamma_dt_RF;  // [nRF], waveform normalised to 1 degree
    RepetitionData dt {};
    RepetitionData gamma_dt_GRz {};
    int nz = 0;
    Array<float> z;  // [nz] slice positions

    pSSFPSequence() = default;
    pSSFPSequence(
        Array<cfloat> RF_train,
        float TR,
        Array<cfloat> gamma_dt_RF,
        RepetitionData dt,
        RepetitionData gamma_dt_GRz,
        Array<float> z) :
        RF_train(RF_train),
        TR(TR),
        nRF(int(gamma_dt_RF.size(0))),
        nTR(int(RF_train.size(0))),
        gamma_dt_RF(gamma_dt_RF),
        dt(dt),
        gamma_dt_GRz(gamma_dt_GRz),
        nz(int(z.size(0))),
        z(z) {}
};

pSSFPSequence make_pssfp_sequence(
    const CompasContext& context,
    View<cfloat> RF_train,
    float TR,
    View<cfloat> gamma_dt_RF,
    RepetitionData dt,
    RepetitionData gamma_dt_GRz,
    View<float> z);

// ---------------------------------------------------------- trajectories --

struct Trajectory: Object {
    int nreadouts = 0;
    int samples_per_readout = 0;
    float delta_t = 0;

    Trajectory(int nr, int ns, float dt) : nreadouts(nr), samples_per_readout(ns), delta_t(dt) {}
};

struct CartesianTrajectory: Trajectory {
    Array<cfloat> k_start;  // [nreadouts]; used by phase_encoding only
    cfloat delta_k;

    CartesianTrajectory(int nr, int ns, float dt, Array<cfloat> k_start, cfloat delta_k) :
        Trajectory(nr, ns, dt),
        k_start(k_start),
        delta_k(delta_k) {}
};

struct SpiralTrajectory: Trajectory {
    Array<cfloat> k_start;  // [nreadouts]
    Array<cfloat> delta_k;  // [nreadouts]

    SpiralTrajectory(int nr, int ns, float dt, Array<cfloat> k_start, Array<cfloat> delta_k) :
        Trajectory(nr, ns, dt),
        k_start(k_start),
        delta_k(delta_k) {}
};

CartesianTrajectory make_cartesian_trajectory(
    const CompasContext& context,
    int nreadouts,
    int samples_per_readout,
    float delta_t,
    HostView<cfloat> k_start,
    cfloat delta_k);

SpiralTrajectory make_spiral_trajectory(
    const CompasContext& context,
    int nreadouts,
    int samples_per_readout,
    float delta_t,
    View<cfloat> k_start,
    View<cfloat> delta_k);

// -------------------------------------------------------------- simulate --

Array<cfloat, 2> simulate_magnetization(
    const CompasContext& context,
    TissueParameters parameters,
    FISPSequence sequence);

Array<cfloat, 2> simulate_magnetization(
    const CompasContext& context,
    TissueParameters parameters,
    pSSFPSequence sequence);

static constexpr float DEFAULT_FINITE_DIFFERENCE_DELTA = 1e-4F;

Array<cfloat, 2> simulate_magnetization_derivative(
    const CompasContext& context,
    int field,
    Array<cfloat, 2> echos,
    TissueParameters parameters,
    FISPSequence sequence,
    float delta = DEFAULT_FINITE_DIFFERENCE_DELTA);

Array<cfloat, 2> simulate_magnetization_derivative(
    const CompasContext& context,
    int field,
    Array<cfloat, 2> echos,
    TissueParameters parameters,
    pSSFPSequence sequence,
    float delta = DEFAULT_FINITE_DIFFERENCE_DELTA);

// Forward-mode derivative of the FISP echoes with respect to one tissue
// parameter (field = T1, T2, B1 or B0 of TissueParameterField): the exact
// derivative of the discrete EPG recursion, computed beside it in one pass.
// No counterpart in the reference, which differentiates by finite
// differences (simulate_magnetization_derivative, derivative.h:21-47).
Array<cfloat, 2> simulate_magnetization_tangent(
    const CompasContext& ctx, int field, TissueParameters parameters, FISPSequence sequence);

// The same for the pSSFP isochromat simulation (FP64 state, FP32 result).
Array<cfloat, 2> simulate_magnetization_tangent(
    const CompasContext& ctx, int field, TissueParameters parameters, pSSFPSequence sequence);

// Derivatives with respect to several fields (distinct T1, T2, B1, B0), one
// array per field in the order given; with echos != nullptr the echoes too.
// pSSFP: T1 + T2 and all four fields (each with or without the echoes) run as
// one fused pass sharing the primal recursion; other sets as several passes.
// FISP: one pass per field (and the plain simulation for the echoes).
std::vector<Array<cfloat, 2>> simulate_magnetization_tangents(
    const CompasContext& ctx,
    const std::vector<int>& fields,
    TissueParameters parameters,
    pSSFPSequence sequence,
    Array<cfloat, 2>* echos = nullptr);
std::vector<Array<cfloat, 2>> simulate_magnetization_tangents(
    const CompasContext& ctx,
    const std::vector<int>& fields,
    TissueParameters parameters,
    FISPSequence sequence,
    Array<cfloat, 2>* echos = nullptr);

// ---------------------------------------------------------------- signal --

// Same enumerators as the reference (signal.h:14-44).  On MI355X:
//   Naive          one output per thread, voxel loop (reference semantics check)
//   Direct         FP32 VALU contraction (the plain, MFMA-free path)
//   MatmulPedantic FP16 hi+lo split operands on MFMA, FP32 accumulate, scaled
//   Matmul         BF16 hi+lo split operands on MFMA (3 products)
//   MatmulFast     same as Matmul (default)
//   MatmulLow      BF16 single product on MFMA (low accuracy, fastest)
enum struct SimulateSignalMethod { Naive, Direct, MatmulPedantic, Matmul, MatmulFast, MatmulLow };

Array<cfloat, 3> magnetization_to_signal(
    const CompasContext& context,
    Array<cfloat, 2> echos,
    TissueParameters parameters,
    const Trajectory& trajectory,
    Array<cfloat, 2> coil_sensitivities,
    SimulateSignalMethod method = SimulateSignalMethod::MatmulFast);

Array<cfloat, 2> phase_encoding(
    const CompasContext& ctx,
    const Array<cfloat, 2>& echos,
    const TissueParameters& parameters,
    const CartesianTrajectory& trajectory);

// -------------------------------------------------------------- jacobian --

// Same enumerators as the reference (product.h:10).  Naive is the per-output
// loop; Direct and Gemm run the FP32 VALU contraction; GemmFast the BF16
// hi+lo MFMA contraction; GemmLow single-product BF16.
enum struct JacobianComputeMethod { Naive, Direct, Gemm, GemmFast, GemmLow };

Array<cfloat, 3> compute_jacobian(
    const CompasContext& ctx,
    Array<cfloat, 2> echos,
    Array<cfloat, 2> delta_echos_T1,
    Array<cfloat, 2> delta_echos_T2,
    TissueParameters parameters,
    CartesianTrajectory trajectory,
    Array<cfloat, 2> coil_sensitivities,
    Array<cfloat, 2> vector,
    JacobianComputeMethod kind = JacobianComputeMethod::Direct);

Array<cfloat, 2> compute_jacobian_hermitian(
    const CompasContext& ctx,
    Array<cfloat, 2> echos,
    Array<cfloat, 2> delta_echos_T1,
    Array<cfloat, 2> delta_echos_T2,
    TissueParameters parameters,
    CartesianTrajectory trajectory,
    Array<cfloat, 2> coil_sensitivities,
    Array<cfloat, 3> vector,
    JacobianComputeMethod kind = JacobianComputeMethod::Direct);

// Spiral trajectory, six parameters (no counterpart in the reference, whose
// Jacobian is Cartesian with four): rows of `vector` and of the J^H.y result
// are (T1, T2, B1, B0, rho_x, rho_y); the T rows are scaled by T1 and T2 as in
// the Cartesian products.  vector is [6][nv] for J.v and [nc][nr][ns] for
// J^H.y.  The mixed-precision methods run the spiral sums on MFMA (BF16
// hi + lo, FP32 accumulation); Naive, Direct and Gemm run them on FP32 VALU.
Array<cfloat, 3> compute_jacobian(
    const CompasContext& ctx,
    Array<cfloat, 2> echos,
    Array<cfloat, 2> delta_echos_T1,
    Array<cfloat, 2> delta_echos_T2,
    Array<cfloat, 2> delta_echos_B1,
    Array<cfloat, 2> delta_echos_B0,
    TissueParameters parameters,
    SpiralTrajectory trajectory,
    Array<cfloat, 2> coil_sensitivities,
    Array<cfloat, 2> vector,
    JacobianComputeMethod kind = JacobianComputeMethod::Direct);

Array<cfloat, 2> compute_jacobian_hermitian(
    const CompasContext& ctx,
    Array<cfloat, 2> echos,
    Array<cfloat, 2> delta_echos_T1,
    Array<cfloat, 2> delta_echos_T2,
    Array<cfloat, 2> delta_echos_B1,
    Array<cfloat, 2> delta_echos_B0,
    TissueParameters parameters,
    SpiralTrajectory trajectory,
    Array<cfloat, 2> coil_sensitivities,
    Array<cfloat, 3> vector,
    JacobianComputeMethod kind = JacobianComputeMethod::Direct);

// ---------------------------------------------------------------- utils --

Array<cfloat, 3> compute_residual(
    const CompasContext& ctx,
    Array<cfloat, 3> lhs,
    Array<cfloat, 3> rhs,
    float* objective_out = nullptr);

// GEMM utility (include/compas/utils/gemm.h:7-71): result = alpha * lhs *
// transpose(rhs) + beta * result with lhs [m][k], rhs [n][k], result [m][n]
// (complex: planar [2][..][..], real plane first).  Arrays are contiguous and
// the result is written in place.  Precision per method for FP32 operands:
// Pedantic = FP32 FMA on VALU, Regular = BF16 hi + lo on MFMA (three products,
// FP32-level), Fast = BF16 on MFMA (one product).  BF16 operands: exact
// products, FP32 accumulation, for every method.
enum struct GemmComputeMethod {
    Pedantic,  // highest precision
    Regular,  // balanced
    Fast  // highest performance
};

using bfloat16_t = __bf16;

void compute_gemm(
    const CompasContext& ctx,
    Array<float, 2> result,
    Array<float, 2> lhs,
    Array<float, 2> rhs,
    float alpha,
    float beta,
    GemmComputeMethod kind = GemmComputeMethod::Regular);

void compute_gemm(
    const CompasContext& ctx,
    Array<float, 2> result,
    Array<bfloat16_t, 2> lhs,
    Array<bfloat16_t, 2> rhs,
    float alpha,
    float beta,
    GemmComputeMethod kind = GemmComputeMethod::Regular);

void compute_complex_gemm(
    const CompasContext& ctx,
    Array<float, 3> result,
    Array<float, 3> lhs,
    Array<float, 3> rhs,
    float alpha,
    float beta,
    GemmComputeMethod kind = GemmComputeMethod::Regular);

void compute_complex_gemm(
    const CompasContext& ctx,
    Array<float, 3> result,
    Array<bfloat16_t, 3> lhs,
    Array<bfloat16_t, 3> rhs,
    float alpha,
    float beta,
    GemmComputeMethod kind = GemmComputeMethod::Regular);

// [n][m] complex <-> planar [2][n][m] (real plane first)
void convert_complex_to_planar(const CompasContext& ctx, Array<float, 3> output, Array<cfloat, 2> input);
void convert_complex_to_planar(const CompasContext& ctx, Array<bfloat16_t, 3> output, Array<cfloat, 2> input);
void convert_planar_to_complex(const CompasContext& ctx, Array<cfloat, 2> output, Array<float, 3> input);

}  // namespace compas
