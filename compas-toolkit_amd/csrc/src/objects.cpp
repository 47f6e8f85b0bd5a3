// make_* constructors of the compas:: data model.
// Reference: src/parameters/tissue.cpp:6-93, include/compas/sequences/fisp.h:44-60,
// pssfp.h:54-72, trajectories/cartesian.h:43-58, spiral.h:41-57.

#include <cstring>

#include "compas/compas.h"

namespace compas {

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
    View<float> z) {
    COMPAS_CHECK(num_voxels >= 0);
    const bool has_z = !z.is_empty();
    const bool has_b0 = !B0.is_empty();
    const bool has_b1 = !B1.is_empty();
    const size_t nv = size_t(num_voxels);
    auto need = [&](const View<float>& v, const char* name) {
        if (v.is_empty() && nv > 0) COMPAS_ERROR((std::string("missing tissue field ") + name).c_str());
        if (!v.is_empty()) COMPAS_CHECK(v.size() >= num_voxels);
    };
    need(T1, "T1");
    need(T2, "T2");
    need(rho_x, "rho_x");
    need(rho_y, "rho_y");
    need(x, "x");
    need(y, "y");

    // host staging of the [NUM_FIELDS][nvoxels] structure-of-arrays map, one upload
    std::vector<float> host(size_t(TissueParameterField::NUM_FIELDS) * nv, 0.0f);
    auto put = [&](int field, const View<float>& v, float fallback) {
        float* dst = host.data() + size_t(field) * nv;
        if (!v.is_empty()) {
            std::memcpy(dst, v.data(), nv * sizeof(float));
        } else {
            for (size_t i = 0; i < nv; i++) dst[i] = fallback;
        }
    };
    put(TissueParameterField::T1, T1, 0.0f);
    put(TissueParameterField::T2, T2, 0.0f);
    put(TissueParameterField::B1, B1, 1.0f);  // tissue.cpp:77-80
    put(TissueParameterField::B0, B0, 0.0f);
    put(TissueParameterField::RHO_X, rho_x, 0.0f);
    put(TissueParameterField::RHO_Y, rho_y, 0.0f);
    put(TissueParameterField::X, x, 0.0f);
    put(TissueParameterField::Y, y, 0.0f);
    put(TissueParameterField::Z, z, 0.0f);
    auto data = ctx.allocate_from<float, 2>(host.data(), {TissueParameterField::NUM_FIELDS, int64_t(nv)});
    TissueParameters p(data, has_z, has_b0, has_b1);
    p.chunk_size = chunk_size > 0 ? chunk_size : num_voxels;
    return p;
}

FISPSequence make_fisp_sequence(
    const CompasContext& ctx,
    View<cfloat> RF_train,
    View<cfloat, 2> sliceprofiles,
    float TR,
    float TE,
    int max_state,
    float TI) {
    return FISPSequence(ctx.allocate(RF_train), ctx.allocate(sliceprofiles), TR, TE, 0.0f, max_state, TI);
}

pSSFPSequence make_pssfp_sequence(
    const CompasContext& ctx,
    View<cfloat> RF_train,
    float TR,
    View<cfloat> gamma_dt_RF,
    RepetitionData dt,
    RepetitionData gamma_dt_GRz,
    View<float> z) {
    COMPAS_CHECK(RF_train.size() > 0);  // pssfp.h:176
    return pSSFPSequence(ctx.allocate(RF_train), TR, ctx.allocate(gamma_dt_RF), dt, gamma_dt_GRz, ctx.allocate(z));
}

CartesianTrajectory make_cartesian_trajectory(
    const CompasContext& ctx,
    int nreadouts,
    int samples_per_readout,
    float delta_t,
    HostView<cfloat> k_start,
    cfloat delta_k) {
    COMPAS_CHECK(k_start.size() == nreadouts);  // cartesian.h:49
    return CartesianTrajectory(nreadouts, samples_per_readout, delta_t, ctx.allocate(k_start), delta_k);
}

SpiralTrajectory make_spiral_trajectory(
    const CompasContext& ctx,
    int nreadouts,
    int samples_per_readout,
    float delta_t,
    View<cfloat> k_start,
    View<cfloat> delta_k) {
    COMPAS_CHECK(k_start.size() == nreadouts);  // spiral.h:50-51
    COMPAS_CHECK(delta_k.size() == nreadouts);
    return SpiralTrajectory(nreadouts, samples_per_readout, delta_t, ctx.allocate(k_start), ctx.allocate(delta_k));
}

}  // namespace compas
