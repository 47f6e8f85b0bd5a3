// Work-table launch of the FISP recursion for the request server: one
// launch runs voxel blocks of several simulations (of different requests,
// voxel counts and sequences) that share the state count.  See
// compas/device/fisp_epg.h for the recursion and fisp.hip for the per-call
// launch; a voxel's arithmetic is the same in both with one lane per voxel.

#include "compas/device/fisp_epg.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

// Workgroup b runs voxel block items[b].y (64 voxels, one per lane) of
// simulation jobs[items[b].x]; every job of the launch has max_state == SPL.
template<int SPL>
__global__ void __launch_bounds__(64) fisp_epg_table(const FispArgs* __restrict__ jobs, const int2* __restrict__ items) {
    const int2 it = items[blockIdx.x];
    const FispArgs a = jobs[it.x];
    const int voxel = it.y * 64 + int(threadIdx.x);
    const float max_angle = fisp_table_max_angle(a);  // while every lane is active
    if (voxel >= a.tissue.nvoxels) return;
    fisp_epg_voxel<SPL, 1>(a, voxel, 0, max_angle);
}

}  // namespace kernels

bool fisp_table_supported(int max_state) {
    // one lane per voxel holds all states in registers up to 32 (the
    // per-call kernel spreads larger state counts over several lanes)
    return max_state >= 1 && max_state <= 32;
}

template<int SPL>
static void launch_table(hipStream_t s, const kernels::FispArgs* jobs, const int2* items, int n) {
    kernels::fisp_epg_table<SPL><<<unsigned(n), 64, 0, s>>>(jobs, items);
}

void launch_fisp_table(const CompasContext& ctx, int spl, const kernels::FispArgs* jobs, const int2* items, int n) {
    if (n <= 0) return;
    using Fn = void (*)(hipStream_t, const kernels::FispArgs*, const int2*, int);
    static const Fn table[33] = {
        nullptr,          &launch_table<1>,  &launch_table<2>,  &launch_table<3>,  &launch_table<4>,
        &launch_table<5>, &launch_table<6>,  &launch_table<7>,  &launch_table<8>,  &launch_table<9>,
        &launch_table<10>, &launch_table<11>, &launch_table<12>, &launch_table<13>, &launch_table<14>,
        &launch_table<15>, &launch_table<16>, &launch_table<17>, &launch_table<18>, &launch_table<19>,
        &launch_table<20>, &launch_table<21>, &launch_table<22>, &launch_table<23>, &launch_table<24>,
        &launch_table<25>, &launch_table<26>, &launch_table<27>, &launch_table<28>, &launch_table<29>,
        &launch_table<30>, &launch_table<31>, &launch_table<32>};
    if (!fisp_table_supported(spl)) COMPAS_ERROR("state count outside the work-table FISP kernels");
    KernelScope ks(ctx, "fisp_epg_table");
    table[spl](ctx.stream(), jobs, items, n);
    check_launch("fisp_epg_table");
}

}  // namespace compas
