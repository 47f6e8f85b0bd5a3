// J^H.y on the spiral trajectory (six parameters) with the sample sums on MFMA.
//
// Per readout r the adjoint needs, for every voxel v and coil c,
//   U[c][v]  = sum_s y[c][r][s] conj(Phi[r][s][v])
//   U'[c][v] = sum_s (s - ns/2) y[c][r][s] conj(Phi[r][s][v])
// (jacobian_spiral.hip states Phi and how U, U' combine into the six rows).
// That is a complex GEMM per readout, rows (coil, U / U') x K = samples x
// N = voxels, whose B operand conj(Phi) is generated, not stored: Phi is a
// geometric sequence in s per voxel.  Here it runs on
// v_mfma_f32_16x16x32_bf16 with BF16 hi + lo operands (FP32 accumulation):
//   * A (16 real rows per group of four coils: U_re, U_im, U'_re, U'_im rows
//     of each coil) is built from y once per workgroup and 256-sample chunk,
//     as (a_re, -a_im) / (a_im, a_re) pairs along the interleaved K split
//     into hi + lo, and kept in LDS in fragment order (one 16-byte read per
//     lane, part and K step);
//   * B (16 voxels per tile) is generated in registers: lane l holds voxel
//     l % 16 and four consecutive samples 4 (l / 16) + j of each 16-sample
//     K step, advanced by Phi^16 per step (four complex multiplies) and
//     restarted from an FP64 phase every 256 samples;
//   * the 16 x 16 accumulator gives each lane one coil's U and U' of its
//     voxel, so the per-readout epilogue (the products with conj(m) and
//     conj(dm_p)) is per lane; coils are summed across the four lane groups
//     once at the end.
// Every lane generates only its own B elements: ~10 VALU instructions per
// (sample, voxel) for the recurrence and the hi/lo split, against ~36 for the
// FP32 VALU form at four coils.
//
// Layout of v_mfma_f32_16x16x32_bf16 (wave64): lane l supplies row (A) or
// column (B) l % 16 and K elements 8 (l / 16) .. 8 (l / 16) + 7; the result
// element j of lane l is row 4 (l / 16) + j, column l % 16.

#include <algorithm>
#include <cstdint>

#include "compas/device/common.h"
#include "compas/device/mfma_tile.h"
#include "compas/internal.h"

namespace compas {
namespace kernels {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int SHM_WAVES = 4;  // waves per workgroup
constexpr int SHM_CHUNK = 256;  // samples of y staged in LDS at a time; Phi restarts at each chunk

__device__ inline mfma::bf16x8 as_bf16x8(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3) {
    const uint4 w = {w0, w1, w2, w3};
    return __builtin_bit_cast(mfma::bf16x8, w);
}

// Phi at sample s: exp((ns/2 - s) dt R2) exp(i (phi0 + s theta)), phase in FP64
__device__ inline cfloat spiral_phi(double phi0, double theta, float R2, float dt, int ns, int s) {
    const float mag = expf((0.5f * float(ns) - float(s)) * dt * R2);
    return dev::polar_reduced(mag, fma(double(s), theta, phi0));
}

// Partial J^H.y over readouts [r0, r1): partial[chunk * groups + group][p][v].
// Grid (ceil(nv / (16 T SHM_WAVES)), readout chunks), block 64 SHM_WAVES.  G
// groups of four coils share the generated operand.
template<int G, int T, bool LO>
__global__ void __launch_bounds__(64 * SHM_WAVES) spiral_hermitian_mfma(
    dev::TissueView t, const cfloat* __restrict__ m, const cfloat* __restrict__ dT1,
    const cfloat* __restrict__ dT2, const cfloat* __restrict__ dB1, const cfloat* __restrict__ dB0,
    const cfloat* __restrict__ coils, int nc, const cfloat* __restrict__ y, const cfloat* __restrict__ k_start,
    const cfloat* __restrict__ delta_k, int nr, int ns, float dt, int r_chunk, cfloat* __restrict__ partial) {
    constexpr int P = LO ? mfma::P_BF16X3 : mfma::P_BF16X1;
    constexpr int KSTEPS = SHM_CHUNK / 16;
    // A fragments of the chunk in fragment order: [group][K step][lane] hi, lo
    __shared__ uint4 afrag_hi[G][KSTEPS][64];
    __shared__ uint4 afrag_lo[LO ? G : 1][LO ? KSTEPS : 1][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = lane & 15, g = lane >> 4;
    const int nv = t.nvoxels;
    const int r0 = blockIdx.y * r_chunk, r1 = min(nr, r0 + r_chunk);
    const int vbase = (blockIdx.x * SHM_WAVES + wave) * (16 * T);

    // voxel constants of the lane's voxel in each tile (out-of-range voxels read voxel 0)
    int vt[T];
    bool okt[T];
    float R2[T], T1v[T], T2v[T];
    double xv[T], yv[T], B0v[T];
    cfloat rho[T], Cg[T][G];
#pragma unroll
    for (int ti = 0; ti < T; ti++) {
        const int v = vbase + 16 * ti + n;
        okt[ti] = v < nv;
        const int vv = okt[ti] ? v : 0;
        vt[ti] = vv;
        T1v[ti] = t.T1(vv);
        T2v[ti] = t.T2(vv);
        R2[ti] = 1.0f / T2v[ti];
        xv[ti] = t.x(vv);
        yv[ti] = t.y(vv);
        B0v[ti] = t.B0(vv);
        rho[ti] = t.rho(vv);
#pragma unroll
        for (int gi = 0; gi < G; gi++) {
            const int c = 4 * gi + g;
            Cg[ti][gi] = c < nc ? coils[size_t(c) * nv + vv] : cfloat {0, 0};
        }
    }
    // per-coil sums over readouts (the lane's coil 4 gi + g of each group)
    cfloat g1[T][G], g2[T][G], gb1[T][G], gb0[T][G], gu[T][G];
#pragma unroll
    for (int ti = 0; ti < T; ti++)
#pragma unroll
        for (int gi = 0; gi < G; gi++) g1[ti][gi] = g2[ti][gi] = gb1[ti][gi] = gb0[ti][gi] = gu[ti][gi] = {0, 0};

    const double half = 0.5 * double(ns) * double(dt);
    const float sp0 = -0.5f * float(ns);

    for (int r = r0; r < r1; r++) {
        const cfloat dk = delta_k[r], k0 = k_start[r];
        double theta[T], phi0[T];
        cfloat step1[T], step16[T];
#pragma unroll
        for (int ti = 0; ti < T; ti++) {
            theta[ti] = double(dk.re) * xv[ti] + double(dk.im) * yv[ti] + double(dt) * dev::TWO_PI * B0v[ti];
            phi0[ti] = -half * dev::TWO_PI * B0v[ti] + double(k0.re) * xv[ti];
            step1[ti] = dev::polar_reduced(expf(-dt * R2[ti]), theta[ti]);
            step16[ti] = dev::polar_reduced(expf(-16.0f * dt * R2[ti]), 16.0 * theta[ti]);
        }
        f32x4 acc[T][G];
#pragma unroll
        for (int ti = 0; ti < T; ti++)
#pragma unroll
            for (int gi = 0; gi < G; gi++) acc[ti][gi] = f32x4 {0.0f, 0.0f, 0.0f, 0.0f};

        for (int sc = 0; sc < ns; sc += SHM_CHUNK) {
            __syncthreads();  // the previous chunk's fragments have been read
            // A fragment of (group gi, K step ks, lane fl): row fl % 16 = coil 4 gi + row / 4 and
            // (U_re, U_im, U'_re, U'_im)[row % 4], samples sc + 16 ks + 4 (fl / 16) + j, as
            // (a_re, -a_im) / (a_im, a_re) pairs split into BF16 hi + lo
            for (int e = threadIdx.x; e < G * KSTEPS * 64; e += 64 * SHM_WAVES) {
                const int fl = e & 63, ks = (e >> 6) % KSTEPS, gi = e / (64 * KSTEPS);
                const int row = fl & 15, c = 4 * gi + (row >> 2), q = row & 3;
                const int s0 = sc + 16 * ks + 4 * (fl >> 4);
                uint32_t hw[4], lw[4];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int s = s0 + j;
                    cfloat a = {0.0f, 0.0f};
                    if (c < nc && s < ns) a = y[(size_t(c) * nr + r) * ns + s];
                    if (q >= 2) a = a * (float(s) + sp0);
                    const float e0 = (q & 1) ? a.im : a.re;
                    const float e1 = (q & 1) ? a.re : -a.im;
                    mfma::split_pair<P>(e0, e1, hw[j], lw[j]);
                }
                afrag_hi[gi][ks][fl] = uint4 {hw[0], hw[1], hw[2], hw[3]};
                if constexpr (LO) afrag_lo[gi][ks][fl] = uint4 {lw[0], lw[1], lw[2], lw[3]};
            }
            __syncthreads();

            // Phi at the lane's four samples of the first K step of the chunk
            cfloat ph[T][4];
#pragma unroll
            for (int ti = 0; ti < T; ti++) {
                ph[ti][0] = spiral_phi(phi0[ti], theta[ti], R2[ti], dt, ns, sc + 4 * g);
#pragma unroll
                for (int j = 1; j < 4; j++) ph[ti][j] = dev::cmul(ph[ti][j - 1], step1[ti]);
            }
            const int kend = min(SHM_CHUNK, ns - sc);
            for (int ks = 0; ks < kend; ks += 16) {
                mfma::bf16x8 ahi[G], alo[G];
#pragma unroll
                for (int gi = 0; gi < G; gi++) {
                    ahi[gi] = __builtin_bit_cast(mfma::bf16x8, afrag_hi[gi][ks >> 4][lane]);
                    if constexpr (LO) alo[gi] = __builtin_bit_cast(mfma::bf16x8, afrag_lo[gi][ks >> 4][lane]);
                }
#pragma unroll
                for (int ti = 0; ti < T; ti++) {
                    // B: conj(Phi) of voxel n at the lane's four samples, as (re, im) pairs
                    uint32_t hw[4], lw[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) mfma::split_pair<P>(ph[ti][j].re, -ph[ti][j].im, hw[j], lw[j]);
                    const mfma::bf16x8 bhi = as_bf16x8(hw[0], hw[1], hw[2], hw[3]);
                    const mfma::bf16x8 blo = as_bf16x8(lw[0], lw[1], lw[2], lw[3]);
#pragma unroll
                    for (int gi = 0; gi < G; gi++) {
                        acc[ti][gi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[gi], bhi, acc[ti][gi], 0, 0, 0);
                        if constexpr (LO) {
                            acc[ti][gi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[gi], blo, acc[ti][gi], 0, 0, 0);
                            acc[ti][gi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[gi], bhi, acc[ti][gi], 0, 0, 0);
                        }
                    }
#pragma unroll
                    for (int j = 0; j < 4; j++) ph[ti][j] = dev::cmul(ph[ti][j], step16[ti]);
                }
            }
        }

        // readout epilogue: the lane holds U, U' of coil 4 gi + g for its voxel
#pragma unroll
        for (int ti = 0; ti < T; ti++) {
            const size_t i = size_t(r) * nv + vt[ti];
            const cfloat me = m[i], a1 = dT1[i], a2 = dT2[i], ab1 = dB1[i], ab0 = dB0[i];
            const float kT2 = dt * R2[ti] * R2[ti];
            const float kB0 = -dev::TWO_PI_F * dt;
#pragma unroll
            for (int gi = 0; gi < G; gi++) {
                const cfloat U = {acc[ti][gi][0], acc[ti][gi][1]};
                const cfloat Up = {acc[ti][gi][2], acc[ti][gi][3]};
                g1[ti][gi] = dev::cfma_conj(U, a1, g1[ti][gi]);
                g2[ti][gi] = dev::cfma_conj(U, a2, g2[ti][gi]);
                g2[ti][gi] = dev::cfma_conj(Up * kT2, me, g2[ti][gi]);
                gb1[ti][gi] = dev::cfma_conj(U, ab1, gb1[ti][gi]);
                gb0[ti][gi] = dev::cfma_conj(U, ab0, gb0[ti][gi]);
                gb0[ti][gi] = dev::cfma_conj(cfloat {-Up.im * kB0, Up.re * kB0}, me, gb0[ti][gi]);
                gu[ti][gi] = dev::cfma_conj(U, me, gu[ti][gi]);
            }
        }
    }

    // fold the lane's coils, sum the four lane groups (coils 4 gi + g), store
#pragma unroll
    for (int ti = 0; ti < T; ti++) {
        cfloat out[5];
#pragma unroll
        for (int p = 0; p < 5; p++) out[p] = {0, 0};
#pragma unroll
        for (int gi = 0; gi < G; gi++) {
            const cfloat C = Cg[ti][gi];
            const cfloat rc = dev::cmul(rho[ti], C);
            out[0] = out[0] + dev::cmul(conj(rc * T1v[ti]), g1[ti][gi]);
            out[1] = out[1] + dev::cmul(conj(rc * T2v[ti]), g2[ti][gi]);
            out[2] = out[2] + dev::cmul(conj(rc), gb1[ti][gi]);
            out[3] = out[3] + dev::cmul(conj(rc), gb0[ti][gi]);
            out[4] = out[4] + dev::cmul(conj(C), gu[ti][gi]);
        }
#pragma unroll
        for (int p = 0; p < 5; p++) {
            out[p].re += __shfl_xor(out[p].re, 16);
            out[p].im += __shfl_xor(out[p].im, 16);
            out[p].re += __shfl_xor(out[p].re, 32);
            out[p].im += __shfl_xor(out[p].im, 32);
        }
        if (g == 0 && okt[ti]) {
            const int v = vbase + 16 * ti + n;
            cfloat* dst = partial + size_t(blockIdx.y) * 6 * size_t(nv) + v;
#pragma unroll
            for (int p = 0; p < 5; p++) dst[size_t(p) * nv] = out[p];
            dst[size_t(5) * nv] = dev::cmul(out[4], cfloat {0.0f, -1.0f});
        }
    }
}

}  // namespace kernels

namespace {

constexpr int SHM_T = 2;  // 16-voxel tiles per wave

template<int G, bool LO>
void launch_shm(dim3 grid, hipStream_t s, const dev::TissueView& tv, const cfloat* m, const cfloat* d1, const cfloat* d2,
                const cfloat* db1, const cfloat* db0, const cfloat* coils, int nc, const cfloat* y, const cfloat* k0,
                const cfloat* dk, int nr, int ns, float dt, int r_chunk, cfloat* partial) {
    kernels::spiral_hermitian_mfma<G, SHM_T, LO><<<grid, 64 * kernels::SHM_WAVES, 0, s>>>(
        tv, m, d1, d2, db1, db0, coils, nc, y, k0, dk, nr, ns, dt, r_chunk, partial);
}

}  // namespace

bool spiral_hermitian_mfma(
    const CompasContext& ctx, const TissueParameters& p, const cfloat* m, const cfloat* d1, const cfloat* d2,
    const cfloat* db1, const cfloat* db0, const cfloat* coils, int nc, const cfloat* y, const SpiralTrajectory& traj,
    ContractionPrecision precision, cfloat* JHy) {
    if (precision == ContractionPrecision::FP32) return false;
    if (nc > 8) return false;  // more coils: the VALU form
    const bool lo = precision == ContractionPrecision::BF16x3 || precision == ContractionPrecision::FP16x3;
    const int nr = traj.nreadouts, ns = traj.samples_per_readout, nv = p.nvoxels;
    const int vblocks = int(dev::div_ceil(nv, 16 * SHM_T * kernels::SHM_WAVES));
    // readout chunks so that about eight waves per SIMD are in the grid
    const int want = std::max(1, std::min(nr, int(dev::div_ceil(8 * 1024, int64_t(vblocks) * kernels::SHM_WAVES))));
    const int r_chunk = int(dev::div_ceil(nr, want));
    const int chunks = int(dev::div_ceil(nr, r_chunk));
    auto partial = ctx.empty<cfloat, 1>({int64_t(chunks) * 6 * nv});
    const auto tv = dev::tissue_view(p);
    {
        KernelScope ks(ctx, "spiral_hermitian_mfma");
        const dim3 grid(vblocks, chunks);
        hipStream_t s = ctx.stream();
        const cfloat *k0 = traj.k_start.data(), *dk = traj.delta_k.data();
        const float dt = traj.delta_t;
        if (nc <= 4) {
            if (lo) launch_shm<1, true>(grid, s, tv, m, d1, d2, db1, db0, coils, nc, y, k0, dk, nr, ns, dt, r_chunk, partial.data());
            else launch_shm<1, false>(grid, s, tv, m, d1, d2, db1, db0, coils, nc, y, k0, dk, nr, ns, dt, r_chunk, partial.data());
        } else {
            if (lo) launch_shm<2, true>(grid, s, tv, m, d1, d2, db1, db0, coils, nc, y, k0, dk, nr, ns, dt, r_chunk, partial.data());
            else launch_shm<2, false>(grid, s, tv, m, d1, d2, db1, db0, coils, nc, y, k0, dk, nr, ns, dt, r_chunk, partial.data());
        }
        check_launch("spiral_hermitian_mfma");
    }
    spiral_hermitian_reduce_partials(ctx, partial.data(), chunks, nv, JHy);
    return true;
}

}  // namespace compas
