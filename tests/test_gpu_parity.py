"""Parity of the HIP path (through the C-ABI) with the oracle on seeded inputs.

Tolerances (relative L2 over the whole output, complex as two reals):
  magnetisation  <= 1e-5   against the FP64 oracle and the FP32 port
  signal, J.v, J^H.y <= 1e-4 against the FP64 oracle
"""

import numpy as np
import pytest

import compas_amd as ca
from compas_amd import workloads as W
from gpu_helpers import banner, context, objects, rel

pytestmark = pytest.mark.gpu

MAG_TOL = 1e-5
SIG_TOL = 1e-4


@pytest.fixture(scope="module")
def ctx():
    return context()


def small_fisp(nv=700, nTR=96, N=16, b0=30.0, seed=3, **kw):
    n = int(np.ceil(np.sqrt(nv)))
    t = W.make_tissue(W.SplitMix64(seed), n, 20.0, b0).slice(0, nv)
    rng = W.SplitMix64(seed + 100)
    RF = (rng.uniform(nTR, 5, 70) * np.exp(1j * rng.uniform(nTR, -0.5, 0.5))).astype(np.complex64)
    seq = W.FispSeq(RF=RF, profiles=np.ones((1, nTR), np.complex64), TR=0.009, TE=0.004, max_state=N, TI=0.02)
    for k, v in kw.items():
        setattr(seq, k, v)
    return t, seq


def run_fisp(t, seq):
    context()
    p = ca.TissueParameters.from_tissue(t)
    s = ca.FispSequence.from_sequence(seq)
    return ca.simulate_magnetization(p, s).collect()


# ---------------------------------------------------------------- FISP ---

@pytest.mark.parametrize("N", [1, 2, 5, 8, 16, 25, 32, 40, 64, 100, 128])
def test_fisp_state_counts(ctx, oracle, N):
    banner(f"fisp N={N}")
    t, seq = small_fisp(nv=300, nTR=70, N=N)
    got = run_fisp(t, seq)
    want = oracle.fisp(t, seq, "f64")
    assert rel(got, want) < MAG_TOL
    assert rel(got, oracle.fisp(t, seq, "f32")) < MAG_TOL


@pytest.mark.parametrize("nv", [1, 63, 64, 65, 4096, 20000])
def test_fisp_voxel_counts_and_layouts(ctx, oracle, nv):
    banner(f"fisp nv={nv}")
    t, seq = small_fisp(nv=nv, nTR=40, N=12)
    got = run_fisp(t, seq)
    assert got.shape == (40, nv)
    assert rel(got, oracle.fisp(t, seq, "f64")) < MAG_TOL


def test_fisp_sequence_options(ctx, oracle):
    banner("fisp usf/reps/TW/no-inversion/slices")
    t, seq = small_fisp(nv=257, nTR=50, N=20)
    seq.undersampling = 3
    seq.repetitions = 3
    seq.TW = 0.15
    prof = np.stack([np.linspace(0.6, 1.1, 50), np.full(50, 0.9 - 0.2j)]).astype(np.complex64)
    seq.profiles = prof
    got = run_fisp(t, seq)
    assert got.shape == (150, 257)
    assert rel(got, oracle.fisp(t, seq, "f64")) < MAG_TOL
    seq.inversion = False
    seq.wait_spoiling = False
    assert rel(run_fisp(t, seq), oracle.fisp(t, seq, "f64")) < MAG_TOL


def test_fisp_b0_zero_branch(ctx, oracle):
    t, seq = small_fisp(nv=200, nTR=60, N=16, b0=None)
    assert np.all(t.B0 == 0)
    assert rel(run_fisp(t, seq), oracle.fisp(t, seq, "f64")) < MAG_TOL


@pytest.mark.parametrize("nTR", [1, 2, 3])
@pytest.mark.parametrize("nslices", [1, 2])
def test_fisp_short_trains(ctx, oracle, nTR, nslices):
    # the recursion reads the RF table entry of the TR after the current one
    # (the next slice's first, or the table's terminating entry after the
    # last TR of the last slice); one slice runs the loop without tests
    banner(f"fisp nTR={nTR} nslices={nslices}")
    t, seq = small_fisp(nv=130, nTR=nTR, N=8)
    if nslices == 2:
        seq.profiles = np.stack([np.linspace(0.6, 1.1, nTR), np.full(nTR, 0.9 - 0.2j)]).astype(np.complex64)
    got = run_fisp(t, seq)
    assert got.shape == (nTR, 130)
    assert rel(got, oracle.fisp(t, seq, "f64")) < MAG_TOL


@pytest.mark.parametrize("peak_deg,nv", [(5000.0, 300), (8000.0, 300), (8000.0, 66000)])
def test_fisp_wide_flip_angles(ctx, oracle, peak_deg, nv):
    # A wave whose half flip angles alpha/2 = deg2rad |RF| B1 / 2 can pass 64
    # takes the TR loop with the library sincosf branch, a wave whose largest
    # stays below it the branch-free polynomial.  B1 in [0.8, 1.2]: 5000 deg
    # keeps alpha/2 <= 52.4 (every wave polynomial, at large arguments);
    # 8000 deg puts alpha/2 in [55.9, 83.8], both sides of 64 within waves.
    # 300 voxels run four lanes per voxel, 66000 one.
    banner(f"fisp flip angles up to {peak_deg} deg, nv={nv}")
    t, seq = small_fisp(nv=nv, nTR=60, N=16)
    RF = seq.RF.copy()
    RF[7::11] = (peak_deg * np.exp(1j * np.linspace(-0.4, 0.4, RF[7::11].size))).astype(np.complex64)
    seq.RF = RF
    half = np.deg2rad(peak_deg) * t.B1 / 2
    assert (half.max() > 64) == (peak_deg > 6000)
    got = run_fisp(t, seq)
    assert rel(got, oracle.fisp(t, seq, "f64")) < MAG_TOL


@pytest.mark.parametrize("nv,options", [(300, False), (300, True), (5000, False)])
def test_fisp_echo_paths(ctx, oracle, nv, options):
    # the kernel samples F+_0 at TE by undoing the decay over TR - TE on the
    # updated state; a wave holding a voxel whose decay over TR - TE is too
    # small to invert applies D(TE) R to state 0 instead.  TE = 0.2 ms and
    # T2 = 0.15 ms: the echo keeps e^(-TE/T2) = 0.26 of the transverse state
    # while e^(-(TR-TE)/T2) underflows FP32.  Voxel counts with different
    # lanes per voxel; options: undersampling 2, 2 repetitions, 2 slices
    banner(f"fisp echo paths (short T2) nv={nv} options={options}")
    t, seq = small_fisp(nv=nv, nTR=60, N=16)
    seq.TE = 0.0002
    if options:
        seq.undersampling = 2
        seq.repetitions = 2
        seq.profiles = np.stack([np.linspace(0.6, 1.1, 60), np.full(60, 0.9 - 0.2j)]).astype(np.complex64)
    t.T2[:64] = 1.5e-4  # the first wave: every lane short
    t.T2[100] = 1.5e-4  # the second wave: one lane short
    got = run_fisp(t, seq)
    want = oracle.fisp(t, seq, "f64")
    assert rel(got, want) < MAG_TOL
    for cols in (slice(0, 64), slice(64, 128), slice(128, nv)):
        assert rel(got[:, cols], want[:, cols]) < MAG_TOL, cols
    # the short-T2 echoes are not ~0: per-voxel RMS against the other voxels'
    assert np.linalg.norm(want[:, :64]) / 8.0 > 0.05 * np.linalg.norm(want[:, 128:]) / np.sqrt(nv - 128)


@pytest.mark.parametrize("index,scale", [(1, 1), (2, 4)])
def test_fisp_workloads(ctx, oracle, index, scale):
    banner(f"fisp workload {index} scale {scale}")
    wl = W.config(index, scale=scale)
    params, seq, _ = objects(wl)
    got = ca.simulate_magnetization(params, seq).collect()
    want = oracle.fisp(wl.tissue, wl.sequence, "f64")
    assert rel(got, want) < MAG_TOL
    assert rel(got, oracle.fisp(wl.tissue, wl.sequence, "f32")) < MAG_TOL


def test_fisp_full_size_voxel_sample(ctx, oracle):
    # 50176 voxels x 1120 TRs x 25 states: every voxel is independent, so a
    # contiguous slice of voxels checked against the oracle checks the full run
    banner("fisp full size 50176 x 1120")
    wl = W.config(2)
    params, seq, _ = objects(wl)
    got = ca.simulate_magnetization(params, seq).collect()
    assert got.shape == (1120, 50176)
    for begin in (0, 25000, 50176 - 300):
        sub = wl.tissue.slice(begin, begin + 300)
        assert rel(got[:, begin:begin + 300], oracle.fisp(sub, wl.sequence, "f64")) < MAG_TOL


def test_fisp_max_state_limit(ctx):
    t, seq = small_fisp(nv=10, nTR=5, N=129)
    with pytest.raises(ca.CompasError, match="max_state"):
        run_fisp(t, seq)


# --------------------------------------------------------------- pSSFP ---

PSSFP_PORT_TOL = 1e-4  # FP64-state GPU recursion against the FP32 port of the reference


def test_pssfp_workload(ctx, oracle):
    """Against FP64 the GPU (FP64 isochromat state) is within 1e-5.  Against the
    FP32 port of the reference the distance is the port's own FP32 drift
    (3.2-3.4e-5 measured at this workload): no FP32 form of this recursion is
    within 1e-5 of FP64, so that distance is bounded at 1e-4 and reported."""
    banner("pssfp workload 3 scale 8")
    wl = W.config(3, scale=8)
    params, seq, _ = objects(wl)
    got = ca.simulate_magnetization(params, seq).collect()
    want = oracle.pssfp(wl.tissue, wl.sequence, "f64")
    port = oracle.pssfp(wl.tissue, wl.sequence, "f32")
    e64, e32, port64 = rel(got, want), rel(got, port), rel(port, want)
    print(f"pssfp rel-L2: gpu vs fp64 {e64:.2e}, gpu vs fp32 port {e32:.2e}, fp32 port vs fp64 {port64:.2e}")
    assert e64 < MAG_TOL
    assert e32 < PSSFP_PORT_TOL


def test_pssfp_complex_z_as_julia_passes_it(ctx, oracle):
    """CompasToolkit.jl:614 passes z as ComplexF32; the library takes its real
    parts.  Same result, bitwise, as a float z, and the oracle's with real z;
    a nonzero imaginary part is rejected."""
    banner("pssfp complex z")
    t = W.make_tissue(W.SplitMix64(15), 10, 20.0, 30.0)
    seq = W.make_pssfp(24, 50.0, nz=7, nRF=5)
    context()
    p = ca.TissueParameters.from_tissue(t)
    as_julia = ca.pSSFPSequence.from_sequence(seq)  # complex64 z
    as_float = ca.pSSFPSequence(seq.RF, seq.TR, seq.gamma_dt_RF, seq.dt, seq.gamma_dt_GRz, seq.z,
                                z_dtype=np.float32)
    a = ca.simulate_magnetization(p, as_julia).collect()
    b = ca.simulate_magnetization(p, as_float).collect()
    assert np.array_equal(a, b)
    assert rel(a, oracle.pssfp(t, seq, "f64")) < MAG_TOL
    with pytest.raises(ca.CompasError, match="must be real"):
        ca.pSSFPSequence(seq.RF, seq.TR, seq.gamma_dt_RF, seq.dt, seq.gamma_dt_GRz,
                         seq.z.astype(np.complex64) + 0.5j)


@pytest.mark.parametrize("nz,nRF", [(1, 1), (3, 5), (35, 25), (60, 7), (130, 3), (3, 300), (2, 600)])
def test_pssfp_shapes(ctx, oracle, nz, nRF):
    banner(f"pssfp nz={nz} nRF={nRF}")
    n = 12
    t = W.make_tissue(W.SplitMix64(5), n, 20.0, 40.0)
    seq = W.make_pssfp(30, 60.0, nz=nz, nRF=nRF)
    context()
    got = ca.simulate_magnetization(ca.TissueParameters.from_tissue(t), ca.pSSFPSequence.from_sequence(seq)).collect()
    assert rel(got, oracle.pssfp(t, seq, "f64")) < MAG_TOL


def test_pssfp_large_angles(ctx, oracle):
    # few RF sub-steps with large flip angles exercise the sincos branch of the rotation
    t = W.make_tissue(W.SplitMix64(8), 10, 20.0, 20.0)
    seq = W.make_pssfp(20, 150.0, nz=4, nRF=1)
    context()
    got = ca.simulate_magnetization(ca.TissueParameters.from_tissue(t), ca.pSSFPSequence.from_sequence(seq)).collect()
    assert rel(got, oracle.pssfp(t, seq, "f64")) < MAG_TOL


# -------------------------------------------------------------- signal ---

@pytest.mark.parametrize("method", ["direct", "matmul_fast", "matmul", "matmul_pedantic"])
def test_cartesian_signal_methods(ctx, oracle, method):
    banner(f"cartesian signal {method}")
    wl = W.config(2, scale=8)
    params, seq, traj = objects(wl)
    echos = oracle.fisp(wl.tissue, wl.sequence, "f64").astype(np.complex64)
    got = ca.magnetization_to_signal(echos, params, traj, wl.coils, method=method).collect()
    want = oracle.signal_cartesian(wl.tissue, wl.trajectory, wl.coils, echos, "f64")
    assert got.shape == want.shape
    assert rel(got, want) < SIG_TOL


@pytest.mark.parametrize("method", ["matmul_fast", "matmul_pedantic"])
def test_cartesian_signal_full_size(ctx, oracle, method):
    # 4 coils x 1120 readouts x 224 samples over 50176 voxels: the FP64 oracle on a
    # subset of readouts (rows are independent), and the FP32 VALU path on all of them
    banner(f"cartesian signal full size {method}")
    wl = W.config(2)
    params, seq, traj = objects(wl)
    echos = ca.simulate_magnetization(params, seq)
    got = ca.magnetization_to_signal(echos, params, traj, wl.coils, method=method).collect()
    plain = ca.magnetization_to_signal(echos, params, traj, wl.coils, method="direct").collect()
    assert rel(got, plain) < SIG_TOL
    e = echos.collect()
    ridx = np.array([0, 1, 333, 777, 1119])
    sub = W.CartesianTraj(len(ridx), traj.samples_per_readout, wl.trajectory.delta_t,
                          wl.trajectory.k_start[ridx], wl.trajectory.delta_k)
    want = oracle.signal_cartesian(wl.tissue, sub, wl.coils, e[ridx].astype(np.complex128), "f64")
    assert rel(got[:, ridx, :], want) < SIG_TOL
    assert rel(plain[:, ridx, :], want) < SIG_TOL


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (5, 3, 7), (129, 65, 33), (200, 300, 1000)])
def test_contraction_ragged_shapes(ctx, oracle, M, N, K):
    # the signal entry point is a contraction echos[M][K] x W[N][K]; tiny grids of voxels
    # with one readout sample make M, N, K arbitrary: compare every method against FP64
    rng = W.SplitMix64(M * 1000 + N * 10 + K)
    n = int(np.ceil(np.sqrt(K)))
    t = W.make_tissue(rng, n, 20.0, 30.0).slice(0, K)
    traj = W.CartesianTraj(M, 1, 5e-6, np.zeros(M, np.complex64), float(np.float32(2 * np.pi / 20.0)))
    coils = ((rng.normal(N * K) + 1j * rng.normal(N * K)) / 3).reshape(N, K).astype(np.complex64)
    echos = ((rng.normal(M * K) + 1j * rng.normal(M * K)) / 3).reshape(M, K).astype(np.complex64)
    context()
    params = ca.TissueParameters.from_tissue(t)
    ctraj = ca.CartesianTrajectory.from_trajectory(traj)
    want = oracle.signal_cartesian(t, traj, coils, echos, "f64")
    for method, tol in (("direct", 1e-5), ("matmul_fast", 1e-4), ("matmul_pedantic", 1e-5), ("matmul_low", 2e-2)):
        got = ca.magnetization_to_signal(echos, params, ctraj, coils, method=method).collect()
        assert rel(got, want) < tol, method


@pytest.mark.parametrize("M,nc,ns,K", [(1, 1, 4, 1), (5, 3, 8, 7), (129, 2, 12, 33), (261, 5, 28, 700),
                                        (200, 1, 64, 1000), (70, 7, 132, 4099)])
def test_cartesian_generated_weights_ragged(ctx, oracle, M, nc, ns, K):
    # samples per readout a multiple of 4: the BF16 methods generate the weights
    # C[c][v] T[s / 4][v] q[v]^(s % 4) in the contraction's staging; ragged
    # readouts, coils and voxels against FP64 and against the FP32 VALU path
    rng = W.SplitMix64(M * 7919 + nc * 31 + ns * 3 + K)
    n = int(np.ceil(np.sqrt(K)))
    t = W.make_tissue(rng, n, 20.0, 30.0).slice(0, K)
    k0 = ((rng.normal(M) - 0.5) * 0.2 + 1j * (rng.normal(M) - 0.5) * 0.2).astype(np.complex64)
    traj = W.CartesianTraj(M, ns, 5e-6, k0, float(np.float32(2 * np.pi / 20.0)))
    coils = ((rng.normal(nc * K) + 1j * rng.normal(nc * K)) / 3).reshape(nc, K).astype(np.complex64)
    echos = ((rng.normal(M * K) + 1j * rng.normal(M * K)) / 3).reshape(M, K).astype(np.complex64)
    context()
    params = ca.TissueParameters.from_tissue(t)
    ctraj = ca.CartesianTrajectory.from_trajectory(traj)
    want = oracle.signal_cartesian(t, traj, coils, echos, "f64")
    plain = ca.magnetization_to_signal(echos, params, ctraj, coils, method="direct").collect()
    assert rel(plain, want) < 1e-5
    for method, tol in (("matmul_fast", 1e-4), ("matmul", 1e-4), ("matmul_low", 2e-2)):
        got = ca.magnetization_to_signal(echos, params, ctraj, coils, method=method).collect()
        assert got.shape == (nc, M, ns)
        assert rel(got, want) < tol, method
        assert rel(got, plain) < tol, method


def test_every_mfma_variant(ctx, oracle):
    # each tile variant the autotuner may pick, on a ragged shape, every precision
    rng = W.SplitMix64(77)
    M, N, K = 261, 131, 700
    n = int(np.ceil(np.sqrt(K)))
    t = W.make_tissue(rng, n, 20.0, 30.0).slice(0, K)
    traj = W.CartesianTraj(M, 1, 5e-6, np.zeros(M, np.complex64), float(np.float32(2 * np.pi / 20.0)))
    coils = ((rng.normal(N * K) + 1j * rng.normal(N * K)) / 3).reshape(N, K).astype(np.complex64)
    echos = ((rng.normal(M * K) + 1j * rng.normal(M * K)) / 3).reshape(M, K).astype(np.complex64)
    context()
    params = ca.TissueParameters.from_tissue(t)
    ctraj = ca.CartesianTrajectory.from_trajectory(traj)
    want = oracle.signal_cartesian(t, traj, coils, echos, "f64")
    try:
        for v in range(ca.contraction_variant_count()):
            ca.set_contraction_variant(v)
            for method, tol in (("matmul_fast", 1e-4), ("matmul_pedantic", 1e-5), ("matmul_low", 2e-2)):
                got = ca.magnetization_to_signal(echos, params, ctraj, coils, method=method).collect()
                assert rel(got, want) < tol, (v, method)
    finally:
        ca.set_contraction_variant(-1)


def test_cartesian_signal_default_abi_call(ctx, oracle):
    wl = W.config(1, scale=2)
    params, seq, traj = objects(wl)
    echos = oracle.fisp(wl.tissue, wl.sequence, "f64").astype(np.complex64)
    got = ca.magnetization_to_signal(echos, params, traj, wl.coils).collect()
    assert rel(got, oracle.signal_cartesian(wl.tissue, wl.trajectory, wl.coils, echos, "f64")) < SIG_TOL


def test_cartesian_signal_low_precision_is_coarse(ctx, oracle):
    wl = W.config(1, scale=2)
    params, seq, traj = objects(wl)
    echos = oracle.fisp(wl.tissue, wl.sequence, "f64").astype(np.complex64)
    got = ca.magnetization_to_signal(echos, params, traj, wl.coils, low_precision=True).collect()
    assert rel(got, oracle.signal_cartesian(wl.tissue, wl.trajectory, wl.coils, echos, "f64")) < 2e-2


def test_spiral_signal(ctx, oracle):
    banner("spiral signal")
    wl = W.config(3, scale=8)
    params, seq, traj = objects(wl)
    rng = W.SplitMix64(42)
    nv, nr = wl.tissue.nvoxels, wl.trajectory.nreadouts
    echos = ((rng.normal(nr * nv) + 1j * rng.normal(nr * nv)) * 0.1).reshape(nr, nv).astype(np.complex64)
    got = ca.magnetization_to_signal(echos, params, traj, wl.coils).collect()
    want = oracle.signal_spiral(wl.tissue, wl.trajectory, wl.coils, echos, "f64")
    assert rel(got, want) < SIG_TOL


def spiral6_inputs(nc, nr, ns, nv, seed):
    rng = W.SplitMix64(seed)
    n = int(np.ceil(np.sqrt(nv)))
    t = W.make_tissue(rng, n, 22.4, 50.0).slice(0, nv)
    traj = W.make_spiral(nr, ns, 22.4, n, 5e-6)

    def cn(*shape, s=1.0):
        k = int(np.prod(shape))
        return ((rng.normal(k) + 1j * rng.normal(k)) * s).reshape(shape).astype(np.complex64)

    coils = cn(nc, nv, s=0.5)
    echos = cn(nr, nv, s=0.1)
    ds = [cn(nr, nv, s=0.01), cn(nr, nv, s=0.1), cn(nr, nv, s=0.1), cn(nr, nv, s=1e-3)]  # T1, T2, B1, B0
    v = rng.normal(6 * nv).reshape(6, nv).astype(np.complex64)  # real perturbation, stored complex
    y = cn(nc, nr, ns, s=0.7)
    return t, traj, coils, echos, ds, v, y


# shapes: coil groups of four (1, 4, 5, 8 coils: one or two groups on MFMA) and
# more than eight (the VALU form), samples across the 256-sample LDS chunks and
# not a multiple of 16, voxels not a multiple of the 32-voxel wave tile
SPIRAL_SHAPES = [(4, 70, 64, 784), (1, 33, 50, 333), (5, 41, 37, 1000), (8, 16, 130, 257), (3, 9, 600, 97),
                 (10, 7, 40, 70), (4, 5, 512, 300), (2, 3, 1024, 130)]


@pytest.mark.parametrize("nc,nr,ns,nv", [(4, 6, 512, 300), (8, 3, 1024, 200), (3, 4, 496, 150)])
def test_spiral_signal_sample_passes(ctx, oracle, nc, nr, ns, nv):
    # sample counts that fill every 512-sample pass of the MFMA spiral sum
    # (its tile loop without tests) and one that does not
    banner(f"spiral signal coils={nc} nr={nr} ns={ns} nv={nv}")
    t, traj, coils, echos, _, _, _ = spiral6_inputs(nc, nr, ns, nv, 11 * nc + ns)
    context()
    params = ca.TissueParameters.from_tissue(t)
    straj = ca.SpiralTrajectory.from_trajectory(traj)
    got = ca.magnetization_to_signal(echos, params, straj, coils).collect()
    assert got.shape == (nc, nr, ns)
    assert rel(got, oracle.signal_spiral(t, traj, coils, echos, "f64")) < SIG_TOL


@pytest.mark.parametrize("method", ["direct", "gemm_fast"])
@pytest.mark.parametrize("nc,nr,ns,nv", SPIRAL_SHAPES)
def test_spiral_jacobian_six_parameters(ctx, oracle, nc, nr, ns, nv, method):
    # J.v and J^H.y on a golden-angle spiral with (T1, T2, B1, B0, rho_x, rho_y)
    banner(f"spiral J 6 params coils={nc} nr={nr} ns={ns} nv={nv} {method}")
    t, traj, coils, echos, ds, v, y = spiral6_inputs(nc, nr, ns, nv, 7 * nc + nr)
    context()
    params = ca.TissueParameters.from_tissue(t)
    straj = ca.SpiralTrajectory.from_trajectory(traj)
    jv = ca.compute_jacobian_spiral(echos, ds, params, straj, coils, v, method=method).collect()
    assert rel(jv, oracle.jacobian_spiral(t, traj, coils, echos, *ds, v, "f64")) < SIG_TOL
    jh = ca.compute_jacobian_hermitian_spiral(echos, ds, params, straj, coils, y, method=method).collect()
    assert rel(jh, oracle.jacobian_hermitian_spiral(t, traj, coils, echos, *ds, y, "f64")) < SIG_TOL
    # the two GPU products are adjoint
    lhs = np.vdot(y.ravel().astype(np.complex128), jv.ravel())
    rhs = np.vdot(jh.ravel().astype(np.complex128), v.ravel().astype(np.complex128))
    assert abs(lhs - rhs) / abs(lhs) < SIG_TOL


def test_spiral_hermitian_single_product_is_coarse(ctx, oracle):
    # gemm_low: BF16 hi only on MFMA -- a few parts in 1e3, far from the hi + lo result
    t, traj, coils, echos, ds, v, y = spiral6_inputs(4, 20, 64, 300, 3)
    context()
    params = ca.TissueParameters.from_tissue(t)
    straj = ca.SpiralTrajectory.from_trajectory(traj)
    want = oracle.jacobian_hermitian_spiral(t, traj, coils, echos, *ds, y, "f64")
    low = ca.compute_jacobian_hermitian_spiral(echos, ds, params, straj, coils, y, method="gemm_low").collect()
    fast = ca.compute_jacobian_hermitian_spiral(echos, ds, params, straj, coils, y).collect()
    assert rel(fast, want) < SIG_TOL
    assert 1e-4 < rel(low, want) < 3e-2


def test_spiral_jacobian_shape_errors(ctx):
    t, traj, coils, echos, ds, v, y = spiral6_inputs(2, 10, 16, 64, 5)
    context()
    params = ca.TissueParameters.from_tissue(t)
    straj = ca.SpiralTrajectory.from_trajectory(traj)
    with pytest.raises(ca.CompasError):
        ca.compute_jacobian_spiral(echos, ds, params, straj, coils, v[:4])  # four rows
    with pytest.raises(ca.CompasError):
        ca.compute_jacobian_hermitian_spiral(echos, ds, params, straj, coils, y[:, :5])
    with pytest.raises(ca.CompasError):
        ca.compute_jacobian_spiral(echos[:5], ds, params, straj, coils, v)


def test_phase_encoding(ctx, oracle):
    wl = W.config(1, scale=2)
    params, seq, traj = objects(wl)
    echos = oracle.fisp(wl.tissue, wl.sequence, "f64").astype(np.complex64)
    got = ca.phase_encoding(echos, params, traj).collect()
    assert rel(got, oracle.phase_encoding(wl.tissue, wl.trajectory, echos, "f64")) < 1e-6


def test_signal_shape_errors(ctx):
    wl = W.config(1, scale=4)
    params, seq, traj = objects(wl)
    bad = np.zeros((3, wl.tissue.nvoxels), np.complex64)
    with pytest.raises(ca.CompasError):
        ca.magnetization_to_signal(bad, params, traj, wl.coils)


# ------------------------------------------------------------ jacobian ---

def jac_inputs(oracle, scale=8):
    wl = W.config(2, scale=scale)
    rng = W.SplitMix64(9)
    nv, nr, ns = wl.tissue.nvoxels, wl.trajectory.nreadouts, wl.trajectory.samples_per_readout
    echos = oracle.fisp(wl.tissue, wl.sequence, "f64").astype(np.complex64)
    d1 = (rng.normal(nr * nv) * 0.01 + 1j * rng.normal(nr * nv) * 0.01).reshape(nr, nv).astype(np.complex64)
    d2 = (rng.normal(nr * nv) * 0.1 + 1j * rng.normal(nr * nv) * 0.1).reshape(nr, nv).astype(np.complex64)
    v = (rng.normal(4 * nv) + 1j * rng.normal(4 * nv)).reshape(4, nv).astype(np.complex64)
    nc = wl.coils.shape[0]
    y = (rng.normal(nc * nr * ns) + 1j * rng.normal(nc * nr * ns)).reshape(nc, nr, ns).astype(np.complex64)
    return wl, echos, d1, d2, v, y


@pytest.mark.parametrize("method", ["direct", "gemm_fast"])
def test_jacobian_product(ctx, oracle, method):
    banner(f"J.v {method}")
    wl, echos, d1, d2, v, y = jac_inputs(oracle)
    params, seq, traj = objects(wl)
    got = ca.compute_jacobian(echos, (d1, d2), params, traj, wl.coils, v, method=method).collect()
    want = oracle.jacobian(wl.tissue, wl.trajectory, wl.coils, echos, d1, d2, v, "f64")
    assert rel(got, want) < SIG_TOL


@pytest.mark.parametrize("method", ["direct", "gemm_fast"])
def test_jacobian_hermitian(ctx, oracle, method):
    banner(f"J^H.y {method}")
    wl, echos, d1, d2, v, y = jac_inputs(oracle)
    params, seq, traj = objects(wl)
    got = ca.compute_jacobian_hermitian(echos, (d1, d2), params, traj, wl.coils, y, method=method).collect()
    want = oracle.jacobian_hermitian(wl.tissue, wl.trajectory, wl.coils, echos, d1, d2, y, "f64")
    assert rel(got, want) < SIG_TOL


@pytest.mark.parametrize("nc,nr,ns,nv", [(1, 77, 50, 333), (2, 64, 51, 257), (3, 77, 50, 333), (4, 33, 64, 70),
                                         (5, 130, 49, 1000), (8, 41, 32, 129)])
def test_jacobian_hermitian_coil_groups(ctx, oracle, nc, nr, ns, nv):
    # the fused MFMA path packs four coils per row group (padding the last group),
    # one or two coils take the per-coil path; ragged readouts, samples and voxels
    banner(f"J^H.y coils={nc} nr={nr} ns={ns} nv={nv}")
    rng = W.SplitMix64(1000 * nc + nr)
    n = int(np.ceil(np.sqrt(nv)))
    t = W.make_tissue(rng, n, 20.0, 40.0).slice(0, nv)
    traj = W.CartesianTraj(nr, ns, 5e-6, np.zeros(nr, np.complex64), float(np.float32(2 * np.pi / 20.0)))

    def cn(*shape, s=1.0):
        k = int(np.prod(shape))
        return ((rng.normal(k) + 1j * rng.normal(k)) * s).reshape(shape).astype(np.complex64)

    coils, echos, d1, d2, y = cn(nc, nv, s=0.5), cn(nr, nv, s=0.3), cn(nr, nv, s=0.01), cn(nr, nv, s=0.1), cn(nc, nr, ns)
    context()
    params = ca.TissueParameters.from_tissue(t)
    ctraj = ca.CartesianTrajectory.from_trajectory(traj)
    want = oracle.jacobian_hermitian(t, traj, coils, echos, d1, d2, y, "f64")
    got = ca.compute_jacobian_hermitian(echos, (d1, d2), params, ctraj, coils, y, method="gemm_fast").collect()
    assert rel(got, want) < SIG_TOL
    try:
        ca.set_contraction_precision("fp16x3")
        got = ca.compute_jacobian_hermitian(echos, (d1, d2), params, ctraj, coils, y, method="gemm_fast").collect()
    finally:
        ca.set_contraction_precision(None)
    assert rel(got, want) < 1e-5


@pytest.mark.parametrize("nc,nr,ns,nv", [(1, 77, 4, 333), (3, 5, 8, 1), (2, 64, 52, 257), (5, 130, 48, 1000),
                                         (8, 41, 36, 129), (4, 33, 51, 70)])
def test_jacobian_product_ragged(ctx, oracle, nc, nr, ns, nv):
    # J.v on ragged readouts, samples, coils and voxels (down to one voxel and
    # four samples), every method against FP64 and the MFMA methods against the
    # FP32 path
    banner(f"J.v coils={nc} nr={nr} ns={ns} nv={nv}")
    rng = W.SplitMix64(7000 * nc + nr + ns)
    n = int(np.ceil(np.sqrt(nv)))
    t = W.make_tissue(rng, n, 20.0, 40.0).slice(0, nv)
    traj = W.CartesianTraj(nr, ns, 5e-6, np.zeros(nr, np.complex64), float(np.float32(2 * np.pi / 20.0)))

    def cn(*shape, s=1.0):
        k = int(np.prod(shape))
        return ((rng.normal(k) + 1j * rng.normal(k)) * s).reshape(shape).astype(np.complex64)

    coils, echos, d1, d2, v = cn(nc, nv, s=0.5), cn(nr, nv, s=0.3), cn(nr, nv, s=0.01), cn(nr, nv, s=0.1), cn(4, nv)
    context()
    params = ca.TissueParameters.from_tissue(t)
    ctraj = ca.CartesianTrajectory.from_trajectory(traj)
    want = oracle.jacobian(t, traj, coils, echos, d1, d2, v, "f64")
    plain = ca.compute_jacobian(echos, (d1, d2), params, ctraj, coils, v, method="direct").collect()
    assert rel(plain, want) < 1e-5
    for method, tol in (("gemm_fast", SIG_TOL), ("gemm_low", 2e-2)):
        got = ca.compute_jacobian(echos, (d1, d2), params, ctraj, coils, v, method=method).collect()
        assert got.shape == (nc, nr, ns)
        assert rel(got, want) < tol, method
        assert rel(got, plain) < tol, method


PRECISION_TOL = {"fp32": 1e-5, "bf16x3": 1e-4, "fp16x3": 1e-5, "bf16x1": 2e-2, "fp16x1": 5e-3}


@pytest.mark.parametrize("precision", list(PRECISION_TOL))
def test_precision_override_every_output(ctx, oracle, precision):
    # signal, J.v and J^H.y with the contraction precision forced, against FP64
    banner(f"precision override {precision}")
    wl, echos, d1, d2, v, y = jac_inputs(oracle)
    params, seq, traj = objects(wl)
    tol = PRECISION_TOL[precision]
    try:
        ca.set_contraction_precision(precision)
        s = ca.magnetization_to_signal(echos, params, traj, wl.coils).collect()
        jv = ca.compute_jacobian(echos, (d1, d2), params, traj, wl.coils, v).collect()
        jh = ca.compute_jacobian_hermitian(echos, (d1, d2), params, traj, wl.coils, y).collect()
    finally:
        ca.set_contraction_precision(None)
    assert rel(s, oracle.signal_cartesian(wl.tissue, wl.trajectory, wl.coils, echos, "f64")) < tol
    assert rel(jv, oracle.jacobian(wl.tissue, wl.trajectory, wl.coils, echos, d1, d2, v, "f64")) < tol
    assert rel(jh, oracle.jacobian_hermitian(wl.tissue, wl.trajectory, wl.coils, echos, d1, d2, y, "f64")) < tol


def test_jacobian_adjoint_on_gpu(ctx, oracle):
    wl, echos, d1, d2, v, y = jac_inputs(oracle)
    params, seq, traj = objects(wl)
    Jv = ca.compute_jacobian(echos, (d1, d2), params, traj, wl.coils, v).collect().astype(np.complex128)
    JHy = ca.compute_jacobian_hermitian(echos, (d1, d2), params, traj, wl.coils, y).collect().astype(np.complex128)
    lhs = np.vdot(y.ravel().astype(np.complex128), Jv.ravel())
    rhs = np.vdot(JHy.ravel(), v.ravel().astype(np.complex128))
    assert abs(lhs - rhs) / abs(lhs) < SIG_TOL


def test_jacobian_shape_errors(ctx, oracle):
    wl, echos, d1, d2, v, y = jac_inputs(oracle, scale=16)
    params, seq, traj = objects(wl)
    with pytest.raises(ca.CompasError):
        ca.compute_jacobian(echos, (d1, d2), params, traj, wl.coils, v[:3])
    with pytest.raises(ca.CompasError):
        ca.compute_jacobian_hermitian(echos, (d1, d2), params, traj, wl.coils, y[:, :, :-1])


# ---------------------------------------------------- derivative, misc ---

@pytest.mark.parametrize("field", [ca.T1, ca.T2, ca.B1, ca.B0])
def test_fisp_finite_difference_derivative(ctx, oracle, field):
    t, seq = small_fisp(nv=128, nTR=40, N=16)
    context()
    p = ca.TissueParameters.from_tissue(t)
    s = ca.FispSequence.from_sequence(seq)
    echos = ca.simulate_magnetization(p, s)
    delta = 1e-4
    got = ca.simulate_magnetization_derivative(field, echos, p, s, delta).collect()
    name = ["T1", "T2", "B1", "B0"][field]
    if field == ca.B0:
        # 1e-4 Hz moves the phase by ~6e-6 rad per TR: below FP32 resolution of m, so B0
        # takes a step that the FP32 forward difference can resolve
        delta = 0.05
        got = ca.simulate_magnetization_derivative(field, echos, p, s, delta).collect()
    tp = t.replace(**{name: (getattr(t, name) + np.float32(delta)).astype(np.float32)})
    want = (oracle.fisp(tp, seq, "f64") - oracle.fisp(t, seq, "f64")) / np.float32(delta)
    # forward difference of FP32 magnetisation at delta = 1e-4: cancellation error ~1e-3
    assert rel(got, want) < 5e-3


@pytest.mark.parametrize("field", [ca.T1, ca.T2, ca.B1, ca.B0])
def test_pssfp_finite_difference_derivative(ctx, oracle, field):
    """compas_simulate_magnetization_derivative_pssfp (derivative.cu:11-94):
    (sim(p + delta e_field) - echos) / delta with p + delta formed in float32,
    against the same forward difference of the FP64 oracle."""
    name = ["T1", "T2", "B1", "B0"][field]
    banner(f"pssfp finite difference {name}")
    t = W.make_tissue(W.SplitMix64(31), 12, 20.0, 30.0)
    seq = W.make_pssfp(30, 60.0, nz=5, nRF=5)
    context()
    p = ca.TissueParameters.from_tissue(t)
    s = ca.pSSFPSequence.from_sequence(seq)
    echos = ca.simulate_magnetization(p, s)
    # 1e-4 Hz moves the phase below FP32 resolution of m: B0 takes a resolvable step
    delta = 0.05 if field == ca.B0 else 1e-4
    got = ca.simulate_magnetization_derivative(field, echos, p, s, delta).collect()
    tp = t.replace(**{name: (getattr(t, name) + np.float32(delta)).astype(np.float32)})
    want = (oracle.pssfp(tp, seq, "f64") - oracle.pssfp(t, seq, "f64")) / np.float32(delta)
    err = rel(got, want)
    print(f"pssfp forward difference {name}: rel-L2 {err:.2e} vs the FP64 forward difference")
    # forward difference of FP32 magnetisation at delta = 1e-4: cancellation error ~1e-3
    assert err < 5e-3


def central_difference(simulate, t, seq, name):
    """FP64 oracle central difference of the echoes w.r.t. one tissue field.

    Steps are exact in float32 (the oracle reads float32 tissue): 2^-16 of the
    value's binade for T1/T2, 2^-16 for B1, 2^-14 Hz for B0; the denominator is
    the difference actually taken.  At these steps the truncation error (h^2/6
    times the third derivative) and the FP64 rounding of the two runs divided
    by 2h both stay below 1e-8 relative.
    """
    v = getattr(t, name).astype(np.float32)
    if name in ("T1", "T2"):
        h = np.exp2(np.floor(np.log2(v)) - 16).astype(np.float32)
    elif name == "B1":
        h = np.full_like(v, 2.0 ** -16)
    else:
        h = np.full_like(v, 2.0 ** -14)
    vp = (v + h).astype(np.float32)
    vm = (v - h).astype(np.float32)
    fp = simulate(t.replace(**{name: vp}), seq, "f64")
    fm = simulate(t.replace(**{name: vm}), seq, "f64")
    return (fp - fm) / (vp.astype(np.float64) - vm.astype(np.float64))[None, :]


TANGENT_TOL = 1e-4  # FP32 recursion against the FP64 derivative


@pytest.mark.parametrize("field", [ca.T1, ca.T2, ca.B1, ca.B0])
@pytest.mark.parametrize("N", [3, 16, 25, 100])
def test_fisp_tangent_derivative(ctx, oracle, field, N):
    name = ["T1", "T2", "B1", "B0"][field]
    banner(f"fisp tangent {name} N={N}")
    t, seq = small_fisp(nv=97, nTR=40, N=N)
    context()
    p = ca.TissueParameters.from_tissue(t)
    s = ca.FispSequence.from_sequence(seq)
    got = ca.simulate_magnetization_tangent(field, p, s).collect()
    assert got.shape == (40, 97)
    assert rel(got, central_difference(oracle.fisp, t, seq, name)) < TANGENT_TOL


@pytest.mark.parametrize("field", [ca.T1, ca.T2, ca.B1, ca.B0])
def test_fisp_tangent_sequence_options(ctx, oracle, field):
    name = ["T1", "T2", "B1", "B0"][field]
    banner(f"fisp tangent {name} usf/reps/TW/slices")
    t, seq = small_fisp(nv=70, nTR=30, N=12)
    seq.undersampling = 2
    seq.repetitions = 3
    seq.TW = 0.12
    seq.profiles = np.stack([np.linspace(0.6, 1.1, 30), np.full(30, 0.9 - 0.2j)]).astype(np.complex64)
    context()
    p = ca.TissueParameters.from_tissue(t)
    for inversion, spoil in ((True, True), (False, False)):
        seq.inversion = inversion
        seq.wait_spoiling = spoil
        s = ca.FispSequence.from_sequence(seq)
        got = ca.simulate_magnetization_tangent(field, p, s).collect()
        assert got.shape == (60, 70)
        assert rel(got, central_difference(oracle.fisp, t, seq, name)) < TANGENT_TOL


@pytest.mark.parametrize("field", [ca.T1, ca.T2, ca.B1, ca.B0])
def test_fisp_tangent_echo_paths(ctx, oracle, field):
    # as test_fisp_echo_paths for the tangents: TE = 0.2 ms, T2 = 0.15 ms in
    # voxels 0..63 (whole waves, whatever the lanes per voxel) and in voxel
    # 100 (one lane of a wave whose other voxels sample from the updated state)
    name = ["T1", "T2", "B1", "B0"][field]
    banner(f"fisp tangent {name} echo paths (short T2)")
    t, seq = small_fisp(nv=150, nTR=40, N=12)
    seq.TE = 0.0002
    t.T2[:64] = 1.5e-4
    t.T2[100] = 1.5e-4
    context()
    p = ca.TissueParameters.from_tissue(t)
    s = ca.FispSequence.from_sequence(seq)
    got = ca.simulate_magnetization_tangent(field, p, s).collect()
    want = central_difference(oracle.fisp, t, seq, name)
    assert rel(got, want) < TANGENT_TOL
    for cols in (slice(0, 64), slice(64, 150)):
        assert rel(got[:, cols], want[:, cols]) < TANGENT_TOL, cols


def test_fisp_tangent_b0_zero_and_errors(ctx, oracle):
    t, seq = small_fisp(nv=50, nTR=30, N=10, b0=None)
    context()
    p = ca.TissueParameters.from_tissue(t)
    s = ca.FispSequence.from_sequence(seq)
    got = ca.simulate_magnetization_tangent(ca.B0, p, s).collect()
    assert rel(got, central_difference(oracle.fisp, t, seq, "B0")) < TANGENT_TOL
    with pytest.raises(ca.CompasError, match="tangent"):
        ca.simulate_magnetization_tangent(4, p, s)


@pytest.mark.parametrize("field", [ca.T1, ca.T2, ca.B1, ca.B0])
@pytest.mark.parametrize("nz,nRF", [(1, 5), (7, 25), (35, 3), (40, 2)])
def test_pssfp_tangent_derivative(ctx, oracle, field, nz, nRF):
    # nz = 40 runs two passes of 32 slice positions; nRF = 2 at 60 degrees
    # takes the sincos branch of the rotation and its closed-form derivative
    name = ["T1", "T2", "B1", "B0"][field]
    banner(f"pssfp tangent {name} nz={nz} nRF={nRF}")
    t = W.make_tissue(W.SplitMix64(6), 9, 20.0, 40.0)
    seq = W.make_pssfp(24, 60.0, nz=nz, nRF=nRF)
    context()
    p = ca.TissueParameters.from_tissue(t)
    got = ca.simulate_magnetization_tangent(field, p, ca.pSSFPSequence.from_sequence(seq)).collect()
    assert got.shape == (24, 81)
    assert rel(got, central_difference(oracle.pssfp, t, seq, name)) < TANGENT_TOL


# field sets of one fused pass (all four, T1 + T2) and of several passes
FIELD_SETS = [(ca.T1, ca.T2, ca.B1, ca.B0), (ca.T1, ca.T2), (ca.T2, ca.B0, ca.T1), (ca.B1,), (ca.B0, ca.T1)]


@pytest.mark.parametrize("fields", FIELD_SETS, ids=lambda f: "".join("T1 T2 B1 B0".split()[i] for i in f))
@pytest.mark.parametrize("echos", [False, True])
@pytest.mark.parametrize("nz,nRF", [(7, 25), (40, 2)])
def test_pssfp_fused_tangents(ctx, oracle, fields, echos, nz, nRF):
    # several tangents (and the echoes) from one call agree with one field
    # at a time, with the central differences of the oracle and, for the
    # echoes, with the plain simulation
    names = ["T1", "T2", "B1", "B0"]
    banner(f"pssfp fused tangents {[names[f] for f in fields]} echos={echos} nz={nz} nRF={nRF}")
    t = W.make_tissue(W.SplitMix64(6), 9, 20.0, 40.0)
    seq = W.make_pssfp(24, 60.0, nz=nz, nRF=nRF)
    context()
    p = ca.TissueParameters.from_tissue(t)
    s = ca.pSSFPSequence.from_sequence(seq)
    out = ca.simulate_magnetization_tangents(fields, p, s, echos=echos)
    assert len(out) == len(fields) + int(echos)
    if echos:
        e = out.pop(0).collect()
        assert e.shape == (24, 81)
        assert rel(e, ca.simulate_magnetization(p, s).collect()) < 1e-6
    for f, d in zip(fields, out):
        got = d.collect()
        assert got.shape == (24, 81)
        assert rel(got, ca.simulate_magnetization_tangent(f, p, s).collect()) < 1e-6
        assert rel(got, central_difference(oracle.pssfp, t, seq, names[f])) < TANGENT_TOL


@pytest.mark.parametrize("fields", FIELD_SETS, ids=lambda f: "".join("T1 T2 B1 B0".split()[i] for i in f))
@pytest.mark.parametrize("echos", [False, True])
@pytest.mark.parametrize("N", [5, 25, 128])
def test_fisp_fused_tangents(ctx, oracle, fields, echos, N):
    # one FISP pass carrying several tangents (two states per lane with four,
    # up to 64 lanes per voxel at N = 128) against one field at a time and the
    # oracle's central differences, with undersampling, repetitions and TW
    names = ["T1", "T2", "B1", "B0"]
    banner(f"fisp fused tangents {[names[f] for f in fields]} echos={echos} N={N}")
    t, seq = small_fisp(nv=61, nTR=24, N=N)
    seq.undersampling = 2
    seq.repetitions = 2
    seq.TW = 0.1
    seq.profiles = np.stack([np.linspace(0.7, 1.1, 24), np.full(24, 0.9 - 0.2j)]).astype(np.complex64)
    context()
    p = ca.TissueParameters.from_tissue(t)
    fs = ca.FispSequence.from_sequence(seq)
    out = ca.simulate_magnetization_tangents(fields, p, fs, echos=echos)
    assert len(out) == len(fields) + int(echos)
    if echos:
        e = out.pop(0).collect()
        assert e.shape == (48, 61)
        assert rel(e, oracle.fisp(t, seq, "f64")) < MAG_TOL
    for f, d in zip(fields, out):
        got = d.collect()
        assert got.shape == (48, 61)
        assert rel(got, ca.simulate_magnetization_tangent(f, p, fs).collect()) < 1e-6
        assert rel(got, central_difference(oracle.fisp, t, seq, names[f])) < TANGENT_TOL


def test_fused_tangents_fisp_and_errors(ctx):
    t, seq = small_fisp(nv=50, nTR=30, N=10)
    context()
    p = ca.TissueParameters.from_tissue(t)
    fs = ca.FispSequence.from_sequence(seq)
    e, d2, d1 = (a.collect() for a in ca.simulate_magnetization_tangents((ca.T2, ca.T1), p, fs, echos=True))
    assert rel(e, ca.simulate_magnetization(p, fs).collect()) < 1e-6
    assert rel(d1, ca.simulate_magnetization_tangent(ca.T1, p, fs).collect()) < 1e-6
    assert rel(d2, ca.simulate_magnetization_tangent(ca.T2, p, fs).collect()) < 1e-6
    assert ca.simulate_magnetization_tangents((), p, fs) == []
    e, = ca.simulate_magnetization_tangents((), p, fs, echos=True)
    assert np.array_equal(e.collect(), ca.simulate_magnetization(p, fs).collect())

    tp = W.make_tissue(W.SplitMix64(3), 4, 20.0, None)
    pp = ca.TissueParameters.from_tissue(tp)
    ps = ca.pSSFPSequence.from_sequence(W.make_pssfp(12, 40.0, nz=3, nRF=4))
    with pytest.raises(ca.CompasError, match="twice"):
        ca.simulate_magnetization_tangents((ca.T1, ca.T1), pp, ps)
    with pytest.raises(ca.CompasError, match="twice"):
        ca.simulate_magnetization_tangents((ca.B1, ca.B1), p, fs)
    with pytest.raises(ca.CompasError, match="tangent"):
        ca.simulate_magnetization_tangents((ca.T1, 4), pp, ps)
    e, = ca.simulate_magnetization_tangents((), pp, ps, echos=True)
    assert np.array_equal(e.collect(), ca.simulate_magnetization(pp, ps).collect())


def test_empty_inputs(ctx):
    # sums over no voxels are zero; per-voxel outputs are empty; no slice
    # profiles (FISP) or slice positions (pSSFP) give zero echoes
    banner("empty inputs")
    context()
    t0 = W.make_tissue(W.SplitMix64(2), 4, 20.0, 30.0).slice(0, 0)
    p0 = ca.TissueParameters.from_tissue(t0)
    _, fseq = small_fisp(nv=4, nTR=12, N=6)
    fs = ca.FispSequence.from_sequence(fseq)
    pseq = ca.pSSFPSequence.from_sequence(W.make_pssfp(12, 40.0, nz=3, nRF=4))
    assert ca.simulate_magnetization(p0, fs).collect().shape == (12, 0)
    assert ca.simulate_magnetization(p0, pseq).collect().shape == (12, 0)
    for f in (ca.T1, ca.T2, ca.B1, ca.B0):
        assert ca.simulate_magnetization_tangent(f, p0, fs).collect().shape == (12, 0)
        assert ca.simulate_magnetization_tangent(f, p0, pseq).collect().shape == (12, 0)

    nc, nr, ns = 2, 12, 5
    echos = np.zeros((nr, 0), np.complex64)
    coils = np.zeros((nc, 0), np.complex64)
    cart = ca.CartesianTrajectory.from_trajectory(
        W.CartesianTraj(nr, ns, 5e-6, np.zeros(nr, np.complex64), float(np.float32(2 * np.pi / 20.0))))
    spiral = ca.SpiralTrajectory.from_trajectory(W.make_spiral(nr, ns, 20.0, 4, 5e-6))
    for method in ("direct", "matmul_fast"):
        s = ca.magnetization_to_signal(echos, p0, cart, coils, method=method).collect()
        assert s.shape == (nc, nr, ns) and not np.any(s)
    s = ca.magnetization_to_signal(echos, p0, spiral, coils).collect()
    assert s.shape == (nc, nr, ns) and not np.any(s)
    v4 = np.zeros((4, 0), np.complex64)
    y = np.ones((nc, nr, ns), np.complex64)
    for method in ("direct", "gemm_fast"):
        jv = ca.compute_jacobian(echos, (echos, echos), p0, cart, coils, v4, method=method).collect()
        assert jv.shape == (nc, nr, ns) and not np.any(jv)
        assert ca.compute_jacobian_hermitian(echos, (echos, echos), p0, cart, coils, y,
                                             method=method).collect().shape == (4, 0)
    d4 = [echos] * 4
    jv = ca.compute_jacobian_spiral(echos, d4, p0, spiral, coils, np.zeros((6, 0), np.complex64)).collect()
    assert jv.shape == (nc, nr, ns) and not np.any(jv)
    assert ca.compute_jacobian_hermitian_spiral(echos, d4, p0, spiral, coils, y).collect().shape == (6, 0)

    # no slice profiles / no slice positions: zero echoes
    t, fseq = small_fisp(nv=70, nTR=12, N=6)
    fseq.profiles = np.zeros((0, 12), np.complex64)
    p = ca.TissueParameters.from_tissue(t)
    e = ca.simulate_magnetization(p, ca.FispSequence.from_sequence(fseq)).collect()
    assert e.shape == (12, 70) and not np.any(e)
    pz = W.make_pssfp(12, 40.0, nz=3, nRF=4)
    pz.z = np.zeros(0, np.float32)
    e = ca.simulate_magnetization(p, ca.pSSFPSequence.from_sequence(pz)).collect()
    assert e.shape == (12, 70) and not np.any(e)


def test_residual(ctx):
    rng = W.SplitMix64(3)
    a = (rng.normal(2 * 30 * 17) + 1j * rng.normal(2 * 30 * 17)).reshape(2, 30, 17).astype(np.complex64)
    b = (rng.normal(2 * 30 * 17) + 1j * rng.normal(2 * 30 * 17)).reshape(2, 30, 17).astype(np.complex64)
    obj, diff = ca.compute_residual(a, b)
    np.testing.assert_allclose(diff.collect(), a - b, rtol=0, atol=1e-6)
    assert abs(obj - np.sum(np.abs(a - b) ** 2)) / obj < 1e-5


def test_arrays_round_trip(ctx):
    x = (np.arange(24).reshape(2, 3, 4) + 0.5j).astype(np.complex64)
    assert np.array_equal(ca.CompasArray.from_numpy(x).collect(), x)
    f = np.linspace(0, 1, 7, dtype=np.float32)
    assert np.array_equal(ca.CompasArray.from_numpy(f).collect(), f)
