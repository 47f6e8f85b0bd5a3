"""Pins of the oracle (oracle/) against closed forms, brute-force ensembles and
library routines.  CPU only.

The reference ships no golden vectors or known-answer tests for this path
(its tests/ holds one empty Catch2 case), so each oracle function is pinned
here by mathematics that a plausible mistake would break:

  FISP  : closed-form first echo; zero flip angle; brute-force isochromat
          ensemble (scipy rotations) for N > nTR; repetitions/TW; slice-profile
          linearity; undersampling expansion; N-independence when N > nTR;
          truncation (N < nTR): N = 1 equals the ideal-spoiling FLASH
          recursion, and echoes 0..2N-1 are exact with echo 2N the first
          changed (a pin that both off-by-one truncations fail).
  pSSFP : closed form of the relaxation-free alternating train; scipy
          rotation-vector ensemble; norm invariance.
  signal: reduction to a centred DFT (numpy.fft); contraction = complex matmul
          of the per-voxel factors; spiral == Cartesian on a Cartesian grid.
  J, J^H: rho columns equal the signal with rho := 1; exact adjoint identity;
          T1/T2 rows equal FP64 central differences of the simulated signal.
"""

import numpy as np
import pytest

from compas_amd import workloads as W
from isochromat_ensemble import fisp_ensemble, pssfp_isochromat


def f32(x):
    """The value a float argument takes after crossing the float32 C interface."""
    return float(np.float32(x))


def rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def small_tissue(nv, seed=7, b0=20.0, n=None):
    rng = W.SplitMix64(seed)
    n = n or int(np.ceil(np.sqrt(nv)))
    t = W.make_tissue(rng, n, 20.0, b0)
    return t.slice(0, nv)


# ---------------------------------------------------------------- FISP ---

def test_fisp_first_echo_closed_form(oracle):
    t = small_tissue(64)
    seq = W.FispSeq(RF=(np.array([30.0 * np.exp(0.4j)] * 5)).astype(np.complex64),
                    profiles=np.ones((1, 5), np.complex64), TR=0.01, TE=0.004, max_state=8, TI=0.03)
    for prec, tol in (("f64", 1e-12), ("f32", 2e-6)):
        e = oracle.fisp(t, seq, prec)
        T1, T2, B1, B0 = (getattr(t, f).astype(np.float64) for f in ("T1", "T2", "B1", "B0"))
        rf = np.complex128(np.complex64(30.0 * np.exp(0.4j)))
        alpha = np.deg2rad(abs(rf)) * B1
        phi = np.angle(rf)
        want = (-1j * np.exp(1j * phi) * np.sin(alpha) * (1 - 2 * np.exp(-f32(seq.TI) / T1))
                * np.exp(-f32(seq.TE) / T2) * np.exp(1j * 2 * np.pi * B0 * f32(seq.TE)))
        assert np.max(np.abs(e[0] - want)) < tol


def test_fisp_zero_flip_is_zero(oracle):
    t = small_tissue(32)
    seq = W.make_fisp(20, 0.0, 8)
    assert np.all(oracle.fisp(t, seq, "f64") == 0)


@pytest.mark.parametrize("inversion", [True, False])
def test_fisp_matches_isochromat_ensemble(oracle, inversion):
    nTR = 24
    t = small_tissue(6, seed=11, b0=30.0)
    rng = W.SplitMix64(99)
    RF = (rng.uniform(nTR, 5, 80) * np.exp(1j * rng.uniform(nTR, -np.pi, np.pi))).astype(np.complex64)
    seq = W.FispSeq(RF=RF, profiles=np.ones((1, nTR), np.complex64), TR=0.009, TE=0.004,
                    max_state=nTR + 2, TI=0.02, inversion=inversion)
    e = oracle.fisp(t, seq, "f64")
    for v in range(t.nvoxels):
        ref = fisp_ensemble(float(t.T1[v]), float(t.T2[v]), float(t.B1[v]), float(t.B0[v]),
                            RF.astype(np.complex128), f32(seq.TR), f32(seq.TE), f32(seq.TI), inversion=inversion)
        assert rel(e[:, v], ref) < 1e-10


def test_fisp_repetitions_match_ensemble(oracle):
    nTR, reps = 10, 3
    t = small_tissue(3, seed=5, b0=10.0)
    RF = (W.sinusoid_train(nTR, 50.0) * np.complex64(np.exp(0.3j))).astype(np.complex64)
    seq = W.FispSeq(RF=RF, profiles=np.ones((1, nTR), np.complex64), TR=0.01, TE=0.005,
                    max_state=reps * nTR + 2, TI=0.05, TW=0.2, repetitions=reps)
    e = oracle.fisp(t, seq, "f64")
    for v in range(t.nvoxels):
        T1, T2, B1, B0 = (float(getattr(t, f)[v]) for f in ("T1", "T2", "B1", "B0"))
        ref = fisp_ensemble(T1, T2, B1, B0, RF.astype(np.complex128), f32(seq.TR), f32(seq.TE), f32(seq.TI),
                            reps=reps, TW=f32(seq.TW))
        assert rel(e[:, v], ref) < 1e-10


def test_fisp_state_count_independent_above_ntr(oracle):
    t = small_tissue(16)
    base = W.make_fisp(20, 45.0, 21)
    e1 = oracle.fisp(t, base, "f64")
    base.max_state = 33
    e2 = oracle.fisp(t, base, "f64")
    assert rel(e1, e2) < 1e-14


def _random_train(nTR, seed):
    """Complex RF train (degrees) whose every flip is nonzero (10-80 degrees, random phase)."""
    rng = W.SplitMix64(seed)
    return (rng.uniform(nTR, 10, 80) * np.exp(1j * rng.uniform(nTR, -np.pi, np.pi))).astype(np.complex64)


def _truncation_seq(RF, N):
    return W.FispSeq(RF=RF, profiles=np.ones((1, len(RF)), np.complex64), TR=0.009, TE=0.004,
                     max_state=N, TI=0.02)


def epg_numpy(t, RF, N, TR, TE, TI, rule="reference"):
    """FP64 EPG over all voxels at once, one slice, inversion, one repetition.

    Excitation is the textbook EPG rotation matrix (Weigel, J. Magn. Reson.
    Imaging 41:266, 2015, eq. 15), not the oracle's element-wise form.  `rule`
    selects the state truncation applied at each dephasing step:
      "reference"  F-_k <- F-_{k+1}, F-_{N-1} <- 0, F+_k <- F+_{k-1} (F+_{N-1} lost),
                   F+_0 <- conj(F-_0)   (epg.cuh:229-298)
      "zero_from_N-2"  as reference but F- zeroed from index N-2 (an off-by-one)
    A run with N+1 states is the other off-by-one (one state too many kept).
    """
    T1, T2, B1, B0 = (getattr(t, f).astype(np.float64) for f in ("T1", "T2", "B1", "B0"))
    nv = T1.shape[0]
    Fp = np.zeros((N, nv), complex)
    Fm = np.zeros((N, nv), complex)
    Z = np.zeros((N, nv), complex)
    Z[0] = 1.0

    def relax(dt):
        E1, E2 = np.exp(-dt / T1), np.exp(-dt / T2)
        r = np.where(B0 == 0, 1.0, np.exp(1j * 2 * np.pi * dt * B0))
        Fp[:] *= E2 * r
        Fm[:] *= E2 * np.conj(r)
        Z[:] *= E1
        Z[0] += 1 - E1

    Z[:] *= -1  # inversion (ignores B1), then the TI interval
    E1 = np.exp(-TI / T1)
    Z[:] *= E1
    Z[0] += 1 - E1
    echo = np.zeros((len(RF), nv), complex)
    for n, rf in enumerate(RF.astype(np.complex128)):
        a = np.deg2rad(abs(rf)) * B1
        ph = np.angle(rf) if abs(rf) > 0 else 0.0
        c2, s2, sa, ca = np.cos(a / 2) ** 2, np.sin(a / 2) ** 2, np.sin(a), np.cos(a)
        e1, e2 = np.exp(1j * ph), np.exp(2j * ph)
        fp, fm, z = Fp.copy(), Fm.copy(), Z.copy()
        Fp[:] = c2 * fp + e2 * s2 * fm - 1j * e1 * sa * z
        Fm[:] = np.conj(e2) * s2 * fp + c2 * fm + 1j * np.conj(e1) * sa * z
        Z[:] = -0.5j * np.conj(e1) * sa * fp + 0.5j * e1 * sa * fm + ca * z
        relax(TE)
        echo[n] = Fp[0]
        relax(TR - TE)
        Fm[:-1] = Fm[1:].copy()
        Fm[N - 1] = 0
        if rule == "zero_from_N-2":
            Fm[max(N - 2, 0):] = 0
        Fp[1:] = Fp[:-1].copy()
        Fp[0] = np.conj(Fm[0])
    return echo


def truncation_horizon_holds(truncated, untruncated, N):
    """With every flip nonzero, the first echo a truncation to N states can
    change is echo 2N: F+ first reaches index N after TR N-1's shift (where the
    truncated run drops it); TR N's excitation moves it into F-_N, which the
    untruncated run shifts to F-_{N-1} (zeroed when truncated); F- then moves
    down one index per TR, reaching F-_0 and F+_0 = conj(F-_0) after TR 2N-1's
    shift, so TR 2N's echo is the first to differ.  Echoes 0..2N-1 are exact."""
    d = np.max(np.abs(truncated - untruncated), axis=1)
    return bool(np.all(d[: 2 * N] < 1e-13) and d[2 * N] > 1e-8)


def test_fisp_single_state_is_ideal_spoiling(oracle):
    """N = 1: the shift zeroes F-_0 and sets F+_0 = conj(F-_0) = 0 every TR, so the
    transverse state is perfectly spoiled and the echo follows the FLASH
    recursion echo_n = -i e^{i phi_n} sin(alpha_n) Z_n E2(TE) r(TE),
    Z_{n+1} = cos(alpha_n) Z_n E1(TR) + 1 - E1(TR), Z_0 = 1 - 2 E1(TI) (inversion) or 1."""
    nTR = 40
    t = small_tissue(24, seed=21, b0=40.0)
    RF = _random_train(nTR, 5)
    T1, T2, B1, B0 = (getattr(t, f).astype(np.float64) for f in ("T1", "T2", "B1", "B0"))
    for inversion in (True, False):
        seq = W.FispSeq(RF=RF, profiles=np.ones((1, nTR), np.complex64), TR=0.0087, TE=0.0041,
                        max_state=1, TI=0.025, inversion=inversion)
        e = oracle.fisp(t, seq, "f64")
        TR, TE, TI = f32(seq.TR), f32(seq.TE), f32(seq.TI)
        Z = 1 - 2 * np.exp(-TI / T1) if inversion else np.ones_like(T1)
        want = np.zeros_like(e)
        for n, rf in enumerate(RF.astype(np.complex128)):
            alpha = np.deg2rad(abs(rf)) * B1
            want[n] = (-1j * np.exp(1j * np.angle(rf)) * np.sin(alpha) * Z * np.exp(-TE / T2)
                       * np.exp(2j * np.pi * B0 * TE))
            Z = np.cos(alpha) * Z * np.exp(-TR / T1) + 1 - np.exp(-TR / T1)
        assert np.max(np.abs(e - want)) < 1e-13


@pytest.mark.parametrize("N", [2, 5, 8, 16])
def test_fisp_truncation_horizon(oracle, N):
    """Truncation to N < nTR states (epg.cuh:229-298) leaves echoes 0..2N-1
    exactly equal to the untruncated run (N > nTR, pinned by the isochromat
    ensemble) and changes echo 2N.  Both off-by-one variants of the rule fail
    this pin: zeroing F- from index N-2 changes echo 2N-2, and keeping N+1
    states leaves echo 2N unchanged (shown on the numpy EPG below)."""
    nTR = 2 * N + 6
    t = small_tissue(16, seed=3 + N, b0=25.0)
    RF = _random_train(nTR, 100 + N)
    full = oracle.fisp(t, _truncation_seq(RF, nTR + 2), "f64")
    trunc = oracle.fisp(t, _truncation_seq(RF, N), "f64")
    assert truncation_horizon_holds(trunc, full, N)


@pytest.mark.parametrize("N", [2, 5, 8])
def test_fisp_truncation_pin_rejects_off_by_one(oracle, N):
    """The horizon pin accepts the reference rule and rejects both off-by-one
    truncations (mistakes planted in a CPU numpy EPG, never in product code)."""
    nTR = 2 * N + 6
    t = small_tissue(16, seed=3 + N, b0=25.0)
    RF = _random_train(nTR, 100 + N)
    TR, TE, TI = f32(0.009), f32(0.004), f32(0.02)
    full = epg_numpy(t, RF, nTR + 2, TR, TE, TI)
    ref_rule = epg_numpy(t, RF, N, TR, TE, TI)
    # the numpy EPG with the reference rule is the oracle's truncated run
    assert np.max(np.abs(ref_rule - oracle.fisp(t, _truncation_seq(RF, N), "f64"))) < 1e-12
    assert truncation_horizon_holds(ref_rule, full, N)
    assert not truncation_horizon_holds(epg_numpy(t, RF, N, TR, TE, TI, "zero_from_N-2"), full, N)
    assert not truncation_horizon_holds(epg_numpy(t, RF, N + 1, TR, TE, TI), full, N)


def test_fisp_slice_profiles_sum(oracle):
    t = small_tissue(16)
    seq = W.make_fisp(30, 60.0, 12)
    prof = np.stack([np.full(30, 0.5 + 0.1j), np.linspace(0.2, 1.1, 30)]).astype(np.complex64)
    seq2 = W.FispSeq(RF=seq.RF, profiles=prof, TR=seq.TR, TE=seq.TE, max_state=12, TI=seq.TI)
    both = oracle.fisp(t, seq2, "f64")
    parts = 0
    for j in range(2):
        sj = W.FispSeq(RF=seq.RF, profiles=prof[j:j + 1], TR=seq.TR, TE=seq.TE, max_state=12, TI=seq.TI)
        parts = parts + oracle.fisp(t, sj, "f64")
    assert rel(both, parts) < 1e-14
    # a profile value scales the flip angle (and rotates the phase) of the RF pulse
    sc = W.FispSeq(RF=(seq.RF * prof[0]).astype(np.complex64), profiles=np.ones((1, 30), np.complex64),
                   TR=seq.TR, TE=seq.TE, max_state=12, TI=seq.TI)
    one = W.FispSeq(RF=seq.RF, profiles=prof[0:1], TR=seq.TR, TE=seq.TE, max_state=12, TI=seq.TI)
    assert rel(oracle.fisp(t, one, "f64"), oracle.fisp(t, sc, "f64")) < 1e-6


def test_fisp_undersampling_expands_echoes(oracle):
    t = small_tissue(9)
    seq = W.make_fisp(16, 40.0, 8)
    e1 = oracle.fisp(t, seq, "f64")
    seq.undersampling = 3
    e3 = oracle.fisp(t, seq, "f64")
    assert e3.shape == (48, 9)
    assert np.array_equal(e3, np.repeat(e1, 3, axis=0))


def test_fisp_f32_port_close_to_f64(oracle):
    wl = W.config(1, scale=4)
    a = oracle.fisp(wl.tissue, wl.sequence, "f64")
    b = oracle.fisp(wl.tissue, wl.sequence, "f32")
    assert rel(b, a) < 1e-5


# --------------------------------------------------------------- pSSFP ---

def test_pssfp_alternating_closed_form(oracle):
    # nz = 1, z = 0, B0 = 0, T1 = T2 = inf, real RF of constant |alpha| with alternating sign and a
    # real positive waveform summing to pi/180 per degree: every sub-rotation is about x, so after
    # the -alpha/2 preparation the magnetisation alternates between angles +-alpha*B1/2 from z, with
    # length |cos(pi*B1)| left by the inversion pulse:  |echo_n| = |cos(pi B1) sin(alpha B1 / 2)|
    nv, nTR, alpha = 5, 12, 40.0
    t = small_tissue(nv, b0=None)
    t = t.replace(T1=np.full(nv, np.inf, np.float32), T2=np.full(nv, np.inf, np.float32),
                  B0=np.zeros(nv, np.float32))
    seq = W.make_pssfp(nTR, 1.0, nz=1, nRF=7)
    seq.RF = (alpha * np.where(np.arange(nTR) % 2 == 0, 1.0, -1.0)).astype(np.complex64)
    seq.z = np.zeros(1, np.float32)
    g = np.abs(seq.gamma_dt_RF.real).astype(np.float64)
    seq.gamma_dt_RF = (g / g.sum() * np.pi / 180).astype(np.complex64)
    e = oracle.pssfp(t, seq, "f64")
    B1 = t.B1.astype(np.float64)
    want = np.abs(np.cos(np.pi * B1) * np.sin(np.deg2rad(alpha) * B1 / 2))
    assert np.max(np.abs(np.abs(e) - want[None, :])) < 1e-6
    # the transverse phase alternates sign with the RF sign (rotation about x only moves m in y-z)
    assert np.max(np.abs(e.real)) < 1e-12


def test_pssfp_norm_preserved_without_relaxation(oracle):
    nv = 4
    t = small_tissue(nv, b0=30.0)
    t = t.replace(T1=np.full(nv, np.inf, np.float32), T2=np.full(nv, np.inf, np.float32))
    seq = W.make_pssfp(15, 50.0, nz=1, nRF=9)
    seq.z = np.array([0.3], np.float32)
    e = oracle.pssfp(t, seq, "f64")
    # |m_xy| <= |m| = |cos(pi B1)| for every echo; equality would need m_z = 0
    B1 = t.B1.astype(np.float64)
    assert np.all(np.abs(e) <= np.abs(np.cos(np.pi * B1))[None, :] + 1e-12)


def test_pssfp_matches_rotvec_ensemble(oracle):
    nv = 3
    t = small_tissue(nv, seed=3, b0=40.0)
    seq = W.make_pssfp(10, 60.0, nz=4, nRF=5)
    e = oracle.pssfp(t, seq, "f64")
    for v in range(nv):
        ref = pssfp_isochromat(float(t.T1[v]), float(t.T2[v]), float(t.B1[v]), float(t.B0[v]),
                               seq.RF.astype(np.complex128), seq.gamma_dt_RF.astype(np.complex128),
                               tuple(float(np.float32(d)) for d in seq.dt),
                               tuple(float(np.float32(g)) for g in seq.gamma_dt_GRz),
                               seq.z.astype(np.float64))
        assert rel(e[:, v], ref) < 1e-10


def test_pssfp_f32_port_close_to_f64(oracle):
    # The FP32 reference algorithm itself drifts ~3.5e-5 from FP64 (even for one isochromat:
    # 25 Rodrigues steps per TR accumulate rounding with little T2 damping), so the port is
    # held to 1e-4 here; the GPU recursion keeps its state in FP64 for that reason.
    wl = W.config(3, scale=8)
    a = oracle.pssfp(wl.tissue, wl.sequence, "f64")
    b = oracle.pssfp(wl.tissue, wl.sequence, "f32")
    assert rel(b, a) < 1e-4


# -------------------------------------------------------------- signal ---

def test_cartesian_signal_is_centred_dft(oracle):
    # R2 = 0, B0 = 0, m = 1, C = 1, dk*dx = 2pi/N: S[r, s] = sum_x rho_y-summed(x) e^{i (s - N/2) dk x}
    n, fov = 16, 16.0
    t = W.make_tissue(W.SplitMix64(1), n, fov, None)
    t = t.replace(T2=np.full(n * n, np.inf, np.float32))
    traj = W.make_cartesian(n, 1, n, fov, 1e-5)
    echos = np.ones((traj.nreadouts, n * n), np.complex64)
    coils = np.ones((1, n * n), np.complex64)
    S = oracle.signal_cartesian(t, traj, coils, echos, "f64")
    rho_img = t.rho.astype(np.complex128).reshape(n, n)  # [iy][ix]
    line = rho_img.sum(axis=0)  # sum over y -> function of ix
    # sum_ix line[ix] exp(2 pi i (s - n/2)(ix - n/2)/n) via an inverse FFT with centring shifts
    want = np.fft.fftshift(np.fft.ifft(np.fft.ifftshift(line))) * n
    for r in range(traj.nreadouts):
        # float32 dk and x carry ~1.2e-7 relative error, on phases up to ~25 rad
        assert rel(S[0, r], want) < 1e-5


def test_cartesian_signal_is_matmul_of_factors(oracle):
    wl = W.config(2, scale=16)
    t, traj, coils = wl.tissue, wl.trajectory, wl.coils
    echos = oracle.fisp(t, wl.sequence, "f64")
    S = oracle.signal_cartesian(t, traj, coils, echos, "f64")
    ns, dt = traj.samples_per_readout, traj.delta_t
    T2, B0, x, y = (getattr(t, f).astype(np.float64) for f in ("T2", "B0", "x", "y"))
    dk = np.float64(np.float32(traj.delta_k))
    dt = np.float64(np.float32(dt))
    R2 = 1 / T2
    fac = np.exp(0.5 * ns * dt * R2 - 0.5j * ns * (dt * 2 * np.pi * B0 + dk * x)) * t.rho
    z = -dt * R2 + 1j * (dk * x + 2 * np.pi * dt * B0)
    B = np.exp(np.arange(ns)[:, None] * z[None, :])  # [s][v]
    for c in range(coils.shape[0]):
        want = (echos * fac[None, :]) @ (B * coils[c][None, :]).T
        assert rel(S[c], want) < 1e-11


def test_spiral_equals_cartesian_on_cartesian_grid(oracle):
    wl = W.config(2, scale=16)
    t, traj, coils = wl.tissue, wl.trajectory, wl.coils
    echos = oracle.fisp(t, wl.sequence, "f64")
    nr, ns = traj.nreadouts, traj.samples_per_readout
    dk = np.float32(traj.delta_k)
    sp = W.SpiralTraj(nr, ns, traj.delta_t,
                      np.full(nr, -ns / 2 * dk, np.complex64), np.full(nr, dk, np.complex64))
    a = oracle.signal_cartesian(t, traj, coils, echos, "f64")
    b = oracle.signal_spiral(t, sp, coils, echos, "f64")
    assert rel(b, a) < 1e-11


def test_phase_encoding_is_exact_elementwise(oracle):
    wl = W.config(1, scale=8)
    t, traj = wl.tissue, wl.trajectory
    echos = oracle.fisp(t, wl.sequence, "f64")
    pe = oracle.phase_encoding(t, traj, echos, "f64")
    want = echos * np.exp(1j * t.y[None, :].astype(np.float64) * traj.k_start.imag[:, None].astype(np.float64))
    assert rel(pe, want) < 1e-14


def test_signal_f32_port_close_to_f64(oracle):
    wl = W.config(1, scale=2)
    e = oracle.fisp(wl.tissue, wl.sequence, "f64")
    a = oracle.signal_cartesian(wl.tissue, wl.trajectory, wl.coils, e, "f64")
    b = oracle.signal_cartesian(wl.tissue, wl.trajectory, wl.coils, e.astype(np.complex64), "f32")
    assert rel(b, a) < 1e-4


# ------------------------------------------------------------ jacobian ---

def _jac_setup(scale=16, nc=None):
    wl = W.config(2, scale=scale)
    coils = wl.coils if nc is None else wl.coils[:nc]
    # round T1/T2 to 12 significant bits so that T * (1 +- 2**-11) is exact in float32:
    # finite-difference steps below are then exact and the FD error is O(h^2) only
    t = wl.tissue
    for f in ("T1", "T2"):
        val = getattr(t, f).astype(np.float64)
        m, e = np.frexp(val)
        t = t.replace(**{f: np.ldexp(np.round(m * 4096) / 4096, e).astype(np.float32)})
    wl.tissue = t
    return wl, coils


H = 2.0 ** -11


def _perturbed(t, name, direction):
    base = getattr(t, name).astype(np.float64)
    out = (base * (1 + H * direction)).astype(np.float32)
    assert np.array_equal(out.astype(np.float64), base * (1 + H * direction))  # exact steps
    return t.replace(**{name: out})


def _fd_derivs(oracle, wl):
    t, seq = wl.tissue, wl.sequence
    out = []
    for f in ("T1", "T2"):
        one = np.ones(t.nvoxels)
        mp = oracle.fisp(_perturbed(t, f, one), seq, "f64")
        mm = oracle.fisp(_perturbed(t, f, -one), seq, "f64")
        out.append((mp - mm) / (2 * H * getattr(t, f).astype(np.float64))[None, :])
    return out


def test_jacobian_rho_columns_equal_signal(oracle):
    wl, coils = _jac_setup()
    t, traj = wl.tissue, wl.trajectory
    nv = t.nvoxels
    echos = oracle.fisp(t, wl.sequence, "f64")
    d1 = d2 = np.zeros_like(echos)
    ones = t.replace(rho=np.ones(nv, np.complex64))
    S1 = oracle.signal_cartesian(ones, traj, coils, echos, "f64")
    v = np.zeros((4, nv), complex)
    v[2] = 1
    assert rel(oracle.jacobian(t, traj, coils, echos, d1, d2, v, "f64"), S1) < 1e-11
    v[2], v[3] = 0, 1
    assert rel(oracle.jacobian(t, traj, coils, echos, d1, d2, v, "f64"), 1j * S1) < 1e-11


def test_jacobian_adjoint_identity(oracle):
    wl, coils = _jac_setup()
    t, traj = wl.tissue, wl.trajectory
    nv = t.nvoxels
    rng = W.SplitMix64(1234)
    echos = oracle.fisp(t, wl.sequence, "f64")
    d1, d2 = _fd_derivs(oracle, wl)
    v = (rng.normal(4 * nv) + 1j * rng.normal(4 * nv)).reshape(4, nv)
    nr, ns = traj.nreadouts, traj.samples_per_readout
    yv = (rng.normal(coils.shape[0] * nr * ns) + 1j * rng.normal(coils.shape[0] * nr * ns)).reshape(coils.shape[0], nr, ns)
    Jv = oracle.jacobian(t, traj, coils, echos, d1, d2, v, "f64")
    JHy = oracle.jacobian_hermitian(t, traj, coils, echos, d1, d2, yv, "f64")
    lhs = np.vdot(yv.ravel(), Jv.ravel())
    rhs = np.vdot(JHy.ravel(), v.ravel())
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


@pytest.mark.parametrize("field", [0, 1])
def test_jacobian_T_rows_equal_finite_difference(oracle, field):
    # J.v with v = random signs on the T1 (T2) row is the derivative of the full pipeline
    # S(FISP(p)) along dT = T * v (the reference's log-parameter scaling)
    wl, coils = _jac_setup()
    t, traj = wl.tissue, wl.trajectory
    nv = t.nvoxels
    echos = oracle.fisp(t, wl.sequence, "f64")
    d1, d2 = _fd_derivs(oracle, wl)
    sign = np.where(W.SplitMix64(77).uniform(nv) < 0.5, -1.0, 1.0)
    v = np.zeros((4, nv), complex)
    v[field] = sign
    Jv = oracle.jacobian(t, traj, coils, echos, d1, d2, v, "f64")
    name = ("T1", "T2")[field]
    sig = []
    for sgn in (1, -1):
        tp = _perturbed(t, name, sgn * sign)
        sig.append(oracle.signal_cartesian(tp, traj, coils, oracle.fisp(tp, wl.sequence, "f64"), "f64"))
    fd = (sig[0] - sig[1]) / (2 * H)
    assert rel(fd, Jv) < 1e-5


# ---------------------------------------------- spiral jacobian, 6 params ---

SPIRAL_FIELDS = ("T1", "T2", "B1", "B0")
B0_STEP = 2.0 ** -16  # Hz; exact in float32 for |B0| < 64 Hz


def _spiral_setup():
    wl, coils = _jac_setup()
    n = int(round(np.sqrt(wl.tissue.nvoxels)))
    traj = W.make_spiral(wl.trajectory.nreadouts, 32, wl.fov, n, 5e-6)
    return wl, coils, traj


def _step(t, name, direction, sgn):
    """Tissue with field `name` moved by sgn * direction along the parameter's
    scaling (T1, T2: relative by H; B1: absolute H; B0: absolute B0_STEP), and
    the step per voxel in the parameter's own units."""
    base = getattr(t, name).astype(np.float64)
    if name in ("T1", "T2"):
        return _perturbed(t, name, sgn * direction), H * base
    h = H if name == "B1" else B0_STEP
    out = (base + sgn * h * direction).astype(np.float32)
    assert np.array_equal(out.astype(np.float64), base + sgn * h * direction)  # exact steps
    return t.replace(**{name: out}), np.full(t.nvoxels, h)


def _spiral_derivs(oracle, wl):
    t, seq = wl.tissue, wl.sequence
    one = np.ones(t.nvoxels)
    out = []
    for f in SPIRAL_FIELDS:
        tp, h = _step(t, f, one, 1)
        tm, _ = _step(t, f, one, -1)
        out.append((oracle.fisp(tp, seq, "f64") - oracle.fisp(tm, seq, "f64")) / (2 * h)[None, :])
    return out


def test_spiral_jacobian_rho_columns_equal_signal(oracle):
    wl, coils, traj = _spiral_setup()
    t = wl.tissue
    nv = t.nvoxels
    echos = oracle.fisp(t, wl.sequence, "f64")
    zero = np.zeros_like(echos)
    S1 = oracle.signal_spiral(t.replace(rho=np.ones(nv, np.complex64)), traj, coils, echos, "f64")
    v = np.zeros((6, nv), complex)
    v[4] = 1
    got = oracle.jacobian_spiral(t, traj, coils, echos, zero, zero, zero, zero, v, "f64")
    assert rel(got, S1) < 1e-11
    v[4], v[5] = 0, 1
    got = oracle.jacobian_spiral(t, traj, coils, echos, zero, zero, zero, zero, v, "f64")
    assert rel(got, 1j * S1) < 1e-11


def test_spiral_jacobian_adjoint_identity(oracle):
    wl, coils, traj = _spiral_setup()
    t = wl.tissue
    nv = t.nvoxels
    rng = W.SplitMix64(4321)
    echos = oracle.fisp(t, wl.sequence, "f64")
    ds = _spiral_derivs(oracle, wl)
    v = (rng.normal(6 * nv) + 1j * rng.normal(6 * nv)).reshape(6, nv)
    nc, nr, ns = coils.shape[0], traj.nreadouts, traj.samples_per_readout
    yv = (rng.normal(nc * nr * ns) + 1j * rng.normal(nc * nr * ns)).reshape(nc, nr, ns)
    Jv = oracle.jacobian_spiral(t, traj, coils, echos, *ds, v, "f64")
    JHy = oracle.jacobian_hermitian_spiral(t, traj, coils, echos, *ds, yv, "f64")
    lhs = np.vdot(yv.ravel(), Jv.ravel())
    rhs = np.vdot(JHy.ravel(), v.ravel())
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


@pytest.mark.parametrize("field", [0, 1, 2, 3])
def test_spiral_jacobian_rows_equal_finite_difference(oracle, field):
    # each J.v row is the derivative of the whole pipeline S_spiral(FISP(p), p) along
    # random signs (T1, T2 scaled by T as in the reference's Cartesian J.v)
    wl, coils, traj = _spiral_setup()
    t = wl.tissue
    nv = t.nvoxels
    echos = oracle.fisp(t, wl.sequence, "f64")
    ds = _spiral_derivs(oracle, wl)
    sign = np.where(W.SplitMix64(99 + field).uniform(nv) < 0.5, -1.0, 1.0)
    v = np.zeros((6, nv), complex)
    v[field] = sign
    Jv = oracle.jacobian_spiral(t, traj, coils, echos, *ds, v, "f64")
    name = SPIRAL_FIELDS[field]
    sig = []
    for sgn in (1, -1):
        tp, _ = _step(t, name, sign, sgn)
        sig.append(oracle.signal_spiral(tp, traj, coils, oracle.fisp(tp, wl.sequence, "f64"), "f64"))
    h = H if name != "B0" else B0_STEP
    fd = (sig[0] - sig[1]) / (2 * h)
    assert rel(fd, Jv) < 1e-5


def test_spiral_jacobian_f32_port_close_to_f64(oracle):
    wl, coils, traj = _spiral_setup()
    t = wl.tissue
    nv = t.nvoxels
    rng = W.SplitMix64(11)
    echos = oracle.fisp(t, wl.sequence, "f64")
    ds = _spiral_derivs(oracle, wl)
    v = (rng.normal(6 * nv) + 1j * rng.normal(6 * nv)).reshape(6, nv)
    a = oracle.jacobian_spiral(t, traj, coils, echos, *ds, v, "f64")
    b = oracle.jacobian_spiral(t, traj, coils, echos, *ds, v, "f32")
    assert rel(b, a) < 1e-4
