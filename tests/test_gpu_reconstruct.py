"""Gauss-Newton reconstruction on the HIP path: noise-free k-space data of a
known phantom, a perturbed start, and convergence back to the phantom.

The loop (compas_amd.reconstruct) composes the products the parity tests pin
(tangent derivatives, signal, residual, J.v, J^H.y); this checks that they
compose into a descent method whose fixed point is the truth: the objective
falls by orders of magnitude and T1, T2, rho return to the phantom's values.
"""

import numpy as np
import pytest

import compas_amd as ca
from compas_amd import reconstruct as R
from compas_amd import workloads as W
from gpu_helpers import banner, context

pytestmark = pytest.mark.gpu


def phantom(n=8, fov=8.0, seed=11):
    rng = W.SplitMix64(seed)
    t = W.make_tissue(rng, n, fov, 20.0)
    return t


def estimate_of(t, **kw):
    e = R.Estimate(t.T1.copy(), t.T2.copy(), t.B1.copy(), t.B0.copy(), t.rho.copy(), t.x, t.y, t.z)
    for k, v in kw.items():
        setattr(e, k, v)
    return e


def test_gauss_newton_cartesian_recovers_phantom():
    banner("gauss-newton cartesian 8x8")
    context()
    n, fov = 8, 8.0
    t = phantom(n, fov)
    seq = ca.FispSequence.from_sequence(W.make_fisp(96, 60.0, 12))
    tr = W.make_cartesian(n, 12, n, fov, 1e-5)
    traj = ca.CartesianTrajectory.from_trajectory(tr)
    coils = W.make_coils(n, fov, 2)
    truth = estimate_of(t)
    params = truth.parameters()
    m = ca.phase_encoding(ca.simulate_magnetization(params, seq), params, traj)
    y = ca.magnetization_to_signal(m, params, traj, coils, method="direct").collect()

    start = estimate_of(t, T1=(t.T1 * 1.25).astype(np.float32), T2=(t.T2 * 0.8).astype(np.float32),
                        rho=(t.rho * 0.9).astype(np.complex64))
    res = R.gauss_newton(y, start, seq, traj, coils, iterations=10, cg_iterations=200, cg_tol=1e-7,
                         method="direct", signal_method="direct")
    obj = res.objective
    print("objective", obj, "cg", res.cg_iterations, "steps", res.steps, flush=True)
    assert obj[-1] < 1e-8 * obj[0], (obj, res.cg_iterations, res.steps)
    assert all(b <= a for a, b in zip(obj, obj[1:])), obj
    est = res.estimate
    assert np.max(np.abs(est.T1 / t.T1 - 1)) < 1e-3
    assert np.max(np.abs(est.T2 / t.T2 - 1)) < 1e-3
    assert np.max(np.abs(est.rho - t.rho)) < 1e-3 * np.max(np.abs(t.rho))


def test_gauss_newton_spiral_with_b1():
    banner("gauss-newton spiral 8x8 with B1")
    context()
    n, fov = 8, 8.0
    t = phantom(n, fov, seed=12)
    seq = ca.FispSequence.from_sequence(W.make_fisp(80, 60.0, 12))
    tr = W.make_spiral(80, 32, fov, n, 5e-6)
    traj = ca.SpiralTrajectory.from_trajectory(tr)
    coils = W.make_coils(n, fov, 2)
    truth = estimate_of(t)
    params = truth.parameters()
    y = ca.magnetization_to_signal(ca.simulate_magnetization(params, seq), params, traj, coils,
                                   method="direct").collect()

    start = estimate_of(t, T1=(t.T1 * 0.85).astype(np.float32), T2=(t.T2 * 1.15).astype(np.float32),
                        B1=(t.B1 + 0.05).astype(np.float32), rho=(t.rho * 1.1).astype(np.complex64))
    # with B1 free the normal equations are ill-conditioned and CG stops at its
    # cap, so the objective falls linearly (about 6x per iteration at the end):
    # 12 iterations bring every voxel's T1, T2 and B1 within the bounds below
    res = R.gauss_newton(y, start, seq, traj, coils, iterations=12, cg_iterations=200, cg_tol=1e-7, fit_b1=True,
                         method="direct", signal_method="direct")
    obj = res.objective
    print("objective", obj, "cg", res.cg_iterations, "steps", res.steps, flush=True)
    assert obj[-1] < 1e-6 * obj[0], (obj, res.cg_iterations, res.steps)
    est = res.estimate
    assert np.max(np.abs(est.T1 / t.T1 - 1)) < 5e-3
    assert np.max(np.abs(est.T2 / t.T2 - 1)) < 5e-3
    assert np.max(np.abs(est.B1 - t.B1)) < 5e-3


def test_gauss_newton_mixed_precision_descends():
    # the default mixed-precision products (BF16 hi + lo on MFMA) drive the same
    # loop: the objective still falls by orders of magnitude
    banner("gauss-newton spiral mixed precision")
    context()
    n, fov = 8, 8.0
    t = phantom(n, fov, seed=13)
    seq = ca.FispSequence.from_sequence(W.make_fisp(80, 60.0, 12))
    traj = ca.SpiralTrajectory.from_trajectory(W.make_spiral(80, 32, fov, n, 5e-6))
    coils = W.make_coils(n, fov, 2)
    params = estimate_of(t).parameters()
    y = ca.magnetization_to_signal(ca.simulate_magnetization(params, seq), params, traj, coils).collect()
    start = estimate_of(t, T1=(t.T1 * 0.9).astype(np.float32), T2=(t.T2 * 1.1).astype(np.float32))
    res = R.gauss_newton(y, start, seq, traj, coils, iterations=6)
    assert res.objective[-1] < 1e-4 * res.objective[0], (res.objective, res.cg_iterations)
    assert np.max(np.abs(res.estimate.T2 / t.T2 - 1)) < 1e-2


def test_gauss_newton_graph_replay_is_identical():
    # the captured normal operator replays the same kernels: the whole fit is
    # bitwise the same with and without graphs
    banner("gauss-newton with and without graph replay")
    context()
    n, fov = 8, 8.0
    t = phantom(n, fov, seed=14)
    seq = ca.FispSequence.from_sequence(W.make_fisp(64, 60.0, 10))
    traj = ca.CartesianTrajectory.from_trajectory(W.make_cartesian(n, 8, n, fov, 1e-5))
    coils = W.make_coils(n, fov, 2)
    params = estimate_of(t).parameters()
    m = ca.phase_encoding(ca.simulate_magnetization(params, seq), params, traj)
    y = ca.magnetization_to_signal(m, params, traj, coils).collect()
    start = estimate_of(t, T1=(t.T1 * 1.2).astype(np.float32), T2=(t.T2 * 0.85).astype(np.float32))
    a = R.gauss_newton(y, start, seq, traj, coils, iterations=4, use_graph=True, cg_on_device=False)
    b = R.gauss_newton(y, start, seq, traj, coils, iterations=4, use_graph=False)
    print("normal operator ms per call: graph %.3f, eager %.3f" % (
        1e3 * a.normal_seconds / a.normal_calls, 1e3 * b.normal_seconds / b.normal_calls), flush=True)
    assert a.objective == b.objective
    assert np.array_equal(a.estimate.T1, b.estimate.T1)
    assert np.array_equal(a.estimate.T2, b.estimate.T2)
    assert np.array_equal(a.estimate.rho, b.estimate.rho)



def test_gauss_newton_device_cg_matches_host_cg():
    # CG with its vectors on the device (torch views of the graph's arrays,
    # event-ordered streams) takes the same steps as CG on the host up to
    # rounding and its every-8-steps convergence test: the same fixed point
    banner("gauss-newton device CG vs host CG")
    context()
    torch = pytest.importorskip("torch")
    assert torch.cuda.is_available()
    n, fov = 8, 8.0
    t = phantom(n, fov, seed=15)
    seq = ca.FispSequence.from_sequence(W.make_fisp(96, 60.0, 12))
    traj = ca.CartesianTrajectory.from_trajectory(W.make_cartesian(n, 12, n, fov, 1e-5))
    coils = W.make_coils(n, fov, 2)
    params = estimate_of(t).parameters()
    m = ca.phase_encoding(ca.simulate_magnetization(params, seq), params, traj)
    y = ca.magnetization_to_signal(m, params, traj, coils, method="direct").collect()
    start = estimate_of(t, T1=(t.T1 * 1.2).astype(np.float32), T2=(t.T2 * 0.85).astype(np.float32),
                        rho=(t.rho * 0.9).astype(np.complex64))
    kw = dict(iterations=8, cg_iterations=200, cg_tol=1e-7, method="direct", signal_method="direct")
    host = R.gauss_newton(y, start, seq, traj, coils, cg_on_device=False, **kw)
    dev = R.gauss_newton(y, start, seq, traj, coils, cg_on_device=True, **kw)
    print("objective host", host.objective, "device", dev.objective, "cg", host.cg_iterations, dev.cg_iterations,
          flush=True)
    assert dev.cg_on_device and not host.cg_on_device
    assert dev.objective[-1] < 1e-8 * dev.objective[0], dev.objective
    # the first CG solve: same system, same start; the device run may take up
    # to 7 more steps before its convergence test, so its step is as close to
    # the solution as the host's or closer
    assert abs(dev.objective[1] - host.objective[1]) <= 1e-3 * host.objective[0], (host.objective, dev.objective)
    for f in ("T1", "T2"):
        a, b = getattr(dev.estimate, f), getattr(host.estimate, f)
        assert np.max(np.abs(a / b - 1)) < 1e-4, f
    assert np.max(np.abs(dev.estimate.rho - host.estimate.rho)) < 1e-4 * np.max(np.abs(t.rho))
