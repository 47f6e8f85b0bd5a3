"""Host logic of the Gauss-Newton loop (compas_amd.reconstruct) without a GPU:
the two conjugate-gradient solvers on small dense systems whose solutions
numpy's direct solver gives, and the multiplicative/additive update rule."""

import numpy as np
import pytest

from compas_amd import reconstruct as R

torch = pytest.importorskip("torch")


def spd(n, cond, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (q * np.geomspace(1.0, cond, n)) @ q.T


def test_host_cg_solves_spd_system():
    a = spd(24, 1e3, 1)
    b = np.random.default_rng(2).standard_normal((4, 6))
    x, k = R._cg(lambda v: (a @ v.reshape(-1)).reshape(v.shape), b, 200, 1e-12)
    assert np.allclose(x.reshape(-1), np.linalg.solve(a, b.reshape(-1)), rtol=1e-8, atol=1e-10)
    assert k <= 200


@pytest.mark.parametrize("shift_rel", [0.0, 0.05])
def test_device_cg_matches_host_cg(shift_rel):
    # same recurrences on torch tensors (here on the CPU): the same solution
    # of (A + shift I) x = b, with the shift taken from the first step
    a = spd(30, 1e4, 3)
    b = np.random.default_rng(4).standard_normal((5, 6))
    host, kh = R._cg(lambda v: (a @ v.reshape(-1)).reshape(v.shape), b, 300, 1e-11, shift_rel)
    at = torch.from_numpy(a)
    dev, kd = R._cg_device(lambda v: (at @ v.reshape(-1)).reshape(v.shape), torch.from_numpy(b), 300, 1e-11,
                           shift_rel, torch)
    assert kd >= kh and kd - kh < 8  # the device test runs every 8 steps
    shift = shift_rel * float(b.reshape(-1) @ a @ b.reshape(-1)) / float(b.reshape(-1) @ b.reshape(-1))
    want = np.linalg.solve(a + shift * np.eye(30), b.reshape(-1)).reshape(b.shape)
    assert np.allclose(dev.numpy(), want, rtol=1e-7, atol=1e-9)
    assert np.allclose(host, want, rtol=1e-7, atol=1e-9)


def test_device_cg_stops_at_nonpositive_curvature():
    # an indefinite operator: host CG stops at the first p.Ap <= 0 and keeps
    # the iterate it had; the device CG latches the stop and returns the same
    # iterate at its next convergence test
    d = np.array([3.0, 2.0, 1.0, -1.0, -2.0])
    b = np.ones((1, 5))
    host, kh = R._cg(lambda v: v * d, b, 50, 1e-14)
    dt = torch.from_numpy(d)
    dev, kd = R._cg_device(lambda v: v * dt, torch.from_numpy(b), 50, 1e-14, 0.0, torch)
    assert kh < 5 and kd == 8
    assert np.allclose(dev.numpy(), host, rtol=1e-12, atol=1e-14)


def test_zero_right_hand_side():
    b = np.zeros((2, 3))
    x, k = R._cg(lambda v: v, b, 10, 1e-6)
    assert k == 0 and not x.any()
    xt, kt = R._cg_device(lambda v: v, torch.zeros(2, 3, dtype=torch.float64), 10, 1e-6, 0.0, torch)
    assert kt == 0 and not xt.any()


def test_update_rule():
    # T1/T2 rows of J are scaled by T (d S / d log T), so a step delta moves
    # log T by delta (clipped at max_log_step); B1, B0, rho move additively
    n = 3
    est = R.Estimate(np.full(n, 800.0, np.float32), np.full(n, 60.0, np.float32), np.ones(n, np.float32),
                     np.zeros(n, np.float32), np.full(n, 1 + 0.5j, np.complex64), np.zeros(n, np.float32),
                     np.zeros(n, np.float32), np.zeros(n, np.float32))
    delta = np.array([[0.1, -2.0, 0.0], [0.0, 0.2, 5.0], [0.01, 0.0, 0.0], [2.0, 0.0, 0.0],
                      [0.5, 0.0, 0.0], [0.0, -0.25, 0.0]])
    out = R._apply_step(est, delta, 0.5, True, 0.7)
    assert np.allclose(out.T1, 800.0 * np.exp([0.05, -0.7, 0.0]), rtol=1e-6)
    assert np.allclose(out.T2, 60.0 * np.exp([0.0, 0.1, 0.7]), rtol=1e-6)
    assert np.allclose(out.B1, [1.005, 1.0, 1.0]) and np.allclose(out.B0, [1.0, 0.0, 0.0])
    assert np.allclose(out.rho, [1.25 + 0.5j, 1 + 0.375j, 1 + 0.5j])
    cart = R._apply_step(est, delta[[0, 1, 4, 5]], 1.0, False, 0.7)  # rows T1 T2 rho_x rho_y
    assert np.allclose(cart.rho, [1.5 + 0.5j, 1 + 0.25j, 1 + 0.5j]) and np.array_equal(cart.B1, est.B1)
