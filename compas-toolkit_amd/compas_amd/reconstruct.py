"""Gauss-Newton reconstruction of tissue parameters from k-space data.

The reference stops at the products a reconstruction needs and leaves the loop
to its Julia users: the objective ||S(p) - y||^2 and residual
(CompasToolkit.jl:1015-1036, src/utils/residual.cu:8-46), J.v
(CompasToolkit.jl compute_jacobian) and J^H.y (compute_jacobian_hermitian).
This module is that caller loop, run on the HIP path:

    per outer iteration
      m, dm/dp       simulate_magnetization + simulate_magnetization_tangents
                     (exact derivatives; pSSFP: all parameters in one pass)
      r = S(m) - y   magnetization_to_signal + compute_residual
      g = Re J^H r   compute_jacobian_hermitian
      (Re J^H J + lambda I) delta = -g    conjugate gradients, one J.v and one
                                          J^H.y per iteration
      p <- p (+) delta, halving the step while the objective does not decrease

Parameterisation follows the products: the T1 and T2 rows of J are scaled by
T1 and T2 (product_kernels.cuh:57-58), so their updates are relative
(T <- T exp(delta)); B1 and B0 (spiral only) and rho_x, rho_y are absolute.
Cartesian data use the four-parameter products (T1, T2, rho_x, rho_y); spiral
data the six-parameter ones (T1, T2, B1, B0, rho_x, rho_y), where B1 and B0
may be held fixed.  With phase_encode (Cartesian), echoes and derivatives pass
through phase_encoding before the signal, as the reference's examples do.

Every product runs on the GPU.  With use_graph (default) the normal operator
J^H(J v) is captured as a HIP graph after each linearisation (one eager
product first, for autotuning) and each CG step replays it with v written
into the graph's input array: the same kernels and results, without the
per-launch host overhead.  The CG vectors ([rows][nvoxels] real) then live
on the device as torch tensors that share memory with the graph's input and
output arrays (DLPack, no copies); torch's stream and the library's are
ordered by events, and the CG scalars stay on the device (convergence is
checked every few steps), so a CG step issues no host round trip.  Without
the graph, or without torch, the vectors live on the host (numpy).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

import time

from . import (B0, B1, T1, T2, CartesianTrajectory, CompasArray, SpiralTrajectory, TissueParameters,
               as_array, capture, compute_jacobian, compute_jacobian_hermitian, compute_jacobian_hermitian_spiral,
               compute_jacobian_spiral, compute_residual, magnetization_to_signal, phase_encoding,
               simulate_magnetization, simulate_magnetization_tangents)


@dataclass
class Estimate:
    """Tissue parameters of a reconstruction (float32 arrays of nvoxels)."""
    T1: np.ndarray
    T2: np.ndarray
    B1: np.ndarray
    B0: np.ndarray
    rho: np.ndarray  # complex64
    x: np.ndarray
    y: np.ndarray
    z: Optional[np.ndarray] = None

    def copy(self) -> "Estimate":
        return Estimate(*(None if a is None else np.array(a, copy=True) for a in
                          (self.T1, self.T2, self.B1, self.B0, self.rho, self.x, self.y, self.z)))

    def parameters(self) -> TissueParameters:
        return TissueParameters(self.T1.shape[0], self.T1, self.T2, self.B1, self.B0, self.rho.real,
                                self.rho.imag, self.x, self.y, self.z)


@dataclass
class GaussNewtonResult:
    estimate: Estimate
    objective: List[float] = field(default_factory=list)  # before each iteration, then the final value
    cg_iterations: List[int] = field(default_factory=list)
    steps: List[float] = field(default_factory=list)
    normal_seconds: float = 0.0  # wall time in the normal operator J^H(J v) (host view; the whole CG on the device)
    normal_calls: int = 0
    cg_on_device: bool = False


class _Model:
    """S(p), the derivatives at p and the products J.v / Re J^H.y."""

    def __init__(self, sequence, trajectory, coils, method, signal_method, phase_encode, fit_b1, fit_b0,
                 use_graph=True):
        self.sequence = sequence
        self.trajectory = trajectory
        self.coils = as_array(coils, np.complex64)
        self.method = method
        self.signal_method = signal_method
        self.spiral = isinstance(trajectory, SpiralTrajectory)
        if not self.spiral and not isinstance(trajectory, CartesianTrajectory):
            raise TypeError("trajectory must be Cartesian or spiral")
        self.phase_encode = phase_encode and not self.spiral
        self.fit_b1, self.fit_b0 = fit_b1, fit_b0
        if not self.spiral and (fit_b1 or fit_b0):
            raise ValueError("the Cartesian products have no B1 / B0 rows")
        self.rows = 6 if self.spiral else 4
        self.use_graph = use_graph
        self.graph = None
        self.v_t = self.out_t = self.mask_t = None  # torch views of the graph's arrays (device CG)
        self._views_of = None
        self.normal_seconds = 0.0
        self.normal_calls = 0

    def _encode(self, a, params):
        return phase_encoding(a, params, self.trajectory) if self.phase_encode else a

    def signal(self, est: Estimate):
        params = est.parameters()
        m = self._encode(simulate_magnetization(params, self.sequence), params)
        return params, m, magnetization_to_signal(m, params, self.trajectory, self.coils, method=self.signal_method)

    def linearise(self, est: Estimate, params, m):
        fields = (T1, T2, B1, B0) if self.spiral else (T1, T2)
        self.params = params
        self.m = m
        self.d = [self._encode(d, params) for d in simulate_magnetization_tangents(fields, params, self.sequence)]
        if self.graph is not None:
            self.graph.close()
            self.graph = None
        if self.use_graph:
            nv = est.T1.shape[0]
            self.v_in = CompasArray.from_numpy(np.zeros((self.rows, nv), np.complex64))
            self._jhy_dev(self.jv(self.v_in)).collect()  # eager once: autotuning outside the capture
            with capture() as g:
                self.normal_out = self._jhy_dev(self.jv(self.v_in))
            self.graph = g

    def jv(self, v) -> CompasArray:
        vc = v if isinstance(v, CompasArray) else as_array(np.ascontiguousarray(v, dtype=np.complex64))
        if self.spiral:
            return compute_jacobian_spiral(self.m, self.d, self.params, self.trajectory, self.coils, vc,
                                           method=self.method)
        return compute_jacobian(self.m, self.d, self.params, self.trajectory, self.coils, vc, method=self.method)

    def _jhy_dev(self, y) -> CompasArray:
        if self.spiral:
            return compute_jacobian_hermitian_spiral(self.m, self.d, self.params, self.trajectory, self.coils, y,
                                                     method=self.method)
        return compute_jacobian_hermitian(self.m, self.d, self.params, self.trajectory, self.coils, y,
                                          method=self.method)

    def jhy(self, y) -> np.ndarray:
        return self.mask(self._jhy_dev(y).collect().real.astype(np.float64))

    def mask(self, g: np.ndarray) -> np.ndarray:
        if self.spiral:
            if not self.fit_b1:
                g[2] = 0.0
            if not self.fit_b0:
                g[3] = 0.0
        return g

    def normal(self, v: np.ndarray) -> np.ndarray:
        t0 = time.perf_counter()
        if self.graph is not None:
            self.v_in.write(np.asarray(v, dtype=np.complex64))
            self.graph.launch()
            out = self.mask(self.normal_out.collect().real.astype(np.float64))
        else:
            out = self.jhy(self.jv(v))
        self.normal_seconds += time.perf_counter() - t0
        self.normal_calls += 1
        return out

    # ---- device-resident normal operator (torch views of the graph's arrays)

    def device_views(self, torch):
        """(input, output) torch views of the captured graph's arrays."""
        if self._views_of is not self.graph:
            self.v_t = self.v_in.to_torch()
            self.out_t = self.normal_out.to_torch()
            self._views_of = self.graph
            mask = np.ones((self.rows, 1))
            if self.spiral:
                mask[2] = 0.0 if not self.fit_b1 else 1.0
                mask[3] = 0.0 if not self.fit_b0 else 1.0
            self.mask_t = torch.tensor(mask, dtype=torch.float64, device=self.v_t.device)
        return self.v_t, self.out_t

    def normal_device(self, v, torch):
        """Re J^H J v (masked) for a device vector v: [rows][nv] float64."""
        v_t, out_t = self.device_views(torch)
        stream = torch.cuda.current_stream(v.device).cuda_stream
        v_t.copy_(v)  # real -> complex, on torch's stream
        self.v_in.context.wait_for_stream(stream)
        self.graph.launch()
        self.v_in.context.make_stream_wait(stream)
        self.normal_calls += 1
        return out_t.real.to(torch.float64) * self.mask_t

    def close(self):
        if self.graph is not None:
            self.graph.close()
            self.graph = None
        self.v_t = self.out_t = None
        self._views_of = None


def _cg(apply, b, iterations, tol, shift_rel=0.0):
    """Conjugate gradients for (A + shift I) x = b, A symmetric positive
    semi-definite; shift = shift_rel times the Rayleigh quotient of b."""
    x = np.zeros_like(b)
    r = b.copy()
    p = r.copy()
    rr = float(np.vdot(r, r).real)
    b_norm = np.sqrt(rr)
    if b_norm == 0.0:
        return x, 0
    shift = None
    k = 0
    for k in range(1, iterations + 1):
        Ap = apply(p)
        if shift is None:
            shift = shift_rel * float(np.vdot(p, Ap).real) / rr
        Ap = Ap + shift * p
        pAp = float(np.vdot(p, Ap).real)
        if pAp <= 0.0:
            break
        alpha = rr / pAp
        x += alpha * p
        r -= alpha * Ap
        rr_new = float(np.vdot(r, r).real)
        if np.sqrt(rr_new) <= tol * b_norm:
            break
        p = r + (rr_new / rr) * p
        rr = rr_new
    return x, k


def _cg_device(apply, b, iterations, tol, shift_rel, torch, check_every=8):
    """_cg with the vectors and scalars on the device: no host round trip per
    step; the residual test (and a non-positive curvature stop) every
    `check_every` steps."""
    x = torch.zeros_like(b)
    r = b.clone()
    p = r.clone()
    rr = torch.sum(r * r)
    b_norm2 = float(rr)
    if b_norm2 == 0.0:
        return x, 0
    shift = None
    live = torch.ones((), dtype=torch.bool, device=b.device)  # false from the first pAp <= 0 on
    k = 0
    for k in range(1, iterations + 1):
        Ap = apply(p)
        if shift is None:
            shift = shift_rel * torch.sum(p * Ap) / rr
        Ap = Ap + shift * p
        pAp = torch.sum(p * Ap)
        live = live & (pAp > 0)
        alpha = torch.where(live, rr / torch.where(live, pAp, torch.ones_like(pAp)), torch.zeros_like(pAp))
        x += alpha * p
        r -= alpha * Ap
        rr_new = torch.sum(r * r)
        if k % check_every == 0 or k == iterations:
            if not bool(live) or float(rr_new) <= tol * tol * b_norm2:
                break
        p = r + (rr_new / rr) * p
        rr = rr_new
    return x, k


def _torch_gpu():
    """torch with a visible GPU, or None."""
    try:
        import torch
    except ImportError:
        return None
    return torch if torch.cuda.is_available() else None


def _apply_step(est: Estimate, delta: np.ndarray, step: float, spiral: bool, max_log_step: float) -> Estimate:
    out = est.copy()
    d = step * delta
    out.T1 = (est.T1 * np.exp(np.clip(d[0], -max_log_step, max_log_step))).astype(np.float32)
    out.T2 = (est.T2 * np.exp(np.clip(d[1], -max_log_step, max_log_step))).astype(np.float32)
    if spiral:
        out.B1 = (est.B1 + d[2]).astype(np.float32)
        out.B0 = (est.B0 + d[3]).astype(np.float32)
        rx, ry = d[4], d[5]
    else:
        rx, ry = d[2], d[3]
    out.rho = (est.rho + (rx + 1j * ry)).astype(np.complex64)
    return out


def gauss_newton(measured, initial: Estimate, sequence, trajectory, coils, iterations: int = 8,
                 cg_iterations: int = 100, cg_tol: float = 1e-6, damping: float = 1e-6,
                 method: str = "gemm_fast", signal_method: str = "matmul_fast", phase_encode: bool = True,
                 fit_b1: bool = False, fit_b0: bool = False, max_log_step: float = 0.7,
                 min_step: float = 1.0 / 64, use_graph: bool = True,
                 cg_on_device: Optional[bool] = None) -> GaussNewtonResult:
    """Minimise ||S(p) - measured||^2 over the tissue parameters, starting from
    `initial`.  `measured` is [ncoils][nreadouts][samples] complex64 (host or
    device array).  `damping` is the Levenberg shift relative to the normal
    operator's Rayleigh quotient along the gradient.  cg_on_device (default:
    with use_graph, when torch sees the GPU) keeps the CG on the device."""
    torch = _torch_gpu() if use_graph and cg_on_device is not False else None
    if cg_on_device and torch is None:
        raise RuntimeError("cg_on_device needs use_graph and torch with a GPU")
    model = _Model(sequence, trajectory, coils, method, signal_method, phase_encode, fit_b1, fit_b0, use_graph)
    y = as_array(measured, np.complex64)
    est = initial.copy()
    result = GaussNewtonResult(est)
    params, m, s = model.signal(est)
    obj, r = compute_residual(s, y)
    for _ in range(iterations):
        result.objective.append(obj)
        model.linearise(est, params, m)
        g = model.jhy(r)
        t0 = time.perf_counter()
        if torch is not None:
            b = torch.from_numpy(-g).to(device=model.device_views(torch)[0].device)  # the library's GPU
            delta_t, k = _cg_device(lambda v: model.normal_device(v, torch), b, cg_iterations, cg_tol, damping, torch)
            delta = delta_t.cpu().numpy()
            model.normal_seconds += time.perf_counter() - t0
        else:
            delta, k = _cg(model.normal, -g, cg_iterations, cg_tol, damping)
        result.cg_iterations.append(k)
        step = 1.0
        while True:
            cand = _apply_step(est, delta, step, model.spiral, max_log_step)
            c_params, c_m, c_s = model.signal(cand)
            c_obj, c_r = compute_residual(c_s, y)
            if c_obj < obj or step <= min_step:
                break
            step *= 0.5
        if not c_obj < obj:  # no decrease along the Gauss-Newton direction: converged
            result.steps.append(0.0)
            break
        result.steps.append(step)
        est, params, m, obj, r = cand, c_params, c_m, c_obj, c_r
    model.close()
    result.estimate = est
    result.objective.append(obj)
    result.normal_seconds = model.normal_seconds
    result.normal_calls = model.normal_calls
    result.cg_on_device = torch is not None
    return result
