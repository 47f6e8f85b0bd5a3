"""ctypes front end of the oracle library (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
benchmark's CPU-baseline leg, never by the product package.

precision="f64" runs the FP64 oracle; precision="f32" runs the FP32 port of
the reference algorithm (the timed CPU baseline).  Inputs use the dataclasses
of compas_amd.workloads; outputs are complex128 / complex64 numpy arrays.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _threads(threads):
    if threads is None:
        return int(os.environ.get("ORACLE_THREADS", os.cpu_count() or 1))
    return int(threads)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _cf32(a):
    """complex64 array -> interleaved float32 view."""
    a = np.ascontiguousarray(a, dtype=np.complex64)
    return a.view(np.float32)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _real_t(precision):
    if precision == "f64":
        return np.float64, np.complex128
    if precision == "f32":
        return np.float32, np.complex64
    raise ValueError(precision)


def _creal(a, precision):
    _, ct = _real_t(precision)
    return np.ascontiguousarray(a, dtype=ct)


_native = None
NATIVE_PATH = os.path.join(HERE, "build", "liboracle_native.so")


def native_lib():
    """The oracle built for this host's ISA (make native: -march=native, FMA
    contraction), used to time the FP32 port on the host that runs the
    benchmark; None when it cannot be built here.  Always rebuilt once per
    process, since a library carried over from another host may not match."""
    global _native
    if _native is None:
        r = subprocess.run(["make", "-s", "-B", "-C", HERE, "native"], capture_output=True, text=True)
        _native = C.CDLL(NATIVE_PATH) if r.returncode == 0 and os.path.exists(NATIVE_PATH) else False
    return _native or None


def _fn(name, precision, native=False):
    if native:
        nl = native_lib()
        if nl is None:
            raise RuntimeError("the native oracle build failed on this host")
        return getattr(nl, f"{name}_{precision}")
    return getattr(lib(), f"{name}_{precision}")


def fisp(tissue, seq, precision="f64", threads=None, native=False):
    rt, ct = _real_t(precision)
    nv = tissue.nvoxels
    nTR = int(seq.RF.shape[0])
    prof = np.ascontiguousarray(seq.profiles, dtype=np.complex64)
    out = np.zeros((nTR * seq.undersampling, nv), dtype=ct)
    f = _fn("fisp_simulate", precision, native)
    args = [_f32(tissue.T1), _f32(tissue.T2), _f32(tissue.B1), _f32(tissue.B0), _cf32(seq.RF), _cf32(prof)]
    f(C.c_int(nv), *[_p(a) for a in args[:4]], C.c_int(nTR), _p(args[4]), C.c_int(prof.shape[0]), _p(args[5]),
      C.c_float(seq.TR), C.c_float(seq.TE), C.c_float(seq.TW), C.c_float(seq.TI),
      C.c_int(seq.max_state), C.c_int(seq.undersampling), C.c_int(seq.repetitions),
      C.c_int(int(seq.inversion)), C.c_int(int(seq.wait_spoiling)), _p(out), C.c_int(_threads(threads)))
    return out


def pssfp(tissue, seq, precision="f64", threads=None, native=False):
    rt, ct = _real_t(precision)
    nv = tissue.nvoxels
    nTR = int(seq.RF.shape[0])
    out = np.zeros((nTR, nv), dtype=ct)
    g = _cf32(seq.gamma_dt_RF)
    rf = _cf32(seq.RF)
    z = _f32(seq.z)
    arrs = [_f32(tissue.T1), _f32(tissue.T2), _f32(tissue.B1), _f32(tissue.B0)]
    _fn("pssfp_simulate", precision, native)(
        C.c_int(nv), *[_p(a) for a in arrs], C.c_int(nTR), _p(rf), C.c_int(seq.gamma_dt_RF.shape[0]), _p(g),
        *[C.c_float(t) for t in seq.dt], *[C.c_float(t) for t in seq.gamma_dt_GRz],
        C.c_int(z.shape[0]), _p(z), _p(out), C.c_int(_threads(threads)))
    return out


def _tissue_arrays(tissue):
    return [_f32(tissue.T2), _f32(tissue.B0), _cf32(tissue.rho), _f32(tissue.x), _f32(tissue.y)]


def signal_cartesian(tissue, traj, coils, echos, precision="f64", threads=None, dk_im=0.0, native=False):
    _, ct = _real_t(precision)
    nv = tissue.nvoxels
    nc = coils.shape[0]
    nr, ns = traj.nreadouts, traj.samples_per_readout
    out = np.zeros((nc, nr, ns), dtype=ct)
    e = _creal(echos, precision)
    cl = _cf32(coils)
    ta = _tissue_arrays(tissue)
    _fn("signal_cartesian", precision, native)(
        C.c_int(nv), *[_p(a) for a in ta], C.c_int(nr), C.c_int(ns), C.c_float(traj.delta_t),
        C.c_float(traj.delta_k), C.c_float(dk_im), C.c_int(nc), _p(cl), _p(e), _p(out), C.c_int(_threads(threads)))
    return out


def signal_spiral(tissue, traj, coils, echos, precision="f64", threads=None, native=False):
    _, ct = _real_t(precision)
    nv = tissue.nvoxels
    nc = coils.shape[0]
    nr, ns = traj.nreadouts, traj.samples_per_readout
    out = np.zeros((nc, nr, ns), dtype=ct)
    e = _creal(echos, precision)
    cl = _cf32(coils)
    ks, dk = _cf32(traj.k_start), _cf32(traj.delta_k)
    ta = _tissue_arrays(tissue)
    _fn("signal_spiral", precision, native)(
        C.c_int(nv), *[_p(a) for a in ta], C.c_int(nr), C.c_int(ns), C.c_float(traj.delta_t),
        _p(ks), _p(dk), C.c_int(nc), _p(cl), _p(e), _p(out), C.c_int(_threads(threads)))
    return out


def phase_encoding(tissue, traj, echos, precision="f64", threads=None):
    _, ct = _real_t(precision)
    nv = tissue.nvoxels
    nr = traj.nreadouts
    out = np.zeros((nr, nv), dtype=ct)
    e = _creal(echos, precision)
    ks = _cf32(traj.k_start)
    y = _f32(tissue.y)
    _fn("phase_encoding", precision)(C.c_int(nv), _p(y), C.c_int(nr), _p(ks), _p(e), _p(out), C.c_int(_threads(threads)))
    return out


def _jac_common(tissue, traj):
    return [_f32(tissue.T1), _f32(tissue.T2), _f32(tissue.B0), _cf32(tissue.rho), _f32(tissue.x), _f32(tissue.y)]


def jacobian(tissue, traj, coils, echos, dT1, dT2, vec, precision="f64", threads=None, dk_im=0.0):
    _, ct = _real_t(precision)
    nv = tissue.nvoxels
    nc = coils.shape[0]
    nr, ns = traj.nreadouts, traj.samples_per_readout
    out = np.zeros((nc, nr, ns), dtype=ct)
    ins = [_creal(a, precision) for a in (echos, dT1, dT2, vec)]
    cl = _cf32(coils)
    ta = _jac_common(tissue, traj)
    _fn("jacobian_product", precision)(
        C.c_int(nv), *[_p(a) for a in ta], C.c_int(nr), C.c_int(ns), C.c_float(traj.delta_t),
        C.c_float(traj.delta_k), C.c_float(dk_im), C.c_int(nc), _p(cl), *[_p(a) for a in ins], _p(out),
        C.c_int(_threads(threads)))
    return out


def jacobian_hermitian(tissue, traj, coils, echos, dT1, dT2, yv, precision="f64", threads=None, dk_im=0.0):
    _, ct = _real_t(precision)
    nv = tissue.nvoxels
    nc = coils.shape[0]
    nr, ns = traj.nreadouts, traj.samples_per_readout
    out = np.zeros((4, nv), dtype=ct)
    ins = [_creal(a, precision) for a in (echos, dT1, dT2, yv)]
    cl = _cf32(coils)
    ta = _jac_common(tissue, traj)
    _fn("jacobian_hermitian", precision)(
        C.c_int(nv), *[_p(a) for a in ta], C.c_int(nr), C.c_int(ns), C.c_float(traj.delta_t),
        C.c_float(traj.delta_k), C.c_float(dk_im), C.c_int(nc), _p(cl), *[_p(a) for a in ins], _p(out),
        C.c_int(_threads(threads)))
    return out


def _spiral6(name, shape, tissue, traj, coils, arrays, precision, threads, native=False):
    _, ct = _real_t(precision)
    nv = tissue.nvoxels
    nc = coils.shape[0]
    out = np.zeros(shape, dtype=ct)
    ins = [_creal(a, precision) for a in arrays]
    cl = _cf32(coils)
    ks, dk = _cf32(traj.k_start), _cf32(traj.delta_k)
    ta = _jac_common(tissue, traj)
    _fn(name, precision, native)(
        C.c_int(nv), *[_p(a) for a in ta], C.c_int(traj.nreadouts), C.c_int(traj.samples_per_readout),
        C.c_float(traj.delta_t), _p(ks), _p(dk), C.c_int(nc), _p(cl), *[_p(a) for a in ins], _p(out),
        C.c_int(_threads(threads)))
    return out


def jacobian_spiral(tissue, traj, coils, echos, dT1, dT2, dB1, dB0, vec, precision="f64", threads=None,
                    native=False):
    """J.v on a spiral trajectory, 6 parameters (T1, T2, B1, B0, rho_x, rho_y); vec [6][nv]."""
    nc = coils.shape[0]
    return _spiral6("jacobian_product_spiral", (nc, traj.nreadouts, traj.samples_per_readout), tissue, traj,
                    coils, (echos, dT1, dT2, dB1, dB0, vec), precision, threads, native)


def jacobian_hermitian_spiral(tissue, traj, coils, echos, dT1, dT2, dB1, dB0, yv, precision="f64", threads=None,
                              native=False):
    """J^H.y on a spiral trajectory, 6 parameters; returns [6][nv]."""
    return _spiral6("jacobian_hermitian_spiral", (6, tissue.nvoxels), tissue, traj, coils,
                    (echos, dT1, dT2, dB1, dB0, yv), precision, threads, native)


def rel_l2(x, ref):
    x = np.asarray(x, dtype=np.complex128)
    ref = np.asarray(ref, dtype=np.complex128)
    den = np.linalg.norm(ref.ravel())
    return float(np.linalg.norm((x - ref).ravel()) / (den if den > 0 else 1.0))
