"""Concurrent C-ABI calls on one context from several host threads.

Julia tasks on several threads may call the library at once with one context
(CompasToolkit.jl keeps it in task-local storage; contexts copied with
with_device share one runtime).  Each thread here runs its own simulate ->
signal -> J^H.y sequence on the shared context; ctypes releases the GIL
inside every library call, so the calls interleave in the library.  Every
result must equal, bitwise, the same sequence run alone beforehand (the
kernels are deterministic and the contraction tile of each shape is fixed by
the alone run).
"""

import threading

import numpy as np
import pytest

import compas_amd as ca
from compas_amd import workloads as W
from gpu_helpers import banner, context

pytestmark = pytest.mark.gpu


def host_case(i):
    rng = W.SplitMix64(0xC0A5_7000 + i)
    nv = [900, 2048, 1500, 700][i % 4]
    nTR = [48, 64, 40, 56][i % 4]
    n = int(np.ceil(np.sqrt(nv)))
    t = W.make_tissue(rng, n, 20.0, 30.0).slice(0, nv)
    seq = W.make_fisp(nTR, 60.0, 12) if i % 2 == 0 else W.make_pssfp(nTR, 50.0, nz=5, nRF=7)
    ns = 40
    tr = W.make_cartesian(n, nTR // n + 1, ns, 20.0, 5e-6)
    tr = W.CartesianTraj(nTR, ns, tr.delta_t, tr.k_start[:nTR], tr.delta_k)
    nc = 1 + i % 4
    coils = ((rng.normal(nc * nv) + 1j * rng.normal(nc * nv)) * 0.5).reshape(nc, nv).astype(np.complex64)
    y = ((rng.normal(nc * nTR * ns) + 1j * rng.normal(nc * nTR * ns)) * 0.5).reshape(nc, nTR, ns).astype(np.complex64)
    return t, seq, tr, coils, y


def run_sequence(case):
    """Objects made and every call issued on the calling thread."""
    t, seq, tr, coils, y = case
    p = ca.TissueParameters.from_tissue(t)
    s = ca.FispSequence.from_sequence(seq) if isinstance(seq, W.FispSeq) else ca.pSSFPSequence.from_sequence(seq)
    traj = ca.CartesianTrajectory.from_trajectory(tr)
    c = ca.CompasArray.from_numpy(coils)
    echos = ca.simulate_magnetization(p, s)
    sig = ca.magnetization_to_signal(echos, p, traj, c, method="matmul_fast")
    d = {f: ca.simulate_magnetization_derivative(getattr(ca, f), echos, p, s) for f in ("T1", "T2")}
    jhy = ca.compute_jacobian_hermitian(echos, d, p, traj, c, ca.CompasArray.from_numpy(y), method="gemm_fast")
    return echos.collect(), sig.collect(), jhy.collect()


def test_concurrent_calls_on_one_context():
    banner("threads: 4 threads x 3 sequences on one context")
    context()
    cases = [host_case(i) for i in range(4)]
    alone = [run_sequence(c) for c in cases]  # also fixes each contraction shape's tile
    results, errors = {}, []

    def worker(i):
        try:
            for rep in range(3):
                results[(i, rep)] = run_sequence(cases[i])
        except Exception as e:  # reported below; a thread's exception would be lost otherwise
            errors.append((i, repr(e)))

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=300)
    assert not any(th.is_alive() for th in threads), "a thread did not finish"
    assert not errors, errors
    for (i, rep), got in results.items():
        for name, a, b in zip(("echos", "signal", "J^H.y"), got, alone[i]):
            assert np.array_equal(a, b), (i, rep, name)
    assert len(results) == 12


def test_errors_stay_with_their_thread():
    """An error raised on one thread is reported to that thread only, while
    another thread's calls on the same context succeed."""
    banner("threads: per-thread error channel under concurrent calls")
    context()
    case = host_case(0)
    bad = ca.CompasArray.from_numpy(np.zeros((1, 3, 4), np.complex64))
    out = {}

    def failing():
        t, seq, tr, coils, y = case
        p = ca.TissueParameters.from_tissue(t)
        s = ca.FispSequence.from_sequence(seq)
        traj = ca.CartesianTrajectory.from_trajectory(tr)
        c = ca.CompasArray.from_numpy(coils)
        echos = ca.simulate_magnetization(p, s)
        d = {f: ca.simulate_magnetization_derivative(getattr(ca, f), echos, p, s) for f in ("T1", "T2")}
        try:
            for _ in range(5):
                ca.compute_jacobian_hermitian(echos, d, p, traj, c, bad, method="gemm_fast")
            out["failing"] = "no error"
        except ca.CompasError as e:
            out["failing"] = str(e)

    def working():
        try:
            out["working"] = [run_sequence(case)[1] for _ in range(3)]
        except Exception as e:
            out["working"] = repr(e)

    ths = [threading.Thread(target=failing), threading.Thread(target=working)]
    for th in ths:
        th.start()
    for th in ths:
        th.join(timeout=300)
    assert out["failing"] != "no error"
    assert isinstance(out["working"], list), out["working"]
    want = run_sequence(case)[1]
    assert all(np.array_equal(s, want) for s in out["working"])
