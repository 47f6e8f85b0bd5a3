"""Graph capture and replay of library calls (compas_graph_*).

A captured sequence must give, on every replay, exactly what the same calls
give when run eagerly -- with unrelated eager work, allocations and host
synchronisations between the replays -- and must follow its input arrays
when they are rewritten between launches.
"""

import numpy as np
import pytest

import compas_amd as ca
from compas_amd import workloads as W
from gpu_helpers import banner, context, objects

pytestmark = pytest.mark.gpu


def eager(params, seq, traj, coils, v, y):
    e = ca.simulate_magnetization(params, seq)
    d1 = ca.simulate_magnetization_tangent(ca.T1, params, seq)
    d2 = ca.simulate_magnetization_tangent(ca.T2, params, seq)
    s = ca.magnetization_to_signal(e, params, traj, coils)
    jv = ca.compute_jacobian(e, (d1, d2), params, traj, coils, v, method="gemm_fast")
    jh = ca.compute_jacobian_hermitian(e, (d1, d2), params, traj, coils, y, method="gemm_fast")
    return e, s, jv, jh


def test_captured_step_replays_match_eager():
    banner("graph capture: simulate + tangents + signal + J.v + J^H.y")
    context()
    wl = W.config(2, scale=4)
    params, seq, traj = objects(wl)
    coils = ca.CompasArray.from_numpy(wl.coils)
    nv = wl.tissue.nvoxels
    nc, nr, ns = wl.coils.shape[0], traj.nreadouts, traj.samples_per_readout
    rng = np.random.default_rng(5)
    v_host = rng.normal(size=(4, nv)).astype(np.complex64)
    y = ca.CompasArray.from_numpy((rng.normal(size=(nc, nr, ns)) + 1j * rng.normal(size=(nc, nr, ns)))
                                  .astype(np.complex64))
    v = ca.CompasArray.from_numpy(v_host)

    want = [a.collect() for a in eager(params, seq, traj, coils, v, y)]  # also tunes the contraction

    with ca.capture() as g:
        outs = eager(params, seq, traj, coils, v, y)
    other = W.config(2, scale=8)
    p2, s2, _ = objects(other)
    for i in range(3):
        g.launch()
        got = [a.collect() for a in outs]
        for name, a, b in zip(("echos", "signal", "J.v", "J^H.y"), got, want):
            assert np.array_equal(a, b), f"replay {i}: {name} differs from the eager result"
        # unrelated eager work, allocations and synchronisation between replays
        junk = ca.simulate_magnetization(p2, s2).collect()
        assert np.isfinite(junk).all()

    # a rewritten input is read by the next replay
    v_new = rng.normal(size=(4, nv)).astype(np.complex64)
    v.write(v_new)
    g.launch()
    got_jv = outs[2].collect()
    ref_jv = eager(params, seq, traj, coils, ca.CompasArray.from_numpy(v_new), y)[2].collect()
    assert np.array_equal(got_jv, ref_jv)
    g.close()


def test_capture_of_pssfp_and_spiral():
    banner("graph capture: pSSFP + spiral signal + spiral J^H.y")
    context()
    wl = W.config(3, scale=8)
    params, seq, traj = objects(wl)
    coils = ca.CompasArray.from_numpy(wl.coils)
    nr, ns = traj.nreadouts, traj.samples_per_readout
    y = ca.CompasArray.from_numpy(wl.extra["y"][:, :nr, :ns].copy())

    def calls():
        e = ca.simulate_magnetization(params, seq)
        ds = [ca.simulate_magnetization_tangent(f, params, seq) for f in (ca.T1, ca.T2, ca.B1, ca.B0)]
        s = ca.magnetization_to_signal(e, params, traj, coils)
        jh = ca.compute_jacobian_hermitian_spiral(e, ds, params, traj, coils, y)
        return e, s, jh

    want = [a.collect() for a in calls()]
    with ca.capture() as g:
        outs = calls()
    for i in range(3):
        g.launch()
        for a, b in zip(outs, want):
            assert np.array_equal(a.collect(), b), f"replay {i} differs"
        ca.simulate_magnetization(params, seq).collect()
    g.close()
