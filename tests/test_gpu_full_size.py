"""Workloads at their full sizes on the GPU, checked against the FP64 oracle
on samples whose values do not depend on the rest of the output: voxel
slices of the simulations and of J^H.y (each voxel's recursion, and each
voxel's J^H.y entry, is independent of the other voxels) and readout
subsets of the signal and J.v (each readout row is independent).

* Workload 3: pSSFP 50176 voxels x 1120 TRs x 35 slice positions x 25 RF
  steps, golden-angle spiral 4 coils x 1120 readouts x 512 samples, the
  six-parameter J.v and J^H.y on the forward-mode derivatives.
* The serving mix at its own sizes: 26 requests of the mix (FISP and pSSFP
  simulate + signal, J^H.y; 1e3..2e5 voxels, 250..2000 TRs, 1..8 coils,
  Cartesian and spiral) submitted at once to the request server, every
  result checked, and a sample re-run alone (bitwise equal).

Tolerances (relative L2): magnetisation 1e-5, signal / J.v / J^H.y 1e-4.
The products are checked with the GPU's own derivative arrays as the
oracle's inputs, so the check is of the products alone.
"""

import numpy as np
import pytest

import compas_amd as ca
from compas_amd import workloads as W
from compas_amd.serving_mix import ServeCatalog, serve_request_spec, serve_submit
from gpu_helpers import banner, context, rel

pytestmark = pytest.mark.gpu

MAG_TOL = 1e-5
SIG_TOL = 1e-4


def spiral_rows(traj, ridx):
    return W.SpiralTraj(len(ridx), traj.samples_per_readout, traj.delta_t, traj.k_start[ridx], traj.delta_k[ridx])


def cartesian_rows(traj, ridx):
    return W.CartesianTraj(len(ridx), traj.samples_per_readout, traj.delta_t, traj.k_start[ridx], traj.delta_k)


def test_workload3_full_size(oracle):
    banner("workload 3 full size: pSSFP 50176 x 1120 x 35, spiral 4 x 1120 x 512, six-parameter products")
    context()
    wl = W.config(3)
    p = ca.TissueParameters.from_tissue(wl.tissue)
    s = ca.pSSFPSequence.from_sequence(wl.sequence)
    traj = ca.SpiralTrajectory.from_trajectory(wl.trajectory)
    coils = ca.CompasArray.from_numpy(wl.coils)
    echos, *d = ca.simulate_magnetization_tangents([ca.T1, ca.T2, ca.B1, ca.B0], p, s, echos=True)
    dd = dict(zip(("T1", "T2", "B1", "B0"), d))
    nv, nr, ns, nc = wl.tissue.nvoxels, 1120, 512, wl.coils.shape[0]
    g = np.random.default_rng(3)
    v = g.standard_normal((6, nv)).astype(np.complex64)
    y = ((g.standard_normal((nc, nr, ns)) + 1j * g.standard_normal((nc, nr, ns))) / np.sqrt(2)).astype(np.complex64)
    sig = ca.magnetization_to_signal(echos, p, traj, coils).collect()
    jv = ca.compute_jacobian_spiral(echos, dd, p, traj, coils, v).collect()
    jhy = ca.compute_jacobian_hermitian_spiral(echos, dd, p, traj, coils, y).collect()
    e = echos.collect()
    dh = [x.collect() for x in d]
    assert e.shape == (nr, nv) and sig.shape == (nc, nr, ns) and jv.shape == (nc, nr, ns) and jhy.shape == (6, nv)

    for begin in (0, 31000, nv - 32):
        sl = slice(begin, begin + 32)
        err = rel(e[:, sl], oracle.pssfp(wl.tissue.slice(begin, begin + 32), wl.sequence, "f64"))
        print(f"echoes voxels {begin}..{begin + 31}: {err:.2e}")
        assert err < MAG_TOL
    ridx = np.array([0, 561, 1119])
    sub = spiral_rows(wl.trajectory, ridx)
    args = [x[ridx].astype(np.complex128) for x in [e] + dh]
    err = rel(sig[:, ridx], oracle.signal_spiral(wl.tissue, sub, wl.coils, args[0], "f64"))
    print(f"signal readouts {list(ridx)}: {err:.2e}")
    assert err < SIG_TOL
    err = rel(jv[:, ridx], oracle.jacobian_spiral(wl.tissue, sub, wl.coils, *args, v, "f64"))
    print(f"J.v readouts {list(ridx)}: {err:.2e}")
    assert err < SIG_TOL
    begin = 20000
    sl = slice(begin, begin + 32)
    cols = [x[:, sl].astype(np.complex128) for x in [e] + dh]
    want = oracle.jacobian_hermitian_spiral(wl.tissue.slice(begin, begin + 32), wl.trajectory, wl.coils[:, sl],
                                            *cols, y, "f64")
    err = rel(jhy[:, sl], want)
    print(f"J^H.y voxels {begin}..{begin + 31}: {err:.2e}")
    assert err < SIG_TOL


# the first 26 requests of the serving mix without the two largest pSSFP
# simulations (ids 12 and 17, 0.5 s of GPU time each), plus two of the
# largest requests of the mix: id 35 (J^H.y, 191569 voxels x 573 TRs, 6
# coils) and id 129 (FISP + spiral signal, 193854 voxels x 852 TRs)
MIX_IDS = [i for i in range(26) if i not in (12, 17)] + [35, 129]


def check_request(oracle, cat, spec, res):
    """res = [echos, signal or J^H.y, finite differences...] of the server."""
    t, n, coils, y = cat.host_request(spec)
    seq = cat.host_sequence(spec["kind"], spec["nTR"])
    traj = cat.host_trajectory(spec["spiral"], spec["nTR"], n)
    e = res[0]
    nv = spec["nv"]
    b = nv // 2
    sl = slice(b, min(nv, b + 32))
    sim = oracle.pssfp if spec["kind"] == "pssfp" else oracle.fisp
    err = rel(e[:, sl], sim(t.slice(sl.start, sl.stop), seq, "f64"))
    assert err < MAG_TOL, (spec, "echoes", err)
    if spec["kind"] != "jhy":
        ridx = np.array([0, spec["nTR"] - 1])
        if spec["spiral"]:
            want = oracle.signal_spiral(t, spiral_rows(traj, ridx), coils, e[ridx].astype(np.complex128), "f64")
        else:
            want = oracle.signal_cartesian(t, cartesian_rows(traj, ridx), coils, e[ridx].astype(np.complex128), "f64")
        err2 = rel(res[1][:, ridx], want)
        assert err2 < SIG_TOL, (spec, "signal", err2)
        return err, err2
    cols = [x[:, sl].astype(np.complex128) for x in [e] + res[2:]]
    tsub = t.slice(sl.start, sl.stop)
    if spec["spiral"]:
        want = oracle.jacobian_hermitian_spiral(tsub, traj, coils[:, sl], *cols, y, "f64")
    else:
        want = oracle.jacobian_hermitian(tsub, traj, coils[:, sl], *cols, y, "f64")
    err2 = rel(res[1][:, sl], want)
    assert err2 < SIG_TOL, (spec, "J^H.y", err2)
    return err, err2


def fetch(server, ticket, spec):
    n = 2 + (0 if spec["kind"] != "jhy" else (4 if spec["spiral"] else 2))
    return [server.result(ticket, k).collect() for k in range(n)]


def test_serving_mix_full_size(oracle):
    banner(f"serving mix at full size: requests {MIX_IDS}")
    context()
    cat = ServeCatalog()
    specs = [serve_request_spec(i) for i in MIX_IDS]
    reqs = [cat.request(sp) for sp in specs]
    server = ca.Server(slots=4, memory_budget=96 << 30)
    tickets = [serve_submit(server, r) for r in reqs]
    for sp, tk in zip(specs, tickets):
        assert server.wait(tk, timeout=240) is not None, sp
    stats = server.batch_stats()
    print("batch stats:", stats)
    assert stats["shared_launches"] >= 1
    kept = {}
    for sp, tk in zip(specs, tickets):
        res = fetch(server, tk, sp)
        server.release(tk)
        em, es = check_request(oracle, cat, sp, res)
        print(f"request {sp['id']} {sp['kind']} nv={sp['nv']} nTR={sp['nTR']} nc={sp['nc']} "
              f"{'spiral' if sp['spiral'] else 'cartesian'}: echoes {em:.2e}, product {es:.2e}", flush=True)
        if sp["id"] in (0, 2, 35, 129):
            kept[sp["id"]] = res
        del res
    # the same requests alone: bitwise equal
    for sp, r in zip(specs, reqs):
        if sp["id"] not in kept:
            continue
        tk = serve_submit(server, r)
        assert server.wait(tk, timeout=240) is not None
        alone = fetch(server, tk, sp)
        server.release(tk)
        for k, (x, y) in enumerate(zip(kept[sp["id"]], alone)):
            assert np.array_equal(x, y), (sp, k)
    server.close()
