"""Request server (continuous batching) through the C-ABI: every result is the
same in the concurrent mix as when the request runs alone, requests share
work-table launches, small requests pass large ones, admission respects the
memory budget, and errors come back per request."""

import numpy as np
import pytest

import compas_amd as ca
from compas_amd import workloads as W
from gpu_helpers import banner, context, rel

pytestmark = pytest.mark.gpu


def make_request(i):
    """Seeded request i: kind, sizes and objects (small shapes)."""
    rng = W.SplitMix64(0xC0A50005 + i)
    kind = ("fisp", "pssfp", "jhy")[i % 3]
    spiral = (i // 3) % 2 == 1
    nv = [300, 1024, 2500, 777][i % 4]
    nTR = [40, 64, 25, 51][(i // 2) % 4]
    nc = 1 + (i * 3) % 5
    n = int(np.ceil(np.sqrt(nv)))
    t = W.make_tissue(rng, n, 20.0, 40.0).slice(0, nv)
    if kind == "pssfp":
        seq = W.make_pssfp(nTR, 50.0, nz=5, nRF=7)
    else:
        seq = W.make_fisp(nTR, 60.0, 12)
    ns = 32 + 7 * (i % 3)
    if spiral:
        traj = W.make_spiral(nTR, ns, 20.0, n, 5e-6)
    else:
        traj = W.make_cartesian(n, max(1, nTR // n + 1), ns, 20.0, 5e-6)
        traj = W.CartesianTraj(nTR, ns, traj.delta_t, traj.k_start[:nTR], traj.delta_k)
    coils = ((rng.normal(nc * nv) + 1j * rng.normal(nc * nv)) * 0.5).reshape(nc, nv).astype(np.complex64)
    y = ((rng.normal(nc * nTR * ns) + 1j * rng.normal(nc * nTR * ns)) * 0.5).reshape(nc, nTR, ns).astype(np.complex64)
    return kind, t, seq, traj, coils, y


def device_objects(req):
    kind, t, seq, traj, coils, y = req
    p = ca.TissueParameters.from_tissue(t)
    s = ca.FispSequence.from_sequence(seq) if isinstance(seq, W.FispSeq) else ca.pSSFPSequence.from_sequence(seq)
    tr = (ca.SpiralTrajectory.from_trajectory(traj) if isinstance(traj, W.SpiralTraj)
          else ca.CartesianTrajectory.from_trajectory(traj))
    return kind, p, s, tr, ca.CompasArray.from_numpy(coils), ca.CompasArray.from_numpy(y)


def run_alone(obj):
    """The same request as plain library calls on the default context."""
    kind, p, s, tr, coils, y = obj
    echos = ca.simulate_magnetization(p, s)
    if kind != "jhy":
        return ca.magnetization_to_signal(echos, p, tr, coils, method="matmul_fast").collect()
    d = {f: ca.simulate_magnetization_derivative(getattr(ca, f), echos, p, s) for f in ("T1", "T2", "B1", "B0")}
    if isinstance(tr, ca.SpiralTrajectory):
        return ca.compute_jacobian_hermitian_spiral(echos, d, p, tr, coils, y).collect()
    return ca.compute_jacobian_hermitian(echos, d, p, tr, coils, y, method="gemm_fast").collect()


def submit(server, obj):
    kind, p, s, tr, coils, y = obj
    if kind == "jhy":
        return server.submit_jacobian_hermitian(p, s, tr, coils, y, method="gemm_fast")
    return server.submit_simulate_signal(p, s, tr, coils, method="matmul_fast")


def collect_all(server, ticket, kind, spiral):
    """Every result of a finished request: echos, signal or J^H.y, and for
    J^H.y the finite differences the server computed."""
    out = [server.result(ticket, 0).collect(), server.result(ticket, 1).collect()]
    if kind == "jhy":
        out += [server.result(ticket, 2 + k).collect() for k in range(4 if spiral else 2)]
    return out


def test_results_equal_alone_and_in_the_mix():
    """18 mixed requests submitted at once share work-table launches; each
    request's results equal, bitwise, the same request served alone, and the
    plain library calls within 1e-6 (echoes, signal, and J^H.y of the same
    derivatives; the derivatives themselves are forward differences whose
    1/delta = 1e4 amplifies any rounding difference of the echoes)."""
    banner("serving: 18 mixed requests vs alone")
    context()
    reqs = [make_request(i) for i in range(18)]
    objs = [device_objects(r) for r in reqs]
    server = ca.Server(slots=4, memory_budget=8 << 30)
    tickets = [submit(server, o) for o in objs]
    times = [server.wait(t, timeout=120) for t in tickets]
    assert all(tm is not None for tm in times)
    spiral = [isinstance(r[3], W.SpiralTraj) for r in reqs]
    mixed = [collect_all(server, t, o[0], sp) for t, o, sp in zip(tickets, objs, spiral)]
    for t in tickets:
        server.release(t)
    stats = server.batch_stats()
    print("batch stats (mix):", stats)
    assert stats["shared_launches"] >= 1 and stats["max_requests"] >= 2
    # a = arrival <= s = admission <= d = done on the server clock
    for a, s_, d in times:
        assert a <= s_ <= d
    for i, o in enumerate(objs):
        t = submit(server, o)
        assert server.wait(t, timeout=120) is not None
        alone = collect_all(server, t, o[0], spiral[i])
        server.release(t)
        for k, (x, y) in enumerate(zip(mixed[i], alone)):
            assert np.array_equal(x, y), (i, o[0], k)
    server.close()
    for i, o in enumerate(objs):
        kind, p, s, tr, coils, y = o
        echos = ca.simulate_magnetization(p, s)
        assert rel(mixed[i][0], echos.collect()) <= 1e-6, (i, kind)
        if kind != "jhy":
            sig = ca.magnetization_to_signal(echos, p, tr, coils, method="matmul_fast").collect()
            assert rel(mixed[i][1], sig) <= 1e-6, (i, kind)
            continue
        d = [ca.CompasArray.from_numpy(x) for x in mixed[i][2:]]
        if isinstance(tr, ca.SpiralTrajectory):
            want = ca.compute_jacobian_hermitian_spiral(echos, dict(zip(("T1", "T2", "B1", "B0"), d)), p, tr,
                                                        coils, y).collect()
        else:
            want = ca.compute_jacobian_hermitian(echos, {"T1": d[0], "T2": d[1]}, p, tr, coils, y,
                                                 method="gemm_fast").collect()
        assert rel(mixed[i][1], want) <= 1e-6, (i, kind)
        fd = ca.simulate_magnetization_derivative(ca.T1, echos, p, s).collect()
        assert rel(mixed[i][2], fd) <= 1e-3, (i, kind)


def test_small_requests_pass_a_large_one():
    """A large pSSFP request (100000 voxels x 500 TRs x 35 positions)
    submitted first does not hold up small requests submitted after it: they
    are admitted at once, share launches with it, and finish before it."""
    banner("serving: small requests alongside a large one")
    context()
    rng = W.SplitMix64(0xC0A5_0077)
    big_t = W.make_tissue(rng, 317, 22.4, 40.0).slice(0, 100000)
    big_p = ca.TissueParameters.from_tissue(big_t)
    big_s = ca.pSSFPSequence.from_sequence(W.make_pssfp(500, 60.0))
    small = [device_objects(make_request(i)) for i in (1, 3, 7, 9, 13, 15)]
    coils = ca.CompasArray.from_numpy(np.ones((1, 100000), np.complex64))
    sp = ca.SpiralTrajectory.from_trajectory(W.make_spiral(500, 32, 22.4, 317, 5e-6))
    server = ca.Server(slots=2, memory_budget=64 << 30)
    tb = server.submit_simulate_signal(big_p, big_s, sp, coils)
    ts = [submit(server, o) for o in small]
    done_big = server.wait(tb, timeout=120)
    done_small = [server.wait(t, timeout=120) for t in ts]
    stats = server.batch_stats()
    server.close()
    print("big", done_big, "small", done_small, stats)
    assert done_big is not None and all(d is not None for d in done_small)
    for a, s_, d in done_small:
        assert s_ - a < 0.05  # admitted while the large request runs
        assert d < done_big[2]  # and finished before it
    assert stats["shared_launches"] >= 1


def test_fisp_beyond_table_states_runs_on_a_slot():
    """FISP with more than 32 states has no work-table kernel: the request
    runs all its calls on a slot stream and gives the plain library result."""
    banner("serving: FISP N=40 on a slot")
    context()
    t = W.make_tissue(W.SplitMix64(0xC0A5_0078), 30, 20.0, 30.0).slice(0, 800)
    seq = W.make_fisp(60, 50.0, 40)
    tr = W.make_cartesian(30, 3, 40, 20.0, 5e-6)
    tr = W.CartesianTraj(60, 40, tr.delta_t, tr.k_start[:60], tr.delta_k)
    p, s = ca.TissueParameters.from_tissue(t), ca.FispSequence.from_sequence(seq)
    traj = ca.CartesianTrajectory.from_trajectory(tr)
    coils = ca.CompasArray.from_numpy(np.ones((2, 800), np.complex64))
    server = ca.Server(slots=2, memory_budget=8 << 30)
    tk = server.submit_simulate_signal(p, s, traj, coils)
    assert server.wait(tk, timeout=60) is not None
    got = server.result(tk).collect()
    server.close()
    echos = ca.simulate_magnetization(p, s)
    want = ca.magnetization_to_signal(echos, p, traj, coils, method="matmul_fast").collect()
    assert rel(got, want) <= 1e-6


def test_memory_budget_serialises_requests():
    banner("serving: budget below one request runs them one at a time")
    context()
    objs = [device_objects(make_request(i)) for i in (0, 3, 6)]
    server = ca.Server(slots=4, memory_budget=1)
    tickets = [submit(server, o) for o in objs]
    times = sorted((server.wait(t, timeout=120) for t in tickets), key=lambda x: x[1])
    server.close()
    for (a0, s0, d0), (a1, s1, d1) in zip(times, times[1:]):
        assert s1 >= d0  # admitted only after the previous one finished


def test_request_errors_are_reported():
    banner("serving: errors")
    context()
    kind, p, s, tr, coils, y = device_objects(make_request(2))
    server = ca.Server(slots=2, memory_budget=1 << 30)
    bad_y = ca.CompasArray.from_numpy(np.zeros((1, 3, 4), np.complex64))
    with pytest.raises(ca.CompasError):
        t = server.submit_jacobian_hermitian(p, s, tr, coils, bad_y)
        server.wait(t, timeout=60)
    t = server.submit_simulate_signal(p, s, tr, coils)
    assert server.wait(t, timeout=60) is not None
    server.close()


def test_request_beyond_device_memory_fails_alone():
    """A request whose echoes cannot be allocated (1e6 voxels x 40000 TRs,
    320 GB, above the budget: admitted when nothing else runs) ends with an
    allocation error; the request submitted beside it and one submitted after
    it complete with the plain library result (the failure does not leak into
    later launches of the dispatcher)."""
    banner("serving: a request beyond device memory")
    context()
    t = W.make_tissue(W.SplitMix64(0xC0A5_0079), 1000, 22.4, 40.0)
    nTR = 40000
    tr = W.CartesianTraj(nTR, 8, 5e-6, np.zeros(nTR, np.complex64), float(np.float32(2 * np.pi / 22.4)))
    p = ca.TissueParameters.from_tissue(t)
    s = ca.FispSequence.from_sequence(W.make_fisp(nTR, 60.0, 12))
    traj = ca.CartesianTrajectory.from_trajectory(tr)
    coils = ca.CompasArray.from_numpy(np.ones((1, t.nvoxels), np.complex64))
    objs = [device_objects(make_request(i)) for i in (0, 1)]
    server = ca.Server(slots=2, memory_budget=8 << 30)
    big = server.submit_simulate_signal(p, s, traj, coils)
    first = submit(server, objs[0])
    assert server.wait(first, timeout=120) is not None
    with pytest.raises(ca.CompasError, match="allocation"):
        server.wait(big, timeout=120)
    after = submit(server, objs[1])
    assert server.wait(after, timeout=120) is not None
    got = server.result(after).collect()
    server.close()
    assert rel(got, run_alone(objs[1])) <= 1e-6


@pytest.mark.parametrize("i", [2, 5])  # Cartesian and spiral J^H.y
def test_inputs_still_queued_on_the_base_stream(i):
    """A residual produced by library calls on the server's base context and
    submitted at once (no host wait): the slot waits for the base stream's
    work queued before the submission, so the result equals the alone path."""
    banner(f"serving: J^H.y on a residual still in flight (request {i})")
    ctx = context()
    kind, p, s, tr, coils, y = device_objects(make_request(i))
    assert kind == "jhy"
    server = ca.Server(slots=2, memory_budget=8 << 30)
    echos = ca.simulate_magnetization(p, s)
    sig = ca.magnetization_to_signal(echos, p, tr, coils, method="matmul_fast")
    # compas_compute_residual without the objective pointer: nothing waits on the host
    res = ca._wrap(ca.lib().compas_compute_residual(ctx.ptr, sig.ptr, y.ptr, None))
    t = server.submit_jacobian_hermitian(p, s, tr, coils, res, method="gemm_fast")
    assert server.wait(t, timeout=120) is not None
    got = server.result(t).collect()
    server.release(t)
    server.close()
    want = run_alone((kind, p, s, tr, coils, res))
    assert np.array_equal(got, want)
