"""Request server (continuous batching) through the C-ABI: every result is the
same in the concurrent mix as when the request runs alone, requests overlap
on the device, admission respects the memory budget, and errors come back
per request."""

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


def test_results_equal_alone_and_in_the_mix():
    banner("serving: 18 mixed requests on 4 slots vs alone")
    context()
    objs = [device_objects(make_request(i)) for i in range(18)]
    server = ca.Server(slots=4, memory_budget=8 << 30)
    tickets = [submit(server, o) for o in objs]
    times = [server.wait(t, timeout=120) for t in tickets]
    assert all(tm is not None for tm in times)
    mixed = [server.result(t).collect() for t in tickets]
    for t in tickets:
        server.release(t)
    server.close()
    for i, o in enumerate(objs):
        alone = run_alone(o)
        assert mixed[i].shape == alone.shape, i
        assert rel(mixed[i], alone) <= 1e-6, (i, o[0])
    # arrival <= start <= done on the server clock
    for a, s, d in times:
        assert a <= s <= d


def test_requests_overlap_and_later_arrivals_start_early():
    banner("serving: continuous admission on 2 slots")
    context()
    objs = [device_objects(make_request(i)) for i in (1, 4, 7, 10, 2, 5, 8, 11)]
    server = ca.Server(slots=2, memory_budget=8 << 30)
    tickets = [submit(server, o) for o in objs]
    times = [server.wait(t, timeout=120) for t in tickets]
    server.close()
    starts = sorted(s for _, s, _ in times)
    dones = sorted(d for _, _, d in times)
    # two requests were running at once at some point
    assert any(times[i][1] < times[j][2] and times[j][1] < times[i][2]
               for i in range(len(times)) for j in range(i + 1, len(times)))
    # the third request started as soon as the first finished, before the last one finished
    assert starts[2] <= dones[-1]
    assert starts[2] >= dones[0] - 1e-3


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
