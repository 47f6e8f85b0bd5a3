"""Request server through its wire format: a request submitted as a frame
gives bitwise the result of the same request submitted as library objects,
failures come back as error frames, and the TCP front end returns pipelined
results in order from several connections."""

import threading

import numpy as np
import pytest

import compas_amd as ca
from compas_amd import wire
from gpu_helpers import banner, context
from test_gpu_serving import device_objects, make_request, submit

pytestmark = pytest.mark.gpu


def frame_of(req, echos=False):
    kind, t, seq, traj, coils, y = req
    if kind == "jhy":
        return wire.encode_request(wire.KIND_JACOBIAN_HERMITIAN, t, seq, traj, coils, y, jacobian_method="gemm_fast",
                                   return_echos=echos)
    return wire.encode_request(wire.KIND_SIMULATE_SIGNAL, t, seq, traj, coils, signal_method="matmul_fast",
                               return_echos=echos)


def object_results(server, reqs):
    objs = [device_objects(r) for r in reqs]
    tickets = [submit(server, o) for o in objs]
    out = []
    for t in tickets:
        assert server.wait(t, timeout=120) is not None
        out.append((server.result(t, 0).collect(), server.result(t, 1).collect()))
        server.release(t)
    return out


def test_wire_requests_equal_object_requests():
    banner("serving: 12 requests as frames vs as objects")
    context()
    reqs = [make_request(i) for i in range(12)]
    server = ca.Server(slots=4, memory_budget=8 << 30)
    want = object_results(server, reqs)
    tickets = [server.submit_wire(frame_of(r, echos=True)) for r in reqs]
    for i, t in enumerate(tickets):
        assert server.wait_done(t)
        res = wire.decode_result(server.result_wire(t))
        server.release(t)
        assert res["ok"] and res["ticket"] == t
        assert res["t_arrival"] <= res["t_start"] <= res["t_done"]
        assert np.array_equal(res["echos"], want[i][0]), f"request {i}: echoes differ"
        key = "jhy" if reqs[i][0] == "jhy" else "signal"
        assert np.array_equal(res[key], want[i][1]), f"request {i}: {key} differs"
    server.close()


def test_rejected_and_failed_frames():
    banner("serving: malformed frame, error result frame")
    context()
    server = ca.Server(slots=2, memory_budget=4 << 30)
    good = frame_of(make_request(2))  # a J^H.y request
    with pytest.raises(ca.CompasError, match="malformed frame"):
        server.submit_wire(good[:-8])
    kind, e = wire.parse_frame(good, wire.REQUEST_MAGIC)
    e["residual"] = e["residual"][:, :, :5]
    with pytest.raises(ca.CompasError, match="residual"):
        server.submit_wire(wire.build_frame(wire.REQUEST_MAGIC, kind, e))
    # accepted by the parser, fails while running (max_state above the limit): an error frame
    kind, e = wire.parse_frame(frame_of(make_request(0)), wire.REQUEST_MAGIC)
    e["max_state"] = np.array(129, np.int32)
    t = server.submit_wire(wire.build_frame(wire.REQUEST_MAGIC, kind, e))
    assert not server.wait_done(t)
    res = wire.decode_result(server.result_wire(t))
    assert not res["ok"] and "max_state" in res["error"]
    server.release(t)
    server.close()


def test_tcp_front_end_pipelined_clients():
    banner("serving: TCP front end, 2 clients x 6 pipelined requests")
    context()
    reqs = [make_request(i) for i in range(12)]
    server = ca.Server(slots=4, memory_budget=8 << 30)
    want = object_results(server, reqs)
    fe = wire.TcpFrontEnd(server)
    results = {}

    def client(ids):
        c = wire.TcpClient(fe.address)
        for i in ids:  # all sent before the first is read back
            c.send(frame_of(reqs[i]))
        c.send(b"not a frame, answered with an error")
        results.update({i: c.recv() for i in ids})
        results[("bad", ids[0])] = c.recv()
        c.close()

    threads = [threading.Thread(target=client, args=(list(range(k, 12, 2)),)) for k in range(2)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=300)
    fe.close()
    server.close()
    for i in range(12):
        r = results[i]
        assert r["ok"] and "echos" not in r  # not asked for
        key = "jhy" if reqs[i][0] == "jhy" else "signal"
        assert np.array_equal(r[key], want[i][1]), f"request {i}"
    for k in (0, 1):
        bad = results[("bad", k)]
        assert not bad["ok"] and "malformed frame" in bad["error"]


def test_spiral_jhy_frame_without_b1_b0():
    """Tissue frames may omit B1 and B0 (defaults 1 and 0).  The finite
    difference of the echoes with respect to an absent field is zero, as in the
    reference, whose perturbed parameters keep has_b1/has_b0
    (derivative.cu:43-45).  B1 enters the spiral signal only through the echoes,
    so its row of the six-parameter J^H.y is zero; B0 also enters the readout
    phase (dPhi/dB0 = (s - ns/2) i 2 pi dt Phi), so its row keeps that term.
    The T1, T2 and rho rows equal those of the frame that sends the defaults."""
    banner("serving: spiral J^H.y frame without B1/B0")
    context()
    kind, t, seq, traj, coils, y = make_request(5)
    assert kind == "jhy" and traj.__class__.__name__ == "SpiralTraj"
    server = ca.Server(slots=2, memory_budget=8 << 30)
    explicit = t.replace(B1=np.ones_like(t.B1), B0=np.zeros_like(t.B0))
    frames = [
        wire.encode_request(wire.KIND_JACOBIAN_HERMITIAN, t, seq, traj, coils, y, jacobian_method="gemm_fast",
                            fields=("T1", "T2", "rho", "x", "y")),
        wire.encode_request(wire.KIND_JACOBIAN_HERMITIAN, explicit, seq, traj, coils, y,
                            jacobian_method="gemm_fast"),
    ]
    out = []
    for f in frames:
        tk = server.submit_wire(f)
        assert server.wait_done(tk)
        res = wire.decode_result(server.result_wire(tk))
        server.release(tk)
        assert res["ok"], res.get("error")
        out.append(res["jhy"])
    server.close()
    absent, given = out
    assert absent.shape == (6, t.nvoxels)
    assert np.all(absent[2] == 0) and np.any(given[2] != 0)
    assert np.all(np.isfinite(absent[3])) and np.any(absent[3] != given[3])
    assert np.array_equal(absent[[0, 1, 4, 5]], given[[0, 1, 4, 5]])
