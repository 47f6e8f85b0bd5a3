"""Wire format of the request server, on the CPU: the Python encoder against
the library's parser (compas_wire_describe reads a frame without a GPU), the
checks on malformed frames, and the Python frame codec."""

import struct

import numpy as np
import pytest

import compas_amd as ca
from compas_amd import wire
from compas_amd import workloads as W


def small(kind=0, spiral=False, pssfp=False, nv=50, seed=4):
    rng = W.SplitMix64(seed)
    n = int(np.ceil(np.sqrt(nv)))
    t = W.make_tissue(rng, n, 20.0, 30.0).slice(0, nv)
    if pssfp:
        seq = W.make_pssfp(12, 40.0, nz=3, nRF=5)
    else:
        RF = (np.linspace(5, 60, 12) * np.exp(0.1j * np.arange(12))).astype(np.complex64)
        seq = W.FispSeq(RF=RF, profiles=np.ones((2, 12), np.complex64), TR=0.009, TE=0.004, max_state=8, TI=0.02,
                        TW=0.1, undersampling=1, repetitions=2)
    nr = seq.nreadouts
    if spiral:
        traj = W.make_spiral(nr, 7, 20.0, 3, 5e-6)
    else:
        traj = W.CartesianTraj(nr, 7, 5e-6, (np.arange(nr) * 0.1 + 0.2j).astype(np.complex64), 0.3)
    coils = (np.arange(2 * nv).reshape(2, nv) * (0.01 + 0.02j)).astype(np.complex64)
    res = None
    if kind == 1:
        res = (np.arange(2 * nr * 7).reshape(2, nr, 7) * (0.5 - 0.25j)).astype(np.complex64)
    return t, seq, traj, coils, res


@pytest.mark.parametrize("kind,spiral,pssfp", [(0, False, False), (1, True, True), (1, False, False),
                                               (0, True, True)])
def test_library_reads_python_frames(kind, spiral, pssfp):
    t, seq, traj, coils, res = small(kind, spiral, pssfp)
    frame = wire.encode_request(kind, t, seq, traj, coils, res, signal_method="direct", jacobian_method="gemm",
                                return_echos=bool(kind))
    d = wire.describe(frame)
    assert d["return_echos"] == kind
    assert d["kind"] == ("simulate_signal" if kind == 0 else "jacobian_hermitian")
    assert d["sequence"] == ("pssfp" if pssfp else "fisp")
    assert d["trajectory"] == ("spiral" if spiral else "cartesian")
    assert (d["nvoxels"], d["nTR"], d["nreadouts"], d["samples"], d["ncoils"]) == (50, 12, 12, 7, 2)
    assert d["signal_method"] == ca.SIGNAL_METHODS["direct"]
    assert d["jacobian_method"] == ca.JACOBIAN_METHODS["gemm"]
    assert d["fields"] == ["T1", "T2", "B1", "B0", "rho_x", "rho_y", "x", "y", "z"]
    # payload sums as the library read them (FP64 sums of the f32 words)
    assert d["sum_T1"] == pytest.approx(float(np.sum(t.T1.astype(np.float32), dtype=np.float64)), rel=1e-12)
    rf = np.asarray(seq.RF, np.complex64).view(np.float32)
    assert d["sum_RF"] == pytest.approx(float(np.sum(rf, dtype=np.float64)), rel=1e-12)
    assert d["sum_coils"] == pytest.approx(float(np.sum(coils.view(np.float32), dtype=np.float64)), rel=1e-12)
    if kind == 1:
        assert d["sum_residual"] == pytest.approx(float(np.sum(res.view(np.float32), dtype=np.float64)), rel=1e-12)
    if pssfp:
        assert (d["nRF"], d["nz"]) == (5, 3)
    else:
        assert (d["nslices"], d["max_state"]) == (2, 8)


def test_optional_tissue_fields():
    t, seq, traj, coils, _ = small()
    frame = wire.encode_request(0, t, seq, traj, coils, fields=("T1", "T2"))
    assert wire.describe(frame)["fields"] == ["T1", "T2", "rho_x", "rho_y", "x", "y"]


def test_python_codec_round_trip():
    entries = {"a": np.arange(5, dtype=np.float32), "b": (np.arange(6) * (1 + 2j)).astype(np.complex64).reshape(2, 3),
               "c": np.array(7, np.int32), "d": np.array(2.5), "e": np.frombuffer(b"hello", np.uint8),
               "f": np.arange(24, dtype=np.int64).reshape(2, 3, 4), "g": np.zeros((0, 4), np.float32)}
    frame = wire.build_frame(wire.RESULT_MAGIC, 1, entries)
    assert len(frame) % 8 == 0
    kind, out = wire.parse_frame(frame, wire.RESULT_MAGIC)
    assert kind == 1 and list(out) == list(entries)
    for k, v in entries.items():
        assert out[k].dtype == v.dtype and out[k].shape == v.shape and np.array_equal(out[k], v)


def _entry_offset(frame, name):
    """Byte offset of an entry's header in a frame."""
    off = 16
    while True:
        raw, code, rank, *dims = struct.unpack_from("<16sII3q", frame, off)
        if raw.split(b"\0", 1)[0].decode() == name:
            return off
        n = int(np.prod(dims[:rank], dtype=np.int64)) * [4, 8, 4, 8, 1, 8][code]
        off += 48 + n + (-n % 8)


def _reject(frame, match):
    with pytest.raises(ca.CompasError, match=match):
        wire.describe(frame)


def test_malformed_frames_are_rejected():
    t, seq, traj, coils, res = small(kind=1)
    good = wire.encode_request(1, t, seq, traj, coils, res)
    wire.describe(good)
    _reject(b"", "shorter than its header")
    _reject(b"XXXX" + good[4:], "magic")
    _reject(good[:4] + struct.pack("<H", 2) + good[6:], "version")
    _reject(good + b"\0" * 8, "trailing")
    for cut in (17, 48, 64, 100, len(good) // 2, len(good) - 8):
        _reject(good[:cut], "past the end")
    # an entry size far beyond the buffer: rejected before any read
    at = _entry_offset(good, "coils")
    for dim, match in ((1 << 30, "past the end"), (1 << 41, "invalid size"), (-5, "invalid size")):
        bad = bytearray(good)
        struct.pack_into("<q", bad, at + 24, dim)
        _reject(bytes(bad), match)
    bad = bytearray(good)
    struct.pack_into("<qq", bad, at + 24, 1 << 30, 1 << 30)  # product beyond 2^40 elements
    _reject(bytes(bad), "too large")
    bad = bytearray(good)
    struct.pack_into("<I", bad, at + 20, 4)  # rank 4
    _reject(bytes(bad), "rank above 3")
    bad = bytearray(good)
    struct.pack_into("<I", bad, at + 16, 9)  # dtype 9
    _reject(bytes(bad), "unknown dtype")
    bad = bytearray(good)
    bad[16:32] = b"x" * 16  # name without its NUL
    _reject(bytes(bad), "not terminated")


def test_inconsistent_requests_are_rejected():
    t, seq, traj, coils, res = small(kind=1)

    def frame(**over):
        kind, e = wire.parse_frame(wire.encode_request(1, t, seq, traj, coils, res), wire.REQUEST_MAGIC)
        for k, v in over.items():
            if v is None:
                del e[k]
            else:
                e[k] = v
        return wire.build_frame(wire.REQUEST_MAGIC, over.pop("_kind", kind), e)

    _reject(frame(T2=None), "no entry `T2`")
    _reject(frame(TR=np.array(1, np.int32)), "`TR` must be f32")
    _reject(frame(coils=coils[:, :10]), r"coils must be \[ncoils\]\[nvoxels\]")
    _reject(frame(residual=res[:, :, :3]), "residual must be")
    _reject(frame(T2=t.T2[:10].astype(np.float32)), "another length")
    _reject(frame(sequence_type=np.array(3, np.int32)), "sequence_type")
    _reject(frame(signal_method=np.array(9, np.int32)), "signal_method")
    _reject(frame(sliceprofiles=np.ones((2, 5), np.complex64)), "sliceprofiles")
    _reject(frame(delta_k=np.ones(3, np.float32)), "delta_k")
    kind_bad = bytearray(frame())
    struct.pack_into("<H", kind_bad, 6, 7)
    _reject(bytes(kind_bad), "request kind")


def test_random_corruption_never_reads_outside_the_frame():
    # flipped bytes anywhere in a frame: the parser accepts or rejects, never crashes
    t, seq, traj, coils, res = small(kind=1, spiral=True, pssfp=True)
    good = wire.encode_request(1, t, seq, traj, coils, res)
    rng = np.random.default_rng(11)
    rejected = 0
    for _ in range(400):
        b = bytearray(good)
        for i in rng.integers(0, len(b), size=int(rng.integers(1, 4))):
            b[int(i)] = int(rng.integers(0, 256))
        try:
            wire.describe(bytes(b))
        except ca.CompasError:
            rejected += 1
    assert rejected > 0


class _StubServer:
    """Host-only stand-in for the library server (the front end's contract:
    submit_wire, wait_done, result_wire, release): 'runs' a request by
    doubling its `x` entry, finishing after a seeded random delay so that
    requests complete out of order."""

    def __init__(self):
        import threading

        self.lock = threading.Lock()
        self.next = 0
        self.results = {}
        self.released = set()
        self.rng = np.random.default_rng(3)

    def submit_wire(self, frame):
        kind, e = wire.parse_frame(frame, wire.REQUEST_MAGIC)
        if "x" not in e:
            raise ValueError("request has no entry `x`")
        with self.lock:
            t = self.next
            self.next += 1
            self.results[t] = (2 * e["x"], float(self.rng.uniform(0, 0.02)))
        return t

    def wait_done(self, t):
        import time

        time.sleep(self.results[t][1])
        return True

    def result_wire(self, t):
        return wire.build_frame(wire.RESULT_MAGIC, 0, {
            "ticket": np.array(t, np.int64), "t_arrival": np.array(0.0), "t_start": np.array(0.0),
            "t_done": np.array(0.0), "signal": self.results[t][0]})

    def release(self, t):
        with self.lock:
            self.released.add(t)


def test_tcp_front_end_order_and_errors_on_the_host():
    import threading

    server = _StubServer()
    fe = wire.TcpFrontEnd(server)
    got = {}

    def client(k):
        c = wire.TcpClient(fe.address)
        frames = [wire.build_frame(wire.REQUEST_MAGIC, 0, {"x": np.full(3, 10 * k + i, np.complex64)})
                  for i in range(15)]
        for f in frames[:7]:
            c.send(f)
        c.send(b"garbage")  # rejected by the parser
        c.send(wire.build_frame(wire.REQUEST_MAGIC, 0, {"y": np.zeros(1, np.float32)}))  # rejected by the server
        for f in frames[7:]:
            c.send(f)
        got[k] = [c.recv() for _ in range(17)]
        c.close()

    threads = [threading.Thread(target=client, args=(k,)) for k in range(3)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=60)
    fe.close()
    for k in range(3):
        res = got[k]
        assert len(res) == 17
        ok = [r for j, r in enumerate(res) if j not in (7, 8)]
        # results in sending order on each connection, whatever order they finished in
        assert [complex(r["signal"][0]) for r in ok] == [2 * (10 * k + i) for i in range(15)]
        assert not res[7]["ok"] and "shorter than its header" in res[7]["error"]
        assert not res[8]["ok"] and "no entry `x`" in res[8]["error"]
    assert len(server.released) == 45
