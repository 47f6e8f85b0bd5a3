"""Wire format of the request server, and a TCP front end for it.

A request crosses a process or network boundary as one self-describing binary
frame of named, typed arrays; its result comes back as one (the layout is
documented in csrc/include/compas/wire.h, which the library parses).  The
reference has no wire or on-disk format: arrays reach it only through
compas_make_array_* copies (julia-bindings/src/bindings.cpp:50-107).

    frame  = header entry*
    header = u32 magic ("CMPQ" | "CMPR"), u16 version (1), u16 kind,
             u32 count, u32 0
    entry  = char name[16], u32 dtype, u32 rank, i64 dims[3],
             data (prod(dims[:rank]) elements, zero-padded to 8 bytes)

On TCP each frame travels as a u64 little-endian byte count followed by the
frame; the server answers every request frame with its result frame, in the
order the requests arrived on that connection, while it runs requests from
all connections concurrently.
"""

from __future__ import annotations

import ctypes as C
import socket
import struct
import threading
from collections import deque
from typing import Dict, Optional, Tuple

import numpy as np

REQUEST_MAGIC = b"CMPQ"
RESULT_MAGIC = b"CMPR"
VERSION = 1
KIND_SIMULATE_SIGNAL = 0
KIND_JACOBIAN_HERMITIAN = 1

_DTYPES = {0: np.dtype("<f4"), 1: np.dtype("<c8"), 2: np.dtype("<i4"), 3: np.dtype("<f8"), 4: np.dtype("u1"),
           5: np.dtype("<i8")}
_CODES = {np.dtype(v).str: k for k, v in _DTYPES.items()}
_ENTRY = struct.Struct("<16sII3q")
_HEADER = struct.Struct("<4sHHII")


class FrameError(ValueError):
    pass


def _entry(name: str, a) -> bytes:
    a = np.asarray(a)
    code = _CODES.get(a.dtype.newbyteorder("<").str if a.dtype.byteorder == ">" else a.dtype.str)
    if code is None:
        raise FrameError(f"entry {name}: unsupported dtype {a.dtype}")
    if a.ndim > 3 or len(name) > 15:
        raise FrameError(f"entry {name}: rank above 3 or name too long")
    dims = list(a.shape) + [1] * (3 - a.ndim)
    data = np.ascontiguousarray(a, dtype=_DTYPES[code]).tobytes()
    return _ENTRY.pack(name.encode(), code, a.ndim, *dims) + data + b"\0" * (-len(data) % 8)


def build_frame(magic: bytes, kind: int, entries: Dict[str, np.ndarray]) -> bytes:
    body = b"".join(_entry(k, v) for k, v in entries.items())
    return _HEADER.pack(magic, VERSION, kind, len(entries), 0) + body


def parse_frame(frame: bytes, magic: bytes) -> Tuple[int, Dict[str, np.ndarray]]:
    """(kind, {name: array}) of a frame; raises FrameError when malformed."""
    mv = memoryview(frame)
    if len(mv) < _HEADER.size:
        raise FrameError("shorter than its header")
    m, version, kind, count, _ = _HEADER.unpack_from(mv, 0)
    if m != magic:
        raise FrameError("wrong magic")
    if version != VERSION:
        raise FrameError(f"unsupported version {version}")
    off, out = _HEADER.size, {}
    for _ in range(count):
        if len(mv) - off < _ENTRY.size:
            raise FrameError("entry header past the end")
        raw, code, rank, d0, d1, d2 = _ENTRY.unpack_from(mv, off)
        off += _ENTRY.size
        name = raw.split(b"\0", 1)[0].decode()
        if code not in _DTYPES or rank > 3:
            raise FrameError(f"entry {name}: bad dtype or rank")
        shape = (d0, d1, d2)[:rank]
        if any(d < 0 for d in shape):
            raise FrameError(f"entry {name}: negative size")
        n = int(np.prod(shape, dtype=np.int64)) * _DTYPES[code].itemsize
        if off + n > len(mv):
            raise FrameError(f"entry {name} runs past the end")
        out[name] = np.frombuffer(mv[off:off + n], dtype=_DTYPES[code]).reshape(shape).copy()
        off += n + (-n % 8)
    if off != len(mv):
        raise FrameError("trailing bytes after the last entry")
    return kind, out


def encode_request(kind: int, tissue, sequence, trajectory, coils, residual=None, signal_method="matmul_fast",
                   jacobian_method="gemm_fast", fields=("T1", "T2", "B1", "B0", "rho", "x", "y", "z"),
                   return_echos: bool = False) -> bytes:
    """Request frame from host data: tissue (workloads.Tissue), sequence
    (FispSeq or PssfpSeq), trajectory (CartesianTraj or SpiralTraj), coils
    [nc][nv] complex, residual [nc][nr][ns] (J^H.y).  `fields` lists the
    tissue fields to send (B1, B0 and z may be left out); with return_echos
    the result frame also carries the echoes [nr][nv]."""
    from . import JACOBIAN_METHODS, SIGNAL_METHODS
    from . import workloads as W

    f32, c64, i32 = np.float32, np.complex64, np.int32
    e: Dict[str, np.ndarray] = {}
    fisp = isinstance(sequence, W.FispSeq)
    spiral = isinstance(trajectory, W.SpiralTraj)
    e["sequence_type"] = np.array(0 if fisp else 1, i32)
    e["trajectory_type"] = np.array(1 if spiral else 0, i32)
    sm = SIGNAL_METHODS[signal_method] if isinstance(signal_method, str) else int(signal_method)
    jm = JACOBIAN_METHODS[jacobian_method] if isinstance(jacobian_method, str) else int(jacobian_method)
    e["signal_method"] = np.array(sm, i32)
    e["jacobian_method"] = np.array(jm, i32)
    e["return_echos"] = np.array(int(bool(return_echos)), i32)
    for name in ("T1", "T2", "B1", "B0"):
        if name in fields:
            e[name] = np.asarray(getattr(tissue, name), f32)
    e["rho_x"] = np.asarray(tissue.rho.real, f32)
    e["rho_y"] = np.asarray(tissue.rho.imag, f32)
    e["x"] = np.asarray(tissue.x, f32)
    e["y"] = np.asarray(tissue.y, f32)
    if "z" in fields:
        e["z"] = np.asarray(tissue.z, f32)
    e["RF"] = np.asarray(sequence.RF, c64)
    e["TR"] = np.array(sequence.TR, f32)
    if fisp:
        e["sliceprofiles"] = np.asarray(sequence.profiles, c64)
        for k in ("TE", "TW", "TI"):
            e[k] = np.array(getattr(sequence, k), f32)
        e["max_state"] = np.array(sequence.max_state, i32)
        e["undersampling"] = np.array(sequence.undersampling, i32)
        e["repetitions"] = np.array(sequence.repetitions, i32)
        e["inversion"] = np.array(int(bool(sequence.inversion)), i32)
        e["wait_spoiling"] = np.array(int(bool(sequence.wait_spoiling)), i32)
    else:
        e["gamma_dt_RF"] = np.asarray(sequence.gamma_dt_RF, c64)
        e["dt"] = np.asarray(sequence.dt, f32)
        e["gamma_dt_GRz"] = np.asarray(sequence.gamma_dt_GRz, f32)
        e["slice_z"] = np.asarray(sequence.z, f32)
    e["samples"] = np.array(trajectory.samples_per_readout, i32)
    e["delta_t"] = np.array(trajectory.delta_t, f32)
    e["k_start"] = np.asarray(trajectory.k_start, c64)
    e["delta_k"] = np.asarray(trajectory.delta_k, c64 if spiral else f32)
    e["coils"] = np.asarray(coils, c64)
    if kind == KIND_JACOBIAN_HERMITIAN:
        if residual is None:
            raise FrameError("a J^H.y request needs the residual")
        e["residual"] = np.asarray(residual, c64)
    return build_frame(REQUEST_MAGIC, kind, e)


def decode_result(frame: bytes) -> dict:
    """{"ok", "ticket", "t_arrival", "t_start", "t_done", and "echos" with
    "signal" or "jhy", or "error"} of a result frame."""
    kind, e = parse_frame(frame, RESULT_MAGIC)
    out = {"ok": kind == 0, "ticket": int(e["ticket"]),
           **{k: float(e[k]) for k in ("t_arrival", "t_start", "t_done")}}
    if kind == 0:
        out.update({k: e[k] for k in ("echos", "signal", "jhy") if k in e})
    else:
        out["error"] = bytes(e["error"]).decode(errors="replace")
    return out


def describe(frame: bytes) -> dict:
    """The library's reading of a request frame (no GPU): kind, sizes, sums."""
    import json

    from . import CompasError, lib

    buf = C.create_string_buffer(4096)
    n = lib().compas_wire_describe(frame, len(frame), buf, len(buf))
    err = lib().compas_last_error()
    if n < 0 or err:
        lib().compas_clear_error()
        raise CompasError((err or b"malformed frame").decode(errors="replace"))
    return json.loads(buf.value.decode())


# ---- TCP front end

_LEN = struct.Struct("<Q")
MAX_FRAME = 1 << 36


def _recv_exact(sock, n) -> Optional[bytes]:
    chunks, got = [], 0
    while got < n:
        b = sock.recv(min(n - got, 1 << 22))
        if not b:
            return None
        chunks.append(b)
        got += len(b)
    return b"".join(chunks)


def send_frame(sock, frame: bytes) -> None:
    sock.sendall(_LEN.pack(len(frame)) + frame)


def recv_frame(sock) -> Optional[bytes]:
    h = _recv_exact(sock, _LEN.size)
    if h is None:
        return None
    (n,) = _LEN.unpack(h)
    if n > MAX_FRAME:
        raise FrameError("frame too large")
    return _recv_exact(sock, n)


class TcpFrontEnd:
    """Serves a library Server over TCP: per connection, a reader submits each
    request frame as it arrives and a writer returns the result frames in
    arrival order; a frame the library rejects gets an error result frame."""

    def __init__(self, server, host: str = "127.0.0.1", port: int = 0):
        self.server = server
        self.sock = socket.socket(socket.AF_INET, socket.SOCK_STREAM)
        self.sock.setsockopt(socket.SOL_SOCKET, socket.SO_REUSEADDR, 1)
        self.sock.bind((host, port))
        self.sock.listen()
        self.address = self.sock.getsockname()
        self._stop = threading.Event()
        self._threads = []
        self._accept = threading.Thread(target=self._accept_loop, daemon=True)
        self._accept.start()

    def _accept_loop(self):
        self.sock.settimeout(0.2)
        while not self._stop.is_set():
            try:
                conn, _ = self.sock.accept()
            except socket.timeout:
                continue
            except OSError:
                break
            t = threading.Thread(target=self._serve, args=(conn,), daemon=True)
            t.start()
            self._threads.append(t)

    def _serve(self, conn):
        queue, cv, done = deque(), threading.Condition(), [False]

        def writer():
            while True:
                with cv:
                    while not queue and not done[0]:
                        cv.wait()
                    if not queue:
                        return
                    item = queue.popleft()
                try:
                    if isinstance(item, bytes):  # a rejected frame: its error result
                        send_frame(conn, item)
                        continue
                    self.server.wait_done(item)
                    send_frame(conn, self.server.result_wire(item))
                    self.server.release(item)
                except OSError:
                    return

        w = threading.Thread(target=writer, daemon=True)
        w.start()
        try:
            while not self._stop.is_set():
                frame = recv_frame(conn)
                if frame is None:
                    break
                try:
                    item = self.server.submit_wire(frame)
                except Exception as exc:  # malformed or inconsistent request
                    item = build_frame(RESULT_MAGIC, 1, {
                        "ticket": np.array(-1, np.int64), "t_arrival": np.array(0.0), "t_start": np.array(0.0),
                        "t_done": np.array(0.0), "error": np.frombuffer(str(exc).encode(), np.uint8)})
                with cv:
                    queue.append(item)
                    cv.notify()
        except (OSError, FrameError):
            pass
        finally:
            with cv:
                done[0] = True
                cv.notify()
            w.join()
            conn.close()

    def close(self):
        self._stop.set()
        self.sock.close()
        self._accept.join()
        for t in self._threads:
            t.join(timeout=30)


class TcpClient:
    """Client of TcpFrontEnd: send() request frames, recv() result frames in
    the order sent (requests may be pipelined)."""

    def __init__(self, address):
        self.sock = socket.create_connection(address)

    def send(self, frame: bytes):
        send_frame(self.sock, frame)

    def recv(self) -> dict:
        frame = recv_frame(self.sock)
        if frame is None:
            raise FrameError("connection closed")
        return decode_result(frame)

    def close(self):
        self.sock.close()
