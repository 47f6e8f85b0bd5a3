"""The serving mix: seeded request specifications (types, voxel counts,
sequence lengths, coils, trajectories) and the resident objects they use.

Request i of the mix (seed 0xC0A50005 + i): FISP simulate + signal 40%,
pSSFP simulate + signal 30%, J^H.y 30%; voxels log-uniform 1e3..2e5; TRs
uniform 250..2000; coils uniform 1..8; Cartesian or spiral trajectory 50/50;
128 samples per readout.  Tissue maps are slices of one 448 x 448 map.
"""

from __future__ import annotations

import numpy as np

import compas_amd as ca
from compas_amd import workloads as W

SERVE_NS = 128  # samples per readout of the serving requests (Cartesian and spiral)


def serve_request_spec(i):
    """Specification of request i of the serving mix (see the module docstring)."""
    rng = W.SplitMix64(0xC0A50005 + i)
    u = rng.uniform(6)
    kind = "fisp" if u[0] < 0.4 else ("pssfp" if u[0] < 0.7 else "jhy")
    nv = int(round(np.exp(np.log(1e3) + u[1] * (np.log(2e5) - np.log(1e3)))))
    nTR = 250 + min(1750, int(u[2] * 1751))
    nc = 1 + min(7, int(u[3] * 8))
    spiral = bool(u[4] < 0.5)
    return {"id": i, "kind": kind, "nv": nv, "nTR": nTR, "nc": nc, "spiral": spiral, "offset": float(u[5])}


class ServeCatalog:
    """Resident objects of the serving run: tissue maps (slices of one
    200704-voxel map), sequences and trajectories per length, coil maps and
    residuals per request -- all uploaded before the clock starts."""

    def __init__(self):
        n = 448
        self.fov = 22.4
        self.master = W.make_tissue(W.SplitMix64(0xC0A50005), n, self.fov, 50.0)
        self.seqs, self.trajs = {}, {}

    def sequence(self, kind, nTR):
        key = (kind == "pssfp", nTR)
        if key not in self.seqs:
            if kind == "pssfp":
                self.seqs[key] = ca.pSSFPSequence.from_sequence(W.make_pssfp(nTR, 70.0))
            else:
                self.seqs[key] = ca.FispSequence.from_sequence(W.make_fisp(nTR, 70.0, 25))
        return self.seqs[key]

    def trajectory(self, spiral, nTR, n):
        key = (spiral, nTR, n)
        if key not in self.trajs:
            if spiral:
                t = W.make_spiral(nTR, SERVE_NS, self.fov, n, 5e-6)
                self.trajs[key] = ca.SpiralTrajectory.from_trajectory(t)
            else:
                t = W.make_cartesian(n, nTR // n + 1, SERVE_NS, self.fov, 5e-6)
                t = W.CartesianTraj(nTR, SERVE_NS, t.delta_t, t.k_start[:nTR], t.delta_k)
                self.trajs[key] = ca.CartesianTrajectory.from_trajectory(t)
        return self.trajs[key]

    def host_sequence(self, kind, nTR):
        return W.make_pssfp(nTR, 70.0) if kind == "pssfp" else W.make_fisp(nTR, 70.0, 25)

    def host_trajectory(self, spiral, nTR, n):
        if spiral:
            return W.make_spiral(nTR, SERVE_NS, self.fov, n, 5e-6)
        t = W.make_cartesian(n, nTR // n + 1, SERVE_NS, self.fov, 5e-6)
        return W.CartesianTraj(nTR, SERVE_NS, t.delta_t, t.k_start[:nTR], t.delta_k)

    def host_request(self, spec):
        """Host data of a request: tissue slice, sequence, trajectory, coils, y."""
        nv, nc = spec["nv"], spec["nc"]
        total = self.master.nvoxels
        off = int(spec["offset"] * (total - nv))
        t = self.master.slice(off, off + nv)
        n = int(np.ceil(np.sqrt(nv)))
        rng = W.SplitMix64(0xC0A50005 + spec["id"] + 0x9E3779B9)
        coils = W.make_coils(n, self.fov, nc)[:, :nv]
        phase = np.exp(1j * rng.uniform(nc) * 2 * np.pi)[:, None]
        y = None
        if spec["kind"] == "jhy":
            k = nc * spec["nTR"] * SERVE_NS
            y = ((rng.normal(k) + 1j * rng.normal(k)) / np.sqrt(2)).reshape(nc, spec["nTR"], SERVE_NS)
            y = y.astype(np.complex64)
        return t, n, (coils * phase).astype(np.complex64), y

    def request(self, spec):
        t, n, coils, y = self.host_request(spec)
        r = {"spec": spec, "params": ca.TissueParameters.from_tissue(t),
             "seq": self.sequence(spec["kind"], spec["nTR"]),
             "traj": self.trajectory(spec["spiral"], spec["nTR"], n),
             "coils": ca.CompasArray.from_numpy(coils)}
        if y is not None:
            r["y"] = ca.CompasArray.from_numpy(y)
        return r

    def frame(self, spec):
        """The request as a wire frame (compas_amd.wire)."""
        from compas_amd import wire

        t, n, coils, y = self.host_request(spec)
        seq = self.host_sequence(spec["kind"], spec["nTR"])
        traj = self.host_trajectory(spec["spiral"], spec["nTR"], n)
        if y is not None:
            return wire.encode_request(wire.KIND_JACOBIAN_HERMITIAN, t, seq, traj, coils, y, jacobian_method="gemm_fast")
        return wire.encode_request(wire.KIND_SIMULATE_SIGNAL, t, seq, traj, coils, signal_method="matmul_fast")


def serve_submit(server, r):
    if r["spec"]["kind"] == "jhy":
        return server.submit_jacobian_hermitian(r["params"], r["seq"], r["traj"], r["coils"], r["y"], "gemm_fast")
    return server.submit_simulate_signal(r["params"], r["seq"], r["traj"], r["coils"], "matmul_fast")


