"""Shared helpers of the GPU parity tests: build library objects from the
seeded workloads and compare against the oracle."""

import numpy as np

import compas_amd as ca
from compas_amd import workloads as W


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def context():
    try:
        return ca.get_context()
    except ca.CompasError:
        return ca.init_context(0)


def objects(wl):
    context()
    params = ca.TissueParameters.from_tissue(wl.tissue)
    if isinstance(wl.sequence, W.FispSeq):
        seq = ca.FispSequence.from_sequence(wl.sequence)
    else:
        seq = ca.pSSFPSequence.from_sequence(wl.sequence)
    if isinstance(wl.trajectory, W.CartesianTraj):
        traj = ca.CartesianTrajectory.from_trajectory(wl.trajectory)
    else:
        traj = ca.SpiralTrajectory.from_trajectory(wl.trajectory)
    return params, seq, traj


def banner(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
