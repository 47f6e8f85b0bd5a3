"""Seeded synthetic inputs for the simulation hot path.

Generator: SplitMix64 (Steele, Lea & Flood, "Fast splittable pseudorandom
number generators", OOPSLA 2014), u = (x >> 40) * 2**-24.  Seeds are
0xC0A50000 + workload index.  Units follow the reference's Julia examples:
cm, rad/cm, s, Hz (CompasToolkit.jl/examples/common.jl:36-72).

Voxel ordering is x-fastest (v = ix + N * iy), matching the Julia examples'
``[x for x in LinRange(...), y in 1:N] |> vec``; positions are
x_i = (i - N/2) * FOV/N, which gives exact Nyquist (dk * dx = 2 pi / N).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """Counter-based SplitMix64 stream; output i depends only on seed + (i+1)*golden."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.count = 0

    def next_u64(self, n: int) -> np.ndarray:
        idx = np.arange(self.count + 1, self.count + n + 1, dtype=np.uint64)
        self.count += n
        with np.errstate(over="ignore"):
            z = self.state + idx * _GOLDEN
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        u = (self.next_u64(n) >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)
        return lo + u * (hi - lo)

    def log_uniform(self, n: int, lo: float, hi: float) -> np.ndarray:
        u = self.uniform(n)
        return np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))

    def normal(self, n: int) -> np.ndarray:
        # Box-Muller on (0, 1] uniforms
        u1 = 1.0 - self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


@dataclass
class Tissue:
    T1: np.ndarray
    T2: np.ndarray
    B1: np.ndarray
    B0: np.ndarray
    rho: np.ndarray  # complex64
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray

    @property
    def nvoxels(self) -> int:
        return int(self.T1.shape[0])

    def slice(self, begin: int, end: int) -> "Tissue":
        return Tissue(*(getattr(self, f)[begin:end] for f in ("T1", "T2", "B1", "B0", "rho", "x", "y", "z")))

    def replace(self, **kw) -> "Tissue":
        d = {f: getattr(self, f) for f in ("T1", "T2", "B1", "B0", "rho", "x", "y", "z")}
        d.update(kw)
        return Tissue(**d)


@dataclass
class FispSeq:
    RF: np.ndarray  # complex64 [nTR], degrees
    profiles: np.ndarray  # complex64 [nslices][nTR]
    TR: float
    TE: float
    max_state: int
    TI: float
    TW: float = 0.0
    undersampling: int = 1
    repetitions: int = 1
    inversion: bool = True
    wait_spoiling: bool = True

    @property
    def nreadouts(self) -> int:
        return int(self.RF.shape[0]) * self.undersampling


@dataclass
class PssfpSeq:
    RF: np.ndarray  # complex64 [nTR]
    TR: float
    gamma_dt_RF: np.ndarray  # complex64 [nRF]
    dt: tuple  # (ex, inv, pr)
    gamma_dt_GRz: tuple  # (ex, inv, pr)
    z: np.ndarray  # float32 [nz]

    @property
    def nreadouts(self) -> int:
        return int(self.RF.shape[0])


@dataclass
class CartesianTraj:
    nreadouts: int
    samples_per_readout: int
    delta_t: float
    k_start: np.ndarray  # complex64 [nr]
    delta_k: float  # real, as passed from Julia


@dataclass
class SpiralTraj:
    nreadouts: int
    samples_per_readout: int
    delta_t: float
    k_start: np.ndarray  # complex64 [nr]
    delta_k: np.ndarray  # complex64 [nr]


@dataclass
class Workload:
    name: str
    tissue: Tissue
    sequence: object
    trajectory: object
    coils: np.ndarray  # complex64 [nc][nv]
    fov: float
    extra: dict = field(default_factory=dict)


def grid_positions(n: int, fov: float):
    dx = fov / n
    coords = (np.arange(n) - n // 2) * dx
    x = np.tile(coords, n)  # x fastest
    y = np.repeat(coords, n)
    return x.astype(np.float32), y.astype(np.float32)


def make_tissue(rng: SplitMix64, n: int, fov: float, b0_range: Optional[float]) -> Tissue:
    nv = n * n
    T1 = rng.log_uniform(nv, 0.1, 4.0)
    T2 = rng.log_uniform(nv, 0.01, 1.0)
    bad = T2 >= T1
    while np.any(bad):  # redraw T2 until T2 < T1
        T2[bad] = rng.log_uniform(int(bad.sum()), 0.01, 1.0)
        bad = T2 >= T1
    B1 = rng.uniform(nv, 0.8, 1.2)
    B0 = rng.uniform(nv, -b0_range, b0_range) if b0_range else np.zeros(nv)
    mag = rng.uniform(nv, 0.5, 1.0)
    ph = rng.uniform(nv, 0.0, 2 * np.pi)
    rho = (mag * np.exp(1j * ph)).astype(np.complex64)
    x, y = grid_positions(n, fov)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return Tissue(f32(T1), f32(T2), f32(B1), f32(B0), rho, x, y, np.zeros(nv, np.float32))


def sinusoid_train(nTR: int, max_deg: float, alternate: bool = False) -> np.ndarray:
    n = np.arange(nTR)
    rf = max_deg * np.sin(np.pi * n / max(nTR - 1, 1))
    if alternate:
        rf = rf * np.where(n % 2 == 0, 1.0, -1.0)
    return rf.astype(np.complex64)


def make_fisp(nTR: int, max_deg: float, max_state: int, TR=8.7e-3, TI=20e-3) -> FispSeq:
    return FispSeq(
        RF=sinusoid_train(nTR, max_deg),
        profiles=np.ones((1, nTR), np.complex64),
        TR=TR, TE=TR / 2, max_state=max_state, TI=TI)


def make_cartesian(n_pe: int, passes: int, ns: int, fov: float, dt: float) -> CartesianTraj:
    nr = n_pe * passes
    dk = 2 * np.pi / fov
    r = np.arange(nr)
    py = (r % n_pe) - n_pe // 2
    k_start = (-ns / 2 * dk + 1j * py * dk).astype(np.complex64)
    return CartesianTraj(nr, ns, dt, k_start, float(np.float32(dk)))


def make_spiral(nr: int, ns: int, fov: float, n: int, dt: float) -> SpiralTraj:
    dx = fov / n
    phi = np.deg2rad(np.arange(nr) * 111.246)
    e = np.exp(1j * phi)
    k_start = (-(np.pi / dx) * e).astype(np.complex64)
    delta_k = ((2 * np.pi / dx) / ns * e).astype(np.complex64)
    return SpiralTraj(nr, ns, dt, k_start, delta_k)


def make_coils(n: int, fov: float, ncoils: int) -> np.ndarray:
    x, y = grid_positions(n, fov)
    x = x.astype(np.float64)
    y = y.astype(np.float64)
    sigma = 0.35 * fov
    centres = [(0.35 * fov, 0.0), (-0.35 * fov, 0.0), (0.0, 0.35 * fov), (0.0, -0.35 * fov)]
    out = []
    for j in range(ncoils):
        cx, cy = centres[j % 4]
        d2 = (x - cx) ** 2 + (y - cy) ** 2
        out.append(np.exp(-d2 / (2 * sigma ** 2)) * np.exp(1j * np.pi * j / 4))
    return np.ascontiguousarray(np.array(out), dtype=np.complex64)


def make_pssfp(nTR: int, max_deg: float, nz: int = 35, nRF: int = 25) -> PssfpSeq:
    t = -1.0 + 2.0 * np.arange(nRF) / (nRF - 1) if nRF > 1 else np.zeros(1)
    w = np.sinc(3 * t) * np.exp(-t ** 2 / (2 * 0.5 ** 2))
    g = (np.pi / 180) * w / w.sum()
    durRF, TR, TI = 1e-3, 10e-3, 20e-3
    return PssfpSeq(
        RF=sinusoid_train(nTR, max_deg, alternate=True),
        TR=TR,
        gamma_dt_RF=g.astype(np.complex64),
        dt=(durRF / nRF, TI, (TR - durRF) / 2),
        gamma_dt_GRz=(0.002 / nRF, 0.0, -0.01),
        z=np.linspace(-1, 1, nz).astype(np.float32))


def config(index: int, scale: int = 1) -> Workload:
    """Workload `index` (1..4) of the benchmark family.

    scale > 1 shrinks the grid side and sequence lengths by that factor for
    parity tests that must finish on the CPU oracle in seconds.
    """
    rng = SplitMix64(0xC0A50000 + index)
    if index == 1:
        n, fov = 64 // scale, 25.6
        tissue = make_tissue(rng, n, fov, None)
        nTR = 256 // scale
        seq = make_fisp(nTR, 60.0, 16)
        traj = make_cartesian(n, 4, 64 // scale, fov, 1e-5)
        coils = np.ones((1, n * n), np.complex64)
        return Workload("fisp-cartesian-4096vox-256tr", tissue, seq, traj, coils, fov)
    if index in (2, 4):
        n, fov = 224 // scale, 22.4
        tissue = make_tissue(rng, n, fov, 50.0)
        nTR = 1120 // scale
        seq = make_fisp(nTR, 70.0, 25)
        traj = make_cartesian(n, 5, 224 // scale, fov, 5e-6)
        coils = make_coils(n, fov, 4)
        return Workload("fisp-cartesian-50176vox-1120tr-4coil", tissue, seq, traj, coils, fov)
    if index == 3:
        n, fov = 224 // scale, 22.4
        tissue = make_tissue(rng, n, fov, 50.0)
        nTR = 1120 // scale
        seq = make_pssfp(nTR, 70.0)
        traj = make_spiral(nTR, 512 // scale, fov, n, 5e-6)
        coils = make_coils(n, fov, 4)
        nv = n * n
        v = rng.normal(6 * nv).reshape(6, nv).astype(np.complex64)
        ns = traj.samples_per_readout
        yv = (rng.normal(4 * nTR * ns) + 1j * rng.normal(4 * nTR * ns)) / np.sqrt(2)
        yv = yv.reshape(4, nTR, ns).astype(np.complex64)
        return Workload("pssfp-spiral-50176vox-1120tr-35iso", tissue, seq, traj, coils, fov, {"v": v, "y": yv})
    raise ValueError(f"unknown workload {index}")
