"""Brute-force isochromat ensembles used to pin the oracle (numpy, FP64).

These are independent formulations of the same physics, not restatements of
the oracle's code:

* FISP: M isochromats with gradient dephasing angles 2*pi*j/M.  Each carries a
  real magnetisation vector, rotated by 3x3 rotation matrices built by
  scipy.spatial.transform.Rotation (rotation about the RF axis
  (cos phi, sin phi, 0) by alpha), relaxed, and dephased by exp(i*theta_j)
  each TR.  The mean transverse magnetisation is the k = 0 EPG state; with
  M > 2*nTR the ensemble has no aliasing and the EPG (untruncated) must agree.
* pSSFP: the same sequence of rotations composed with scipy rotation vectors
  (Rotation.from_rotvec), which is the library form of Rodrigues' formula.
"""

import numpy as np
from scipy.spatial.transform import Rotation


def fisp_ensemble(T1, T2, B1, B0, RF, TR, TE, TI, M=None, inversion=True, reps=1, TW=0.0):
    """Single voxel, single slice profile, wait spoiling; samples the last repetition."""
    nTR = len(RF)
    if M is None:
        M = 2 * nTR * reps + 8
    theta = 2 * np.pi * np.arange(M) / M
    m = np.zeros((M, 3))
    m[:, 2] = 1.0

    def relax(m, t):
        E1, E2 = np.exp(-t / T1), np.exp(-t / T2)
        mxy = (m[:, 0] + 1j * m[:, 1]) * E2 * np.exp(1j * 2 * np.pi * B0 * t)
        out = np.empty_like(m)
        out[:, 0], out[:, 1] = mxy.real, mxy.imag
        out[:, 2] = m[:, 2] * E1 + (1 - E1)
        return out

    echos = np.zeros(nTR, complex)
    for rep in range(reps):
        if inversion:
            m[:, 2] = -m[:, 2]
            m[:, 0:2] = 0
            m = relax(m, TI)
        for n in range(nTR):
            rf = complex(RF[n])
            alpha = np.deg2rad(abs(rf)) * B1
            phi = np.angle(rf) if abs(rf) > 0 else 0.0
            axis = np.array([np.cos(phi), np.sin(phi), 0.0])
            R = Rotation.from_rotvec(alpha * axis).as_matrix()
            m = m @ R.T
            m = relax(m, TE)
            echos[n] = np.mean(m[:, 0] + 1j * m[:, 1])
            m = relax(m, TR - TE)
            mxy = (m[:, 0] + 1j * m[:, 1]) * np.exp(1j * theta)
            m[:, 0], m[:, 1] = mxy.real, mxy.imag
        m[:, 0:2] = 0  # wait spoiling
        m = relax(m, TW)
    return echos


def pssfp_isochromat(T1, T2, B1, B0, RF, g_RF, dt, gz, zs):
    """Sum over slice positions of the transverse magnetisation, one voxel."""
    dt_ex, dt_inv, dt_pr = dt
    g_ex, _, g_pr = gz
    nTR = len(RF)
    echos = np.zeros(nTR, complex)

    def relax(m, t):
        E1, E2 = np.exp(-t / T1), np.exp(-t / T2)
        return np.array([m[0] * E2, m[1] * E2, m[2] * E1 + 1 - E1])

    def zrot(m, g, z, t):
        ang = -g * z - t * 2 * np.pi * B0
        return Rotation.from_rotvec([0.0, 0.0, ang]).apply(m)

    def excite(m, th, z):
        for gi in g_RF:
            zap = th * gi
            a = np.array([zap.real * B1, -zap.imag * B1, -g_ex * z - dt_ex * 2 * np.pi * B0])
            m = Rotation.from_rotvec(a).apply(m)
            m = relax(m, dt_ex)
        return m

    for z in zs:
        m = np.array([0.0, 0.0, np.cos(np.pi * B1)])
        m = relax(m, dt_inv)
        m = excite(m, -complex(RF[0]) / 2, z)
        m = relax(zrot(m, 2 * g_pr, z, dt_pr), dt_pr)
        for n in range(nTR):
            m = excite(m, complex(RF[n]), z)
            m = relax(zrot(m, g_pr, z, dt_pr), dt_pr)
            echos[n] += m[0] + 1j * m[1]
            m = relax(zrot(m, g_pr, z, dt_pr), dt_pr)
    return echos
