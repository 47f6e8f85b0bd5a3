"""The allocation-free FISP entry point (compas_simulate_magnetization_fisp_noalloc)
against compas_simulate_magnetization_fisp, through the C-ABI.

Both run the same kernels on the same inputs, so the echoes must agree
bitwise; the FP64 oracle bounds both at the magnetisation tolerance (1e-5
relative L2).  Two calls in a row share the library's static RF-table array:
the second, with fewer voxels, TRs and slice profiles, must not see the
first one's table.
"""

import ctypes as C

import numpy as np
import pytest

import compas_amd as ca
from compas_amd import workloads as W
from gpu_helpers import banner, context, rel

pytestmark = pytest.mark.gpu

MAG_TOL = 1e-5
MAX_ENTRIES = 65536  # (slice profile, TR) pairs of the static RF table


def fisp_case(nv, nTR, N, nslices, seed):
    n = int(np.ceil(np.sqrt(nv)))
    t = W.make_tissue(W.SplitMix64(seed), n, 20.0, 40.0).slice(0, nv)
    rng = W.SplitMix64(seed + 100)
    RF = (rng.uniform(nTR, 5, 70) * np.exp(1j * rng.uniform(nTR, -0.5, 0.5))).astype(np.complex64)
    prof = np.stack([np.linspace(0.7 + 0.1 * j, 1.1, nTR) * np.exp(0.2j * j) for j in range(nslices)])
    seq = W.FispSeq(RF=RF, profiles=prof.astype(np.complex64), TR=0.009, TE=0.004, max_state=N, TI=0.02)
    return t, seq


def noalloc(params, seq, out: ca.CompasArray, capacity: int) -> None:
    ok = ca.lib().compas_simulate_magnetization_fisp_noalloc(
        ca.get_context().ptr, params.ptr, seq.ptr, C.c_void_p(out.device_pointer), C.c_int64(capacity))
    ca._check()
    assert ok == 1


def test_noalloc_matches_allocating_entry_point(oracle):
    context()
    nr0, nv0 = 3 * 96, 3000
    # one output buffer, sized for the first call and reused by the second
    out = ca.CompasArray.from_numpy(np.full((nr0, nv0), np.nan + 1j * np.nan, np.complex64))
    cases = [
        dict(nv=nv0, nTR=96, N=20, nslices=3, seed=11),  # first call: 288 table entries
        dict(nv=517, nTR=40, N=7, nslices=1, seed=12),  # second call: smaller in every dimension
    ]
    for case in cases:
        banner(f"fisp noalloc {case}")
        t, s = fisp_case(**case)
        params = ca.TissueParameters.from_tissue(t)
        seq = ca.FispSequence.from_sequence(s)
        want = ca.simulate_magnetization(params, seq).collect()
        nr, nv = want.shape
        noalloc(params, seq, out, nr0 * nv0)
        got = out.collect().reshape(-1)[: nr * nv].reshape(nr, nv)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), "not bitwise equal"
        assert rel(got, oracle.fisp(t, s, "f64")) < MAG_TOL


def test_noalloc_rejects_oversize_and_short_buffers():
    context()
    banner("fisp noalloc error paths")
    out = ca.CompasArray.from_numpy(np.zeros((4, 100), np.complex64))
    t, s = fisp_case(nv=100, nTR=4, N=3, nslices=1, seed=5)
    params = ca.TissueParameters.from_tissue(t)
    seq = ca.FispSequence.from_sequence(s)
    # capacity below [nTR][nv]: rejected on the host before any launch
    ok = ca.lib().compas_simulate_magnetization_fisp_noalloc(
        ca.get_context().ptr, params.ptr, seq.ptr, C.c_void_p(out.device_pointer), C.c_int64(399))
    assert ok == 0
    with pytest.raises(ca.CompasError, match="echo buffer too small"):
        ca._check()
    # more (slice, TR) pairs than the static table holds: rejected on the host
    big = W.FispSeq(RF=np.full(MAX_ENTRIES // 2 + 1, 10.0, np.complex64),
                    profiles=np.ones((2, MAX_ENTRIES // 2 + 1), np.complex64),
                    TR=0.009, TE=0.004, max_state=3, TI=0.02)
    bseq = ca.FispSequence.from_sequence(big)
    ok = ca.lib().compas_simulate_magnetization_fisp_noalloc(
        ca.get_context().ptr, params.ptr, bseq.ptr, C.c_void_p(out.device_pointer), C.c_int64(400))
    assert ok == 0
    with pytest.raises(ca.CompasError, match="too large for the allocation-free"):
        ca._check()
    # the buffer is untouched by the rejected calls
    assert np.array_equal(out.collect(), np.zeros((4, 100), np.complex64))
