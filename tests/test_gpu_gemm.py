"""GEMM utility (reference include/compas/utils/gemm.h:7-71, src/utils/gemm.cu)
through the C-ABI against numpy FP64 matrix products on ragged shapes.

Tolerances (relative L2 of the whole result): Pedantic (FP32 FMA) 1e-6,
Regular (BF16 hi + lo, three MFMA products) 1e-5, Fast (BF16, one product)
1e-2; BF16 operands (exact products, FP32 accumulation) 1e-6 against the
FP64 product of the same BF16 values.  The converters are exact (BF16: round
to nearest even, compared bit for bit).
"""

import numpy as np
import pytest

import compas_amd as ca
from gpu_helpers import banner, context, rel

pytestmark = pytest.mark.gpu

TOL = {"pedantic": 1e-6, "regular": 1e-5, "fast": 1e-2}
SHAPES = [(1, 1, 1), (5, 3, 7), (129, 65, 33), (200, 300, 1000), (257, 130, 4099), (1120, 224, 2048)]


def rng(seed):
    return np.random.default_rng(seed)


@pytest.mark.parametrize("method", list(TOL))
@pytest.mark.parametrize("m,n,k", SHAPES)
def test_real_gemm(m, n, k, method):
    banner(f"gemm real {m}x{n}x{k} {method}")
    context()
    g = rng(m * 7 + n * 3 + k)
    L = g.standard_normal((m, k)).astype(np.float32)
    R = g.standard_normal((n, k)).astype(np.float32)
    C0 = g.standard_normal((m, n)).astype(np.float32)
    alpha, beta = 0.75, -1.25
    res = ca.CompasArray.from_numpy(C0)
    ca.compute_gemm(res, ca.CompasArray.from_numpy(L), ca.CompasArray.from_numpy(R), alpha, beta, method)
    want = alpha * (L.astype(np.float64) @ R.astype(np.float64).T) + beta * C0.astype(np.float64)
    err = rel(res.collect(), want)
    print(f"gemm {m}x{n}x{k} {method}: rel-L2 {err:.2e}")
    assert err < TOL[method]


@pytest.mark.parametrize("method", list(TOL))
def test_beta_zero_does_not_read_result(method):
    banner(f"gemm beta=0 {method}")
    context()
    g = rng(1)
    L = g.standard_normal((70, 90)).astype(np.float32)
    R = g.standard_normal((40, 90)).astype(np.float32)
    res = ca.CompasArray.from_numpy(np.full((70, 40), np.nan, np.float32))
    ca.compute_gemm(res, ca.CompasArray.from_numpy(L), ca.CompasArray.from_numpy(R), 2.0, 0.0, method)
    got = res.collect()
    assert np.all(np.isfinite(got))
    assert rel(got, 2.0 * (L.astype(np.float64) @ R.astype(np.float64).T)) < TOL[method]


@pytest.mark.parametrize("method", list(TOL))
@pytest.mark.parametrize("m,n,k", [(3, 5, 9), (129, 65, 33), (300, 200, 3001)])
def test_real_gemm_bf16(m, n, k, method):
    banner(f"gemm bf16 {m}x{n}x{k} {method}")
    context()
    g = rng(m + n + k)
    Lb = ca.to_bf16_bits(g.standard_normal((m, k)))
    Rb = ca.to_bf16_bits(g.standard_normal((n, k)))
    res = ca.CompasArray.from_numpy(np.zeros((m, n), np.float32))
    ca.compute_gemm(res, ca.CompasArray.from_bf16_bits(Lb), ca.CompasArray.from_bf16_bits(Rb), 1.0, 0.0, method)
    want = ca.from_bf16_bits(Lb).astype(np.float64) @ ca.from_bf16_bits(Rb).astype(np.float64).T
    assert rel(res.collect(), want) < 1e-6


def planar(z):
    return np.stack([z.real, z.imag]).astype(np.float32)


@pytest.mark.parametrize("method", list(TOL))
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (129, 65, 33), (224, 120, 2049)])
def test_complex_gemm(m, n, k, method):
    banner(f"gemm complex {m}x{n}x{k} {method}")
    context()
    g = rng(m * n + k)
    L = (g.standard_normal((m, k)) + 1j * g.standard_normal((m, k))).astype(np.complex64)
    R = (g.standard_normal((n, k)) + 1j * g.standard_normal((n, k))).astype(np.complex64)
    C0 = (g.standard_normal((m, n)) + 1j * g.standard_normal((m, n))).astype(np.complex64)
    alpha, beta = -0.5, 0.25
    res = ca.CompasArray.from_numpy(planar(C0))
    ca.compute_complex_gemm(res, ca.CompasArray.from_numpy(planar(L)), ca.CompasArray.from_numpy(planar(R)),
                            alpha, beta, method)
    got = res.collect()
    want = alpha * (L.astype(np.complex128) @ R.astype(np.complex128).T) + beta * C0.astype(np.complex128)
    err = rel(got[0] + 1j * got[1].astype(np.float64), want)
    print(f"complex gemm {m}x{n}x{k} {method}: rel-L2 {err:.2e}")
    assert err < TOL[method]


def test_complex_gemm_bf16_through_converters():
    """The reference's benchmark path (benchmarks/compute_gemm.cu): complex
    operands converted to planar BF16, the planar product, converted back."""
    banner("gemm complex bf16 via converters")
    context()
    g = rng(5)
    m, n, k = 150, 96, 1500
    L = (g.standard_normal((m, k)) + 1j * g.standard_normal((m, k))).astype(np.complex64)
    R = (g.standard_normal((n, k)) + 1j * g.standard_normal((n, k))).astype(np.complex64)
    Lp = ca.CompasArray.from_bf16_bits(np.zeros((2, m, k), np.uint16))
    Rp = ca.CompasArray.from_bf16_bits(np.zeros((2, n, k), np.uint16))
    ca.convert_complex_to_planar(Lp, ca.CompasArray.from_numpy(L))
    ca.convert_complex_to_planar(Rp, ca.CompasArray.from_numpy(R))
    lb, rb = Lp.collect(), Rp.collect()
    assert np.array_equal(lb, ca.to_bf16_bits(planar(L)))
    assert np.array_equal(rb, ca.to_bf16_bits(planar(R)))
    res = ca.CompasArray.from_numpy(np.zeros((2, m, n), np.float32))
    ca.compute_complex_gemm(res, Lp, Rp, 1.0, 0.0, "regular")
    out = ca.CompasArray.from_numpy(np.zeros((m, n), np.complex64))
    ca.convert_planar_to_complex(out, res)
    lf, rf = ca.from_bf16_bits(lb).astype(np.float64), ca.from_bf16_bits(rb).astype(np.float64)
    want = (lf[0] + 1j * lf[1]) @ (rf[0] + 1j * rf[1]).T
    assert rel(out.collect(), want) < 1e-6


def test_converters_are_exact():
    banner("gemm converters")
    context()
    g = rng(9)
    z = (g.standard_normal((33, 47)) + 1j * g.standard_normal((33, 47))).astype(np.complex64)
    p = ca.CompasArray.from_numpy(np.zeros((2, 33, 47), np.float32))
    ca.convert_complex_to_planar(p, ca.CompasArray.from_numpy(z))
    assert np.array_equal(p.collect(), planar(z))
    back = ca.CompasArray.from_numpy(np.zeros((33, 47), np.complex64))
    ca.convert_planar_to_complex(back, p)
    assert np.array_equal(back.collect(), z)


def test_gemm_errors():
    banner("gemm errors")
    context()
    a = ca.CompasArray.from_numpy(np.zeros((4, 5), np.float32))
    b = ca.CompasArray.from_numpy(np.zeros((3, 6), np.float32))  # k differs
    r = ca.CompasArray.from_numpy(np.zeros((4, 3), np.float32))
    with pytest.raises(ca.CompasError, match="assertion failed"):
        ca.compute_gemm(r, a, b)
    b = ca.CompasArray.from_numpy(np.zeros((3, 5), np.float32))
    with pytest.raises(ca.CompasError, match="GemmComputeMethod"):
        ca.lib().compas_compute_gemm(ca.get_context().ptr, r.ptr, a.ptr, b.ptr, 1.0, 0.0, 7)
        ca._check()
    with pytest.raises(ca.CompasError, match="invalid array"):  # complex result for a real product
        ca.compute_gemm(ca.CompasArray.from_numpy(np.zeros((4, 3), np.complex64)), a, b)
    with pytest.raises(ca.CompasError, match="assertion failed"):  # planar axis must be 2
        ca.compute_complex_gemm(ca.CompasArray.from_numpy(np.zeros((3, 4, 3), np.float32)),
                                ca.CompasArray.from_numpy(np.zeros((2, 4, 5), np.float32)),
                                ca.CompasArray.from_numpy(np.zeros((2, 3, 5), np.float32)))
