"""Zero-copy interop: library arrays over torch memory (DLPack in), torch
tensors over library arrays (DLPack out), raw device-pointer wrapping with a
release hook, and the checks on what cannot be wrapped."""

import gc

import numpy as np
import pytest

import compas_amd as ca
from compas_amd import workloads as W
from gpu_helpers import banner, context, objects

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def test_torch_tensor_as_library_input():
    banner("interop: torch coils and vector into signal and J^H.y")
    context()
    wl = W.config(2, scale=8)
    params, seq, traj = objects(wl)
    echos = ca.simulate_magnetization(params, seq)
    want = ca.magnetization_to_signal(echos, params, traj, wl.coils).collect()
    coils_t = torch.from_numpy(wl.coils).to("cuda")
    coils = ca.CompasArray.from_dlpack(coils_t)
    assert coils.shape == wl.coils.shape and coils.dtype == np.complex64
    assert coils.device_pointer == coils_t.data_ptr()  # no copy
    del coils_t  # the array keeps the tensor's memory alive
    gc.collect()
    got = ca.magnetization_to_signal(echos, params, traj, coils).collect()
    assert np.array_equal(got, want)


def test_library_array_as_torch_tensor():
    banner("interop: library result as a torch tensor, shared memory")
    context()
    wl = W.config(2, scale=8)
    params, seq, _ = objects(wl)
    echos = ca.simulate_magnetization(params, seq)
    host = echos.collect()
    t = echos.to_torch()
    assert t.is_cuda and t.dtype == torch.complex64 and tuple(t.shape) == host.shape
    assert t.data_ptr() == echos.device_pointer
    assert np.array_equal(t.cpu().numpy(), host)
    t.mul_(2)  # a write through torch is a write to the library's array
    torch.cuda.synchronize()
    assert np.array_equal(echos.collect(), 2 * host)
    # the tensor keeps the memory after the library handle is gone
    del echos
    gc.collect()
    assert np.array_equal(t.cpu().numpy(), 2 * host)


def test_stream_ordering_both_ways():
    banner("interop: torch side-stream producer and consumer")
    context()
    n = 1 << 20
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        x = torch.full((n,), 3.0, device="cuda", dtype=torch.float32)
        for _ in range(20):
            x = x * 1.5 - 1.0  # pending work on the side stream
        a = ca.CompasArray.from_dlpack(x)  # the library's stream waits for it
    want = np.full(n, 3.0, np.float64)
    for _ in range(20):
        want = want * 1.5 - 1.0
    assert np.allclose(a.collect(), want.astype(np.float32), rtol=1e-6)
    with torch.cuda.stream(s):
        y = torch.from_dlpack(a)  # the side stream waits for the library
        z = (y + 1.0).sum()
    s.synchronize()
    assert np.isclose(z.item(), float(np.sum(want.astype(np.float32) + 1.0)), rtol=1e-5)


def test_raw_pointer_wrap_and_release_hook():
    banner("interop: raw device pointer with a release hook")
    context()
    t = torch.arange(64, dtype=torch.float32, device="cuda")
    released = []
    a = ca.CompasArray.from_device(t.data_ptr(), np.float32, (8, 8), release=lambda: released.append(1))
    assert np.array_equal(a.collect(), np.arange(64, dtype=np.float32).reshape(8, 8))
    b = a.to_torch()  # a second owner of the same buffer
    del a
    gc.collect()
    assert released == []
    del b
    gc.collect()
    assert released == [1]


def test_unwrappable_memory_is_rejected():
    banner("interop: misaligned, host, wrong dtype, strided")
    context()
    t = torch.zeros(1024, dtype=torch.complex64, device="cuda")
    with pytest.raises(ca.CompasError, match="aligned"):
        ca.CompasArray.from_dlpack(t[1:])
    with pytest.raises(ca.CompasError, match="float32 or complex64"):
        ca.CompasArray.from_dlpack(torch.zeros(16, dtype=torch.float64, device="cuda"))
    with pytest.raises(ca.CompasError, match="ROCm device memory"):
        ca.CompasArray.from_dlpack(torch.zeros(16, dtype=torch.float32))
    with pytest.raises(ca.CompasError, match="contiguous"):
        ca.CompasArray.from_dlpack(torch.zeros(32, 32, dtype=torch.float32, device="cuda").t())
    host = np.zeros(64, np.float32)
    with pytest.raises(ca.CompasError, match="device memory"):
        ca.CompasArray.from_device(host.ctypes.data & ~255, np.float32, (16,))
    # a rejected tensor stays usable by its owner
    t2 = torch.ones(16, dtype=torch.float64, device="cuda")
    with pytest.raises(ca.CompasError):
        ca.CompasArray.from_dlpack(t2)
    assert float(t2.sum()) == 16.0
