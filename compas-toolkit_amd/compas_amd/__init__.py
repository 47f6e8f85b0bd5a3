"""Python host interface of libcompas_amd.so, mirroring CompasToolkit.jl.

The names, argument order and meaning follow the Julia package the
reference ships (CompasToolkit.jl/src/CompasToolkit.jl:98-1036), so tests
read like the reference's Julia examples.  Arrays are numpy arrays in C
order: echos [nreadouts][nvoxels], coils [ncoils][nvoxels], signal
[ncoils][nreadouts][nsamples] (Julia passes the same memory with reversed
dimensions).

Everything goes through the C-ABI declared in include/compas_amd.h; there is
no fallback: if the library is missing or no GPU is present the calls raise.
"""

from __future__ import annotations

import contextlib
import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "libcompas_amd.so")

SIGNAL_METHODS = {"naive": 0, "direct": 1, "matmul_pedantic": 2, "matmul": 3, "matmul_fast": 4, "matmul_low": 5}
JACOBIAN_METHODS = {"naive": 0, "direct": 1, "gemm": 2, "gemm_fast": 3, "gemm_low": 4}

# TissueParameterField (tissue_view.cuh:13-26)
T1, T2, B1, B0, RHO_X, RHO_Y, X, Y, Z = range(9)


class CompasError(RuntimeError):
    pass


_lib = None
_vp = C.c_void_p
_f = C.c_float
_i = C.c_int
_i32 = C.c_int32
_i64p = C.POINTER(C.c_int64)

_SIGNATURES = {
    "compas_version": (C.c_char_p, []),
    "compas_make_context": (_vp, []),
    "compas_copy_context_for_device": (_vp, [_vp, _i]),
    "compas_destroy_context": (None, [_vp]),
    "compas_destroy_object": (None, [_vp]),
    "compas_synchronize": (None, [_vp]),
    "compas_make_array_float": (_vp, [_vp, _vp, _i, _i64p]),
    "compas_read_array_float": (None, [_vp, _vp, _vp, C.c_int64]),
    "compas_make_array_complex": (_vp, [_vp, _vp, _i, _i64p]),
    "compas_read_array_complex": (None, [_vp, _vp, _vp, C.c_int64]),
    "compas_destroy_array": (None, [_vp]),
    "compas_make_tissue_parameters": (_vp, [_vp, _i] + [_vp] * 9),
    "compas_make_fisp_sequence": (_vp, [_vp, _vp, _f, _f, _f, _i, _f, _i, _i]),
    "compas_make_pssfp_sequence": (_vp, [_vp, _f, _vp, _f, _f, _f, _f, _f, _f, _vp]),
    "compas_make_cartesian_trajectory": (_vp, [_i, _i, _f, _vp, _f]),
    "compas_make_spiral_trajectory": (_vp, [_i, _i, _f, _vp, _vp]),
    "compas_simulate_magnetization_fisp": (_vp, [_vp, _vp, _vp]),
    "compas_simulate_magnetization_pssfp": (_vp, [_vp, _vp, _vp]),
    "compas_simulate_magnetization_fisp_noalloc": (_i, [_vp, _vp, _vp, _vp, C.c_int64]),
    "compas_magnetization_to_signal_cartesian": (_vp, [_vp, _i, _vp, _vp, _vp, _vp, _i32]),
    "compas_magnetization_to_signal_spiral": (_vp, [_vp, _i, _vp, _vp, _vp, _vp]),
    "compas_compute_jacobian": (_vp, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32]),
    "compas_compute_jacobian_hermitian": (_vp, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32]),
    "phase_encoding": (_vp, [_vp, _vp, _vp, _vp]),
    "compas_compute_residual": (_vp, [_vp, _vp, _vp, C.POINTER(C.c_float)]),
    "compas_simulate_magnetization_derivative_pssfp": (_vp, [_vp, _i, _vp, _vp, _f, _vp]),
    "compas_simulate_magnetization_derivative_fisp": (_vp, [_vp, _i, _vp, _vp, _f, _vp]),
    # extensions
    "compas_last_error": (C.c_char_p, []),
    "compas_clear_error": (None, []),
    "compas_make_fisp_sequence_ex": (_vp, [_vp, _vp, _f, _f, _f, _i, _f, _i, _i, _i, _i]),
    "compas_array_info": (_i, [_vp, C.POINTER(C.c_int), _i64p]),
    "compas_make_array_bf16": (_vp, [_vp, _vp, _i, _i64p]),
    "compas_read_array_bf16": (None, [_vp, _vp, _vp, C.c_int64]),
    "compas_compute_gemm": (_i, [_vp, _vp, _vp, _vp, _f, _f, _i]),
    "compas_compute_complex_gemm": (_i, [_vp, _vp, _vp, _vp, _f, _f, _i]),
    "compas_convert_complex_to_planar": (_i, [_vp, _vp, _vp]),
    "compas_convert_planar_to_complex": (_i, [_vp, _vp, _vp]),
    "compas_magnetization_to_signal_method": (_vp, [_vp, _vp, _vp, _vp, _vp, _i]),
    "compas_compute_jacobian_method": (_vp, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i]),
    "compas_compute_jacobian_hermitian_method": (_vp, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i]),
    "compas_set_kernel_timing": (None, [_vp, _i]),
    "compas_read_kernel_timings": (_i, [_vp, C.c_char_p, _i]),
    "compas_trim_memory": (None, [_vp]),
    "compas_bandwidth_probe": (_f, [_vp, C.c_int64]),
    "compas_set_contraction_variant": (None, [_i]),
    "compas_contraction_variant_count": (_i, []),
    "compas_contraction_tuning_report": (_i, [C.c_char_p, _i]),
    "compas_set_contraction_precision": (None, [_i]),
    "compas_compute_jacobian_spiral": (_vp, [_vp] * 10),
    "compas_compute_jacobian_hermitian_spiral": (_vp, [_vp] * 10),
    "compas_compute_jacobian_spiral_method": (_vp, [_vp] * 10 + [_i]),
    "compas_compute_jacobian_hermitian_spiral_method": (_vp, [_vp] * 10 + [_i]),
    "compas_simulate_magnetization_tangent_fisp": (_vp, [_vp, _i, _vp, _vp]),
    "compas_simulate_magnetization_tangent_pssfp": (_vp, [_vp, _i, _vp, _vp]),
    "compas_simulate_magnetization_tangents_fisp": (_i, [_vp, C.POINTER(C.c_int32), _i32, _i32, _vp, _vp,
                                                         C.POINTER(_vp)]),
    "compas_simulate_magnetization_tangents_pssfp": (_i, [_vp, C.POINTER(C.c_int32), _i32, _i32, _vp, _vp,
                                                          C.POINTER(_vp)]),
    "compas_context_stream": (_vp, [_vp]),
    "compas_context_wait_stream": (None, [_vp, _vp]),
    "compas_stream_wait_context": (None, [_vp, _vp]),
    "compas_wrap_device_array": (_vp, [_vp, _vp, _i32, _i32, _i64p, _vp, _vp]),
    "compas_array_device_pointer": (_vp, [_vp]),
    "compas_array_device": (_i32, [_vp]),
    "compas_array_from_dlpack": (_vp, [_vp, _vp]),
    "compas_array_to_dlpack": (_vp, [_vp]),
    "compas_dlpack_delete": (None, [_vp]),
    "compas_graph_begin": (None, [_vp]),
    "compas_graph_end": (_vp, [_vp]),
    "compas_graph_abort": (None, [_vp]),
    "compas_graph_launch": (None, [_vp]),
    "compas_graph_destroy": (None, [_vp]),
    "compas_write_array_complex": (None, [_vp, _vp, _vp, C.c_int64]),
    "compas_copy_array": (None, [_vp, _vp, _vp]),
    "compas_server_create": (_vp, [_vp, _i, C.c_int64]),
    "compas_server_destroy": (None, [_vp]),
    "compas_server_submit_simulate_signal": (C.c_int64, [_vp, _vp, _vp, _vp, _vp, _i]),
    "compas_server_submit_jacobian_hermitian": (C.c_int64, [_vp, _vp, _vp, _vp, _vp, _vp, _i]),
    "compas_server_wait": (_i, [_vp, C.c_int64, C.c_double, C.POINTER(C.c_double)]),
    "compas_server_result": (_vp, [_vp, C.c_int64, _i]),
    "compas_server_release": (None, [_vp, C.c_int64]),
    "compas_server_batch_stats": (_i, [_vp, _i64p]),
    "compas_server_now": (C.c_double, [_vp]),
    "compas_server_submit_wire": (C.c_int64, [_vp, C.c_char_p, C.c_int64]),
    "compas_server_result_wire": (C.c_int64, [_vp, C.c_int64, _vp, C.c_int64]),
    "compas_wire_describe": (C.c_int64, [C.c_char_p, C.c_int64, C.c_char_p, C.c_int64]),
    "compas_server_set_kernel_timing": (None, [_vp, _i]),
    "compas_server_read_kernel_timings": (_i, [_vp, C.c_char_p, _i]),
}


def set_contraction_variant(variant: int):
    """Force an MFMA tile variant (>= 0) or restore autotuning (-1)."""
    lib().compas_set_contraction_variant(int(variant))


CONTRACTION_PRECISIONS = {"fp32": 0, "bf16x3": 1, "fp16x3": 2, "bf16x1": 3, "fp16x1": 4}


def set_contraction_precision(precision):
    """Override the contraction precision every signal/Jacobian method selects
    (a name from CONTRACTION_PRECISIONS), or clear the override (None)."""
    p = -1 if precision is None else CONTRACTION_PRECISIONS[precision]
    lib().compas_set_contraction_precision(p)
    _check()


def contraction_variant_count() -> int:
    return int(lib().compas_contraction_variant_count())


def contraction_tuning_report() -> str:
    buf = C.create_string_buffer(1 << 16)
    lib().compas_contraction_tuning_report(buf, len(buf))
    return buf.value.decode()


def library_path() -> str:
    return os.environ.get("COMPAS_AMD_LIBRARY", LIB_PATH)


def lib():
    """Load libcompas_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        path = library_path()
        if not os.path.exists(path):
            raise CompasError(f"{path} not found: build it with `python compas-toolkit_amd/build.py`")
        handle = C.CDLL(path, mode=C.RTLD_GLOBAL)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def _check(result=None):
    err = lib().compas_last_error()
    if err:
        lib().compas_clear_error()
        raise CompasError(err.decode(errors="replace"))
    return result


def _parse_timings(read) -> dict:
    buf = C.create_string_buffer(1 << 20)  # one call: reading resets the counters
    n = read(buf)
    _check()
    if n > len(buf):
        raise CompasError("kernel timing report truncated")
    out = {}
    for line in buf.value.decode().splitlines():
        name, count, total = line.rsplit(" ", 2)
        out[name] = (int(count), float(total))
    return out


def version() -> str:
    return lib().compas_version().decode()


class Context:
    """A compas context bound to one device (CompasToolkit.jl:56-91)."""

    def __init__(self, ptr=None, device: int = 0):
        if ptr is None:
            ptr = _check(lib().compas_make_context())
            if device != 0:
                base = ptr
                ptr = _check(lib().compas_copy_context_for_device(base, device))
                lib().compas_destroy_context(base)
        self.ptr = ptr

    def with_device(self, device: int) -> "Context":
        return Context(_check(lib().compas_copy_context_for_device(self.ptr, device)))

    def synchronize(self):
        lib().compas_synchronize(self.ptr)
        _check()

    @property
    def stream(self) -> int:
        """The HIP stream the library's calls on this context run on."""
        s = lib().compas_context_stream(self.ptr)
        _check()
        return int(s or 0)

    def wait_for_stream(self, stream: int) -> None:
        """Library calls queued after this wait for the work queued so far on
        `stream` (an event; the host does not block)."""
        lib().compas_context_wait_stream(self.ptr, C.c_void_p(int(stream)))
        _check()

    def make_stream_wait(self, stream: int) -> None:
        """Work queued on `stream` after this waits for the library calls
        queued so far on this context."""
        lib().compas_stream_wait_context(self.ptr, C.c_void_p(int(stream)))
        _check()

    def set_kernel_timing(self, enabled: bool):
        lib().compas_set_kernel_timing(self.ptr, int(enabled))
        _check()

    def read_kernel_timings(self) -> dict:
        """Per-kernel (launches, total ms) since the last read; resets the counters."""
        return _parse_timings(lambda buf: lib().compas_read_kernel_timings(self.ptr, buf, len(buf)))

    def trim_memory(self):
        lib().compas_trim_memory(self.ptr)

    def bandwidth_probe(self, nbytes: int) -> float:
        """Milliseconds of one streaming copy of `nbytes` (kernel stream_copy)."""
        ms = lib().compas_bandwidth_probe(self.ptr, int(nbytes))
        _check()
        return float(ms)

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.compas_destroy_context(self.ptr)
            self.ptr = None


_CTX: Optional[Context] = None


def init_context(device: int = 0) -> Context:
    """CompasToolkit.init_context: create and install the current context."""
    global _CTX
    _CTX = Context(device=device)
    return _CTX


def get_context() -> Context:
    if _CTX is None:
        raise CompasError("COMPAS toolkit has not been initialized, use `init_context` before usage")
    return _CTX


class CompasArray:
    """Device array owned by the library (CompasToolkit.jl:140-256)."""

    def __init__(self, ptr, context: Context):
        self.ptr = ptr
        self.context = context
        dtype = C.c_int(0)
        sizes = (C.c_int64 * 3)()
        rank = _check(lib().compas_array_info(ptr, C.byref(dtype), sizes))
        # bfloat16 arrays are exchanged as their raw bits (uint16)
        self.bf16 = dtype.value == 2
        self.dtype = {1: np.complex64, 2: np.uint16}.get(dtype.value, np.float32)
        self.shape = tuple(int(sizes[i]) for i in range(rank))

    @classmethod
    def from_numpy(cls, a, context: Optional[Context] = None) -> "CompasArray":
        ctx = context or get_context()
        a = np.asarray(a)
        if np.iscomplexobj(a):
            a = np.ascontiguousarray(a, dtype=np.complex64)
            fn = lib().compas_make_array_complex
        else:
            a = np.ascontiguousarray(a, dtype=np.float32)
            fn = lib().compas_make_array_float
        sizes = (C.c_int64 * a.ndim)(*a.shape)
        ptr = _check(fn(ctx.ptr, a.ctypes.data_as(C.c_void_p), a.ndim, sizes))
        return cls(ptr, ctx)

    @classmethod
    def from_bf16_bits(cls, bits, context: Optional[Context] = None) -> "CompasArray":
        """A bfloat16 array from raw bf16 bit patterns (uint16, see to_bf16_bits)."""
        ctx = context or get_context()
        b = np.ascontiguousarray(bits, dtype=np.uint16)
        sizes = (C.c_int64 * b.ndim)(*b.shape)
        ptr = _check(lib().compas_make_array_bf16(ctx.ptr, b.ctypes.data_as(C.c_void_p), b.ndim, sizes))
        return cls(ptr, ctx)

    def write(self, a) -> None:
        """Overwrite the contents with a complex array of the same size (blocking upload)."""
        a = np.ascontiguousarray(np.asarray(a, dtype=np.complex64))
        if a.size != int(np.prod(self.shape)) or self.dtype != np.complex64:
            raise CompasError("write needs a complex array of the same size")
        lib().compas_write_array_complex(self.context.ptr, self.ptr, a.ctypes.data_as(C.c_void_p), a.size)
        _check()

    def copy_from(self, src: "CompasArray") -> None:
        """Stream-ordered device copy from an array of the same type and size."""
        lib().compas_copy_array(self.context.ptr, self.ptr, src.ptr)
        _check()

    def collect(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        n = out.size
        if self.dtype == np.complex64:
            lib().compas_read_array_complex(self.context.ptr, self.ptr, out.ctypes.data_as(C.c_void_p), n)
        elif self.bf16:
            lib().compas_read_array_bf16(self.context.ptr, self.ptr, out.ctypes.data_as(C.c_void_p), n)
        else:
            lib().compas_read_array_float(self.context.ptr, self.ptr, out.ctypes.data_as(C.c_void_p), n)
        _check()
        return out

    # ---- zero-copy interop

    @classmethod
    def from_device(cls, device_ptr: int, dtype, shape, release=None,
                    context: Optional[Context] = None) -> "CompasArray":
        """An array over device memory owned elsewhere (no copy): 256-byte
        aligned, row-major, on the context's GPU.  `release()` is called once
        the last array over it is gone."""
        ctx = context or get_context()
        code = 1 if np.dtype(dtype) == np.complex64 else 0 if np.dtype(dtype) == np.float32 else None
        if code is None:
            raise CompasError("dtype must be float32 or complex64")
        sizes = (C.c_int64 * len(shape))(*shape)
        cb = None
        if release is not None:
            key = [None]

            def hook(_user, key=key):
                try:
                    release()
                finally:
                    _RELEASE_HOOKS.pop(id(key), None)

            cb = _RELEASE_FN(hook)
            key[0] = cb
            _RELEASE_HOOKS[id(key)] = key  # kept alive until the library calls it
        ptr = lib().compas_wrap_device_array(ctx.ptr, C.c_void_p(device_ptr), code, len(shape), sizes,
                                             C.cast(cb, C.c_void_p) if cb else None, None)
        if not ptr and cb is not None:
            _RELEASE_HOOKS.pop(id(key), None)
        return _wrap_in(ptr, ctx)

    @classmethod
    def from_dlpack(cls, obj, context: Optional[Context] = None) -> "CompasArray":
        """Zero-copy array over a DLPack producer's memory (a torch tensor on
        the GPU, float32 or complex64, contiguous, 256-byte aligned).  The
        producer's pending work is ordered before the library's stream."""
        ctx = context or get_context()
        if hasattr(obj, "__dlpack__"):
            if hasattr(obj, "__dlpack_device__") and int(obj.__dlpack_device__()[0]) != 10:  # kDLROCM
                raise CompasError("DLPack tensor is not ROCm device memory")
            stream = lib().compas_context_stream(ctx.ptr)
            _check()
            capsule = obj.__dlpack__(stream=int(stream or 0))
        else:
            capsule = obj
        if not _capsule_is_valid(capsule, _NAME_DLTENSOR):
            raise CompasError("expected an unconsumed DLPack capsule (dltensor)")
        managed = _capsule_get_pointer(capsule, _NAME_DLTENSOR)
        ptr = lib().compas_array_from_dlpack(ctx.ptr, C.c_void_p(managed))
        _check()
        if not ptr:
            raise CompasError("library returned no array")
        _capsule_set_name(capsule, _NAME_USED)  # the array owns the tensor now
        return cls(ptr, ctx)

    def __dlpack__(self, *, stream=None, max_version=None, dl_device=None, copy=None):
        """DLPack export (legacy capsule, kDLROCM).  `stream`: the consumer's
        stream, made to wait for the library's queued work (None: the host
        waits; -1: no ordering)."""
        if copy:
            raise BufferError("CompasArray exports its memory without copying")
        if dl_device is not None and tuple(dl_device) != self.__dlpack_device__():
            raise BufferError("CompasArray lives on another device")
        if stream is None:
            self.context.synchronize()
        elif stream != -1:
            lib().compas_stream_wait_context(self.context.ptr, C.c_void_p(int(stream)))
            _check()
        managed = lib().compas_array_to_dlpack(self.ptr)
        _check()
        return _capsule_new(managed, _NAME_DLTENSOR, _DLPACK_DTOR)

    def __dlpack_device__(self):
        dev = lib().compas_array_device(self.ptr)
        _check()
        return (10, int(dev))  # kDLROCM

    def to_torch(self):
        """Zero-copy torch tensor over this array's memory (shares it)."""
        import torch

        return torch.from_dlpack(self)

    @property
    def device_pointer(self) -> int:
        p = lib().compas_array_device_pointer(self.ptr)
        _check()
        return int(p or 0)

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.compas_destroy_array(self.ptr)
            self.ptr = None

    def __repr__(self):
        return f"CompasArray({np.dtype(self.dtype).name}, {self.shape})"


# PyCapsule helpers for DLPack (ctypes over the CPython API).  A capsule keeps
# the name pointer it is given, so the names live in module globals.
_NAME_DLTENSOR = b"dltensor"
_NAME_USED = b"used_dltensor"
_RELEASE_FN = C.CFUNCTYPE(None, C.c_void_p)
_RELEASE_HOOKS: dict = {}
_capsule_new_fn = C.pythonapi.PyCapsule_New
_capsule_new_fn.restype = C.py_object
_capsule_new_fn.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
_capsule_ptr_fn = C.pythonapi.PyCapsule_GetPointer
_capsule_ptr_fn.restype = C.c_void_p
_capsule_ptr_fn.argtypes = [C.py_object, C.c_char_p]
_capsule_valid_fn = C.pythonapi.PyCapsule_IsValid
_capsule_valid_fn.restype = C.c_int
_capsule_valid_fn.argtypes = [C.py_object, C.c_char_p]
_capsule_setname_fn = C.pythonapi.PyCapsule_SetName
_capsule_setname_fn.restype = C.c_int
_capsule_setname_fn.argtypes = [C.py_object, C.c_char_p]
_capsule_valid_raw = C.PYFUNCTYPE(C.c_int, C.c_void_p, C.c_char_p)(("PyCapsule_IsValid", C.pythonapi))
_capsule_ptr_raw = C.PYFUNCTYPE(C.c_void_p, C.c_void_p, C.c_char_p)(("PyCapsule_GetPointer", C.pythonapi))


@C.CFUNCTYPE(None, C.c_void_p)
def _DLPACK_DTOR(capsule):
    # an exported capsule nobody consumed: free the tensor it holds
    if _capsule_valid_raw(capsule, _NAME_DLTENSOR) and _lib is not None:
        _lib.compas_dlpack_delete(_capsule_ptr_raw(capsule, _NAME_DLTENSOR))


def _capsule_new(ptr, name, dtor):
    return _capsule_new_fn(C.c_void_p(ptr), name, C.cast(dtor, C.c_void_p))


def _capsule_get_pointer(capsule, name):
    return _capsule_ptr_fn(capsule, name)


def _capsule_is_valid(capsule, name):
    return bool(_capsule_valid_fn(capsule, name))


def _capsule_set_name(capsule, name):
    if _capsule_setname_fn(capsule, name) != 0:
        raise CompasError("could not mark the DLPack capsule as consumed")


def _wrap_in(ptr, ctx) -> "CompasArray":
    _check()
    if not ptr:
        raise CompasError("library returned no array")
    return CompasArray(ptr, ctx)


def as_array(a, dtype=None, context: Optional[Context] = None) -> CompasArray:
    if isinstance(a, CompasArray):
        return a
    if dtype is not None:
        a = np.asarray(a, dtype=dtype)
    return CompasArray.from_numpy(a, context)


class _Object:
    def __init__(self, ptr, keep=()):
        self.ptr = ptr
        self._keep = keep

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.compas_destroy_object(self.ptr)
            self.ptr = None


class TissueParameters(_Object):
    """CompasToolkit.jl:437-505; every field is given (z defaults to zeros)."""

    def __init__(self, nvoxels, T1, T2, B1, B0, rho_x, rho_y, x, y, z=None):
        ctx = get_context()
        fields = [T1, T2, B1, B0, rho_x, rho_y, x, y, np.zeros(nvoxels) if z is None else z]
        arrs = [np.ascontiguousarray(np.broadcast_to(np.asarray(f, dtype=np.float32), (nvoxels,))) for f in fields]
        ptr = _check(lib().compas_make_tissue_parameters(ctx.ptr, int(nvoxels), *[a.ctypes.data_as(C.c_void_p) for a in arrs]))
        super().__init__(ptr)
        self.nvoxels = int(nvoxels)

    @classmethod
    def from_tissue(cls, t) -> "TissueParameters":
        return cls(t.nvoxels, t.T1, t.T2, t.B1, t.B0, t.rho.real, t.rho.imag, t.x, t.y, t.z)


class FispSequence(_Object):
    """CompasToolkit.jl:542-575 (slice_profiles in C order: [nslices][nTR])."""

    def __init__(self, RF_train, slice_profiles, TR, TE, max_state, TI, undersampling_factor=1, repetitions=1,
                 TW=0.0, inversion_prepulse=True, wait_spoiling=True):
        rf = as_array(np.asarray(RF_train, dtype=np.complex64))
        prof = as_array(np.asarray(slice_profiles, dtype=np.complex64).reshape(-1, rf.shape[0]))
        args = (rf.ptr, prof.ptr, float(TR), float(TE), float(TW), int(max_state), float(TI),
                int(undersampling_factor), int(repetitions))
        if inversion_prepulse and wait_spoiling:
            ptr = _check(lib().compas_make_fisp_sequence(*args))  # the reference's entry point
        else:
            ptr = _check(lib().compas_make_fisp_sequence_ex(*args, int(bool(inversion_prepulse)),
                                                            int(bool(wait_spoiling))))
        super().__init__(ptr, keep=(rf, prof))
        self.nreadouts = rf.shape[0] * int(undersampling_factor)

    @classmethod
    def from_sequence(cls, s) -> "FispSequence":
        return cls(s.RF, s.profiles, s.TR, s.TE, s.max_state, s.TI, s.undersampling, s.repetitions, s.TW,
                   s.inversion, s.wait_spoiling)


class pSSFPSequence(_Object):
    """CompasToolkit.jl:594-631.  Like the Julia wrapper (:614), z is converted
    to complex64 before the call; the library takes its real parts (a complex z
    with a nonzero imaginary part is an error).  z_dtype=np.float32 passes a
    float array instead, as C++ callers of make_pssfp_sequence do."""

    def __init__(self, RF_train, TR, gamma_dt_RF, dt, gamma_dt_GRz, z, z_dtype=np.complex64):
        rf = as_array(np.asarray(RF_train, dtype=np.complex64))
        g = as_array(np.asarray(gamma_dt_RF, dtype=np.complex64))
        zz = as_array(np.asarray(z, dtype=z_dtype))
        ptr = _check(lib().compas_make_pssfp_sequence(rf.ptr, float(TR), g.ptr, *[float(t) for t in dt],
                                                      *[float(t) for t in gamma_dt_GRz], zz.ptr))
        super().__init__(ptr, keep=(rf, g, zz))
        self.nreadouts = rf.shape[0]

    @classmethod
    def from_sequence(cls, s) -> "pSSFPSequence":
        return cls(s.RF, s.TR, s.gamma_dt_RF, s.dt, s.gamma_dt_GRz, s.z)


class CartesianTrajectory(_Object):
    """CompasToolkit.jl:280-320 (real delta_k)."""

    def __init__(self, nreadouts, samples_per_readout, delta_t, k_start, delta_k):
        ks = as_array(np.asarray(k_start, dtype=np.complex64))
        ptr = _check(lib().compas_make_cartesian_trajectory(int(nreadouts), int(samples_per_readout),
                                                            float(delta_t), ks.ptr, float(np.real(delta_k))))
        super().__init__(ptr, keep=(ks,))
        self.nreadouts = int(nreadouts)
        self.samples_per_readout = int(samples_per_readout)

    @classmethod
    def from_trajectory(cls, t) -> "CartesianTrajectory":
        return cls(t.nreadouts, t.samples_per_readout, t.delta_t, t.k_start, t.delta_k)


class SpiralTrajectory(_Object):
    """CompasToolkit.jl:333-375."""

    def __init__(self, nreadouts, samples_per_readout, delta_t, k_start, delta_k):
        ks = as_array(np.asarray(k_start, dtype=np.complex64))
        dk = as_array(np.asarray(delta_k, dtype=np.complex64))
        ptr = _check(lib().compas_make_spiral_trajectory(int(nreadouts), int(samples_per_readout), float(delta_t),
                                                         ks.ptr, dk.ptr))
        super().__init__(ptr, keep=(ks, dk))
        self.nreadouts = int(nreadouts)
        self.samples_per_readout = int(samples_per_readout)

    @classmethod
    def from_trajectory(cls, t) -> "SpiralTrajectory":
        return cls(t.nreadouts, t.samples_per_readout, t.delta_t, t.k_start, t.delta_k)


def _wrap(ptr) -> CompasArray:
    _check()
    if not ptr:
        raise CompasError("library returned no array")
    return CompasArray(ptr, get_context())


def simulate_magnetization(parameters: TissueParameters, sequence) -> CompasArray:
    ctx = get_context()
    if isinstance(sequence, FispSequence):
        return _wrap(lib().compas_simulate_magnetization_fisp(ctx.ptr, parameters.ptr, sequence.ptr))
    if isinstance(sequence, pSSFPSequence):
        return _wrap(lib().compas_simulate_magnetization_pssfp(ctx.ptr, parameters.ptr, sequence.ptr))
    raise TypeError("sequence must be a FispSequence or pSSFPSequence")


def simulate_magnetization_derivative(field, echos, parameters, sequence, delta=1e-4) -> CompasArray:
    ctx = get_context()
    e = as_array(echos, np.complex64)
    fn = (lib().compas_simulate_magnetization_derivative_fisp if isinstance(sequence, FispSequence)
          else lib().compas_simulate_magnetization_derivative_pssfp)
    return _wrap(fn(ctx.ptr, int(field), e.ptr, parameters.ptr, float(delta), sequence.ptr))


def simulate_magnetization_tangent(field, parameters, sequence) -> CompasArray:
    """Forward-mode derivative of the echoes w.r.t. T1, T2, B1 or B0: the exact
    derivative of the discrete recursion (no finite-difference step), computed
    beside the FISP EPG or pSSFP isochromat recursion in one pass."""
    ctx = get_context()
    if isinstance(sequence, FispSequence):
        fn = lib().compas_simulate_magnetization_tangent_fisp
    elif isinstance(sequence, pSSFPSequence):
        fn = lib().compas_simulate_magnetization_tangent_pssfp
    else:
        raise TypeError("expected a FispSequence or pSSFPSequence")
    return _wrap(fn(ctx.ptr, int(field), parameters.ptr, sequence.ptr))


def simulate_magnetization_tangents(fields, parameters, sequence, echos: bool = False):
    """Forward-mode derivatives w.r.t. several distinct fields (T1, T2, B1,
    B0) in one call: a list of arrays in the order of `fields`, preceded by
    the echoes when `echos` is true.  For pSSFP, T1 + T2 and all four fields
    (with or without the echoes) run as one pass that shares the isochromat
    recursion among its tangents."""
    ctx = get_context()
    if isinstance(sequence, FispSequence):
        fn = lib().compas_simulate_magnetization_tangents_fisp
    elif isinstance(sequence, pSSFPSequence):
        fn = lib().compas_simulate_magnetization_tangents_pssfp
    else:
        raise TypeError("expected a FispSequence or pSSFPSequence")
    f = (C.c_int32 * max(1, len(fields)))(*[int(x) for x in fields])
    out = (_vp * (len(fields) + 1))()
    n = fn(ctx.ptr, f, len(fields), 1 if echos else 0, parameters.ptr, sequence.ptr, out)
    _check()
    if n != len(fields) + (1 if echos else 0):
        raise CompasError("library returned an unexpected number of arrays")
    return [CompasArray(out[i], ctx) for i in range(n)]


def simulate_magnetization_derivatives(echos, parameters, sequence, delta=1e-4):
    return {"T1": simulate_magnetization_derivative(T1, echos, parameters, sequence, delta),
            "T2": simulate_magnetization_derivative(T2, echos, parameters, sequence, delta)}


def magnetization_to_signal(echos, parameters, trajectory, coils, low_precision=False, method=None) -> CompasArray:
    ctx = get_context()
    e = as_array(echos, np.complex64)
    c = as_array(coils, np.complex64)
    if method is not None:
        m = SIGNAL_METHODS[method] if isinstance(method, str) else int(method)
        return _wrap(lib().compas_magnetization_to_signal_method(ctx.ptr, e.ptr, parameters.ptr, c.ptr,
                                                                 trajectory.ptr, m))
    nc = c.shape[0]
    if isinstance(trajectory, CartesianTrajectory):
        return _wrap(lib().compas_magnetization_to_signal_cartesian(ctx.ptr, nc, e.ptr, parameters.ptr, c.ptr,
                                                                    trajectory.ptr, int(bool(low_precision))))
    return _wrap(lib().compas_magnetization_to_signal_spiral(ctx.ptr, nc, e.ptr, parameters.ptr, c.ptr,
                                                             trajectory.ptr))


def compute_jacobian(echos, dechos, parameters, trajectory, coils, v, low_precision=False, method=None):
    ctx = get_context()
    e = as_array(echos, np.complex64)
    d1 = as_array(dechos["T1"] if isinstance(dechos, dict) else dechos[0], np.complex64)
    d2 = as_array(dechos["T2"] if isinstance(dechos, dict) else dechos[1], np.complex64)
    c = as_array(coils, np.complex64)
    vv = as_array(v, np.complex64)
    if method is not None:
        m = JACOBIAN_METHODS[method] if isinstance(method, str) else int(method)
        return _wrap(lib().compas_compute_jacobian_method(ctx.ptr, e.ptr, d1.ptr, d2.ptr, parameters.ptr, c.ptr,
                                                          trajectory.ptr, vv.ptr, m))
    return _wrap(lib().compas_compute_jacobian(ctx.ptr, c.shape[0], e.ptr, d1.ptr, d2.ptr, parameters.ptr, c.ptr,
                                               trajectory.ptr, vv.ptr, int(bool(low_precision))))


def compute_jacobian_hermitian(echos, dechos, parameters, trajectory, coils, y, low_precision=False, method=None):
    ctx = get_context()
    e = as_array(echos, np.complex64)
    d1 = as_array(dechos["T1"] if isinstance(dechos, dict) else dechos[0], np.complex64)
    d2 = as_array(dechos["T2"] if isinstance(dechos, dict) else dechos[1], np.complex64)
    c = as_array(coils, np.complex64)
    yy = as_array(y, np.complex64)
    if method is not None:
        m = JACOBIAN_METHODS[method] if isinstance(method, str) else int(method)
        return _wrap(lib().compas_compute_jacobian_hermitian_method(ctx.ptr, e.ptr, d1.ptr, d2.ptr, parameters.ptr,
                                                                    trajectory.ptr, c.ptr, yy.ptr, m))
    return _wrap(lib().compas_compute_jacobian_hermitian(ctx.ptr, c.shape[0], e.ptr, d1.ptr, d2.ptr, parameters.ptr,
                                                         trajectory.ptr, c.ptr, yy.ptr, int(bool(low_precision))))


SPIRAL_FIELDS = ("T1", "T2", "B1", "B0")


def _spiral_derivs(dechos):
    if isinstance(dechos, dict):
        dechos = [dechos[f] for f in SPIRAL_FIELDS]
    return [as_array(d, np.complex64) for d in dechos]


def compute_jacobian_spiral(echos, dechos, parameters, trajectory, coils, v, method=None) -> CompasArray:
    """J.v on a spiral trajectory, six parameters (T1, T2, B1, B0, rho_x, rho_y).
    dechos: derivatives of the echoes for T1, T2, B1, B0 (dict or sequence);
    v: [6][nv] complex.  Returns [nc][nr][ns].  method: a JACOBIAN_METHODS key
    (default gemm_fast, as the C-ABI entry point)."""
    ctx = get_context()
    e = as_array(echos, np.complex64)
    d = _spiral_derivs(dechos)
    c = as_array(coils, np.complex64)
    vv = as_array(v, np.complex64)
    args = (ctx.ptr, e.ptr, *[x.ptr for x in d], parameters.ptr, trajectory.ptr, c.ptr, vv.ptr)
    if method is None:
        return _wrap(lib().compas_compute_jacobian_spiral(*args))
    m = JACOBIAN_METHODS[method] if isinstance(method, str) else int(method)
    return _wrap(lib().compas_compute_jacobian_spiral_method(*args, m))


def compute_jacobian_hermitian_spiral(echos, dechos, parameters, trajectory, coils, y, method=None) -> CompasArray:
    """J^H.y on a spiral trajectory, six parameters; y: [nc][nr][ns].  Returns
    [6][nv].  method: a JACOBIAN_METHODS key (default gemm_fast: MFMA)."""
    ctx = get_context()
    e = as_array(echos, np.complex64)
    d = _spiral_derivs(dechos)
    c = as_array(coils, np.complex64)
    yy = as_array(y, np.complex64)
    args = (ctx.ptr, e.ptr, *[x.ptr for x in d], parameters.ptr, trajectory.ptr, c.ptr, yy.ptr)
    if method is None:
        return _wrap(lib().compas_compute_jacobian_hermitian_spiral(*args))
    m = JACOBIAN_METHODS[method] if isinstance(method, str) else int(method)
    return _wrap(lib().compas_compute_jacobian_hermitian_spiral_method(*args, m))


class Graph:
    """A captured sequence of library calls (see `capture`); launch() replays it
    on the context's stream, rewriting the arrays the captured calls returned."""

    def __init__(self, ctx: "Context"):
        self.context = ctx
        self.ptr = None

    def launch(self) -> None:
        if not self.ptr:
            raise CompasError("graph is not captured")
        lib().compas_graph_launch(self.ptr)
        _check()

    def close(self) -> None:
        if self.ptr:
            lib().compas_graph_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@contextlib.contextmanager
def capture():
    """Record the library calls of the block into a Graph instead of running them:

        with capture() as g:
            echos = simulate_magnetization(p, s)
            signal = magnetization_to_signal(echos, p, traj, coils)
        g.launch()          # runs both; signal.collect() reads the result

    The block must not read arrays back or upload new ones; run it once eagerly
    first so that autotuning is done."""
    ctx = get_context()
    g = Graph(ctx)
    lib().compas_graph_begin(ctx.ptr)
    _check()
    try:
        yield g
    except BaseException:
        lib().compas_graph_abort(ctx.ptr)
        _check_quiet()
        raise
    g.ptr = _check(lib().compas_graph_end(ctx.ptr))


def _check_quiet():
    try:
        _check()
    except CompasError:
        pass


class Server:
    """Request server of the library (continuous batching on one GPU):
    requests admitted within `memory_budget` bytes of estimated device memory
    share work-table simulation launches; their contractions run on `slots`
    private streams."""

    def __init__(self, slots: int = 4, memory_budget: int = 64 << 30):
        ctx = get_context()
        self.ptr = _check(lib().compas_server_create(ctx.ptr, int(slots), int(memory_budget)))

    def submit_simulate_signal(self, parameters, sequence, trajectory, coils, method="matmul_fast") -> int:
        c = as_array(coils, np.complex64)
        m = SIGNAL_METHODS[method] if isinstance(method, str) else int(method)
        t = lib().compas_server_submit_simulate_signal(self.ptr, parameters.ptr, sequence.ptr, trajectory.ptr,
                                                      c.ptr, m)
        _check()
        return int(t)

    def submit_jacobian_hermitian(self, parameters, sequence, trajectory, coils, y, method="gemm_fast") -> int:
        c = as_array(coils, np.complex64)
        yy = as_array(y, np.complex64)
        m = JACOBIAN_METHODS[method] if isinstance(method, str) else int(method)
        t = lib().compas_server_submit_jacobian_hermitian(self.ptr, parameters.ptr, sequence.ptr, trajectory.ptr,
                                                         c.ptr, yy.ptr, m)
        _check()
        return int(t)

    def wait(self, ticket: int, timeout: float = 600.0):
        """(arrival, start, done) seconds on the server clock, or None on timeout."""
        times = (C.c_double * 3)()
        rc = lib().compas_server_wait(self.ptr, int(ticket), float(timeout), times)
        _check()
        return (times[0], times[1], times[2]) if rc == 1 else None

    def result(self, ticket: int, which: int = 1) -> CompasArray:
        """which = 0: echos; 1: signal (simulate_signal) or J^H.y (jacobian_hermitian);
        2..: the finite differences of a jacobian_hermitian request (T1, T2, then B1, B0 for spiral)."""
        return _wrap(lib().compas_server_result(self.ptr, int(ticket), int(which)))

    def batch_stats(self) -> dict:
        """Work-table launches so far, their work items, the most requests one
        launch served, and the launches that served two or more requests."""
        out = (C.c_int64 * 4)()
        lib().compas_server_batch_stats(self.ptr, out)
        _check()
        return {"launches": out[0], "items": out[1], "max_requests": out[2], "shared_launches": out[3]}

    def release(self, ticket: int):
        lib().compas_server_release(self.ptr, int(ticket))
        _check()

    def submit_wire(self, frame: bytes) -> int:
        """Submit a request frame (compas_amd.wire); returns its ticket."""
        frame = bytes(frame)
        t = lib().compas_server_submit_wire(self.ptr, frame, len(frame))
        _check()
        return int(t)

    def wait_done(self, ticket: int, poll: float = 1.0) -> bool:
        """Block until the request has finished; True if it succeeded (a
        failure is reported by its result frame, not raised)."""
        while True:
            rc = lib().compas_server_wait(self.ptr, int(ticket), float(poll), None)
            if rc != 0:
                ok = lib().compas_last_error() is None
                lib().compas_clear_error()
                return ok and rc == 1

    def result_wire(self, ticket: int) -> bytes:
        """Result frame of a finished request (its arrays read back)."""
        n = lib().compas_server_result_wire(self.ptr, int(ticket), None, 0)
        _check()
        buf = C.create_string_buffer(int(n))
        n2 = lib().compas_server_result_wire(self.ptr, int(ticket), buf, int(n))
        _check()
        if n2 != n:
            raise CompasError("result frame changed size")
        return buf.raw[:n]

    def now(self) -> float:
        return float(lib().compas_server_now(self.ptr))

    def set_kernel_timing(self, enabled: bool):
        lib().compas_server_set_kernel_timing(self.ptr, int(enabled))
        _check()

    def read_kernel_timings(self) -> dict:
        """Per-kernel (launches, total ms) over all slots since the last read."""
        return _parse_timings(lambda buf: lib().compas_server_read_kernel_timings(self.ptr, buf, len(buf)))

    def close(self):
        if self.ptr:
            lib().compas_server_destroy(self.ptr)
            self.ptr = None
            _check()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ GEMM ---

GEMM_METHODS = {"pedantic": 0, "regular": 1, "fast": 2}  # GemmComputeMethod (utils/gemm.h:7-11)


def to_bf16_bits(x) -> np.ndarray:
    """float32 values rounded to the nearest bfloat16 (ties to even), as raw bits."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return r.astype(np.uint16)


def from_bf16_bits(b) -> np.ndarray:
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def _gemm_call(fn, result, lhs, rhs, alpha, beta, method):
    ctx = get_context()
    ok = fn(ctx.ptr, result.ptr, lhs.ptr, rhs.ptr, float(alpha), float(beta), GEMM_METHODS[method])
    _check()
    if ok != 1:
        raise CompasError("GEMM failed")


def compute_gemm(result: CompasArray, lhs: CompasArray, rhs: CompasArray, alpha=1.0, beta=0.0,
                 method="regular") -> None:
    """result = alpha * lhs @ rhs.T + beta * result, in place (compute_gemm, utils/gemm.h:21-37)."""
    _gemm_call(lib().compas_compute_gemm, result, lhs, rhs, alpha, beta, method)


def compute_complex_gemm(result: CompasArray, lhs: CompasArray, rhs: CompasArray, alpha=1.0, beta=0.0,
                         method="regular") -> None:
    """Planar complex [2][m][n] = alpha * lhs @ rhs.T + beta * result (compute_complex_gemm, utils/gemm.h:39-55)."""
    _gemm_call(lib().compas_compute_complex_gemm, result, lhs, rhs, alpha, beta, method)


def convert_complex_to_planar(output: CompasArray, input: CompasArray) -> None:
    ok = lib().compas_convert_complex_to_planar(get_context().ptr, output.ptr, input.ptr)
    _check()
    if ok != 1:
        raise CompasError("conversion failed")


def convert_planar_to_complex(output: CompasArray, input: CompasArray) -> None:
    ok = lib().compas_convert_planar_to_complex(get_context().ptr, output.ptr, input.ptr)
    _check()
    if ok != 1:
        raise CompasError("conversion failed")


def phase_encoding(echos, parameters, trajectory) -> CompasArray:
    ctx = get_context()
    e = as_array(echos, np.complex64)
    return _wrap(lib().phase_encoding(ctx.ptr, e.ptr, parameters.ptr, trajectory.ptr))


def compute_residual(lhs, rhs):
    ctx = get_context()
    a = as_array(lhs, np.complex64)
    b = as_array(rhs, np.complex64)
    obj = C.c_float(0.0)
    diff = _wrap(lib().compas_compute_residual(ctx.ptr, a.ptr, b.ptr, C.byref(obj)))
    return float(obj.value), diff
