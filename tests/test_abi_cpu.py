"""C-ABI checks that need no GPU: the library loads, exports every entry point
declared in include/compas_amd.h, keeps jl_errorf as a weak import, and
reports errors through compas_last_error instead of aborting."""

import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "compas_amd.h")

# julia-bindings/src/bindings.cpp exports exactly these 26 symbols
REFERENCE_EXPORTS = [
    "compas_version", "compas_make_context", "compas_copy_context_for_device", "compas_destroy_context",
    "compas_destroy_object", "compas_synchronize", "compas_make_array_float", "compas_read_array_float",
    "compas_make_array_complex", "compas_read_array_complex", "compas_destroy_array",
    "compas_make_tissue_parameters", "compas_make_fisp_sequence", "compas_make_pssfp_sequence",
    "compas_make_cartesian_trajectory", "compas_make_spiral_trajectory", "compas_simulate_magnetization_fisp",
    "compas_simulate_magnetization_pssfp", "compas_magnetization_to_signal_cartesian",
    "compas_magnetization_to_signal_spiral", "compas_compute_jacobian", "compas_compute_jacobian_hermitian",
    "phase_encoding", "compas_compute_residual", "compas_simulate_magnetization_derivative_pssfp",
    "compas_simulate_magnetization_derivative_fisp",
]


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"typedef[^;]*;", "", text)  # function-pointer types are not entry points
    return sorted(set(re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", text)) - {"defined"})


@pytest.fixture(scope="module")
def lib():
    import compas_amd

    return compas_amd.lib()


def dynamic_symbols():
    import compas_amd

    out = subprocess.run(["nm", "-D", compas_amd.library_path()], capture_output=True, text=True, check=True).stdout
    syms = {}
    for line in out.splitlines():
        parts = line.split()
        if len(parts) >= 2:
            syms[parts[-1]] = parts[-2]
    return syms


def test_header_declares_reference_exports():
    declared = header_functions()
    for name in REFERENCE_EXPORTS:
        assert name in declared, name


def test_library_exports_every_declared_symbol(lib):
    syms = dynamic_symbols()
    for name in header_functions():
        assert syms.get(name) == "T", f"{name} not exported"
        assert hasattr(lib, name)


def test_jl_errorf_is_a_weak_import():
    syms = dynamic_symbols()
    assert syms.get("jl_errorf") in ("w", "v"), syms.get("jl_errorf")


def test_version_string(lib):
    import compas_amd

    assert compas_amd.version().startswith("compas-amd")


def test_errors_are_reported_not_raised_through_c(lib):
    import compas_amd

    # null handles are rejected with a message on the error channel
    assert lib.compas_make_fisp_sequence(None, None, 0.01, 0.005, 0.0, 8, 0.1, 1, 1) is None
    err = lib.compas_last_error()
    assert err and b"null handle" in err
    lib.compas_clear_error()
    assert lib.compas_last_error() is None
    assert lib.compas_compute_residual(None, None, None, None) is None
    assert b"null" in lib.compas_last_error()
    lib.compas_clear_error()


def test_python_wrapper_raises_compas_error(lib):
    import compas_amd

    with pytest.raises(compas_amd.CompasError):
        compas_amd._check(lib.compas_simulate_magnetization_fisp(None, None, None))


def test_no_gpu_context_is_an_error_here():
    import torch  # noqa: F401  (torch reports the device state of this machine)

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import compas_amd

    with pytest.raises(compas_amd.CompasError, match="no HIP device"):
        compas_amd.init_context()


def test_error_channel_is_per_thread(lib):
    """Julia tasks call the library from several threads at once (CompasToolkit.jl:7-14,
    104-136): each thread reads back the error of its own last call.  ctypes
    releases the GIL inside the calls, so the calls overlap in the library."""
    import threading

    calls = [
        (lambda: lib.compas_make_fisp_sequence(None, None, 0.01, 0.005, 0.0, 8, 0.1, 1, 1), b"`RF_train`"),
        (lambda: lib.compas_compute_residual(None, None, None, None), b"null context"),
        (lambda: lib.compas_simulate_magnetization_fisp(None, None, None), b"null context"),
    ]
    failures = []
    barrier = threading.Barrier(len(calls) * 2)

    def worker(call, want):
        barrier.wait()
        for _ in range(2000):
            call()
            err = lib.compas_last_error()
            if not err or want not in err:
                failures.append((want, err))
                return

    threads = [threading.Thread(target=worker, args=c) for c in calls * 2]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not failures, failures[:3]
