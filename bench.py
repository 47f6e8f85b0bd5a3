"""Benchmark of the compas simulation hot path on one MI355X.

Workload (default): FISP EPG simulation of 50176 voxels (224 x 224) x 1120
TRs with 25 states, then the Cartesian k-space signal of 4 coils x 1120
readouts x 224 samples -- one "step" is simulate_magnetization followed by
magnetization_to_signal with all inputs resident in HBM, called through the
C-ABI of libcompas_amd.so.

Prints one JSON line (rank 0).  The K timed steps are independent requests
run through the library's request server on `--pipeline` slots (separate HIP
streams), so the VALU-bound EPG simulation of one step overlaps the MFMA
contraction of another; `value` is voxel.TR/s over those K steps.  The same K
steps are also timed one after another on one stream (`sequential`), which
gives the per-step latency percentiles and kernel durations without overlap.
k-space samples/s, per-kernel times (HIP events on the stream each kernel is
launched on), the roofline of the dominant kernel (from its durations in the
sequential phase, where each launch has the GPU to itself; the overlapped
figure is under `pipeline`) and the CPU baseline (the FP32 port of the
reference algorithm on the host cores, on a bounded sample) are reported
beside it.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
# The request server drives the base stream, its slot streams and three
# simulation streams; with HIP's default of four hardware queues some of them
# would share a queue and run one after another.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"
sys.path.insert(0, os.path.join(ROOT, "compas-toolkit_amd"))

import compas_amd as ca  # noqa: E402
from compas_amd import workloads as W  # noqa: E402
from compas_amd.serving_mix import SERVE_NS, ServeCatalog, serve_request_spec, serve_submit  # noqa: E402

PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector (published)
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector (published)
PEAK_BF16_TFLOPS = 2500.0  # MI355X dense BF16/FP16 MFMA (published)
PEAK_HBM_GBS = 8000.0


def algorithmic_work(kernel: str, wl, method: str):
    """(amount per launch, unit, bound, peak) of a kernel of this workload."""
    nv = wl.tissue.nvoxels
    nTR = int(wl.sequence.RF.shape[0])
    nr, ns = wl.trajectory.nreadouts, wl.trajectory.samples_per_readout
    nc = wl.coils.shape[0]
    if kernel == "fisp_epg":
        # 82 N flop per voxel.TR (excite 54 N + two rotate_decay 14 N)
        return 82.0 * wl.sequence.max_state * nv * nTR / 1e12, "TFLOP/s", "valu", PEAK_FP32_TFLOPS
    if kernel in ("contraction_fp32", "contraction_mfma"):
        flops = 8.0 * nc * nr * ns * nv / 1e12
        if kernel == "contraction_mfma":
            return flops, "TFLOP/s", "mfma", PEAK_BF16_TFLOPS
        return flops, "TFLOP/s", "valu", PEAK_FP32_TFLOPS
    return None


def measured_traffic(kernel: str, workload: int):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (two rocprofv3 --pmc runs, FETCH_SIZE and WRITE_SIZE, scaled by the same
    counters of the stream_copy dispatch of known size), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as fh:
            data = json.load(fh)
    except (OSError, ValueError):
        return None
    entry = data.get(f"workload{workload}", {}).get(kernel)
    if not entry:
        return None
    return entry.get("bytes_per_launch")


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs), q))


def median_time(fn, reps=5):
    """(last result, median seconds, all seconds) of `reps` timed calls."""
    ts = []
    out = None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t0)
    return out, float(np.median(ts)), ts


def oracle_for_timing():
    """(pyoracle, native, threads): the FP32 port is timed with the oracle built
    for this host (-march=native) when that builds, else the portable build."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    pyoracle.build()
    threads = int(os.environ.get("ORACLE_THREADS", os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    return pyoracle, pyoracle.native_lib() is not None, threads


def cpu_baseline(wl, gpu_echos, gpu_signal, seconds_budget=20.0, other_signals=None):
    """FP32 port of the reference algorithm (oracle/, test infrastructure) timed on
    the host cores on a bounded sample (one calibration run as warm-up, then the
    median of 5 runs), plus the accuracy of the GPU outputs against the FP64
    oracle on the same sample."""
    pyoracle, native, threads = oracle_for_timing()
    nv = wl.tissue.nvoxels
    nTR = int(wl.sequence.RF.shape[0])
    nr, ns = wl.trajectory.nreadouts, wl.trajectory.samples_per_readout
    nc = wl.coils.shape[0]

    # simulation: contiguous voxel slice (voxels are independent), sized from a short
    # calibration run so that 5 runs take about a third of the time budget
    ncal = max(64, nv // 64)
    t0 = time.perf_counter()
    pyoracle.fisp(wl.tissue.slice(0, ncal), wl.sequence, "f32", threads, native=native)
    t_cal = max(time.perf_counter() - t0, 1e-6)
    nsim = int(min(nv, max(ncal, ncal * (seconds_budget / 15) / t_cal)))
    sub = wl.tissue.slice(0, nsim)
    e32, t_sim, sim_runs = median_time(lambda: pyoracle.fisp(sub, wl.sequence, "f32", threads, native=native))
    e32p = pyoracle.fisp(sub, wl.sequence, "f32", threads)  # portable build: the parity proxy
    e64 = pyoracle.fisp(sub, wl.sequence, "f64", threads)
    mag_err64 = pyoracle.rel_l2(gpu_echos[:, :nsim], e64)
    mag_err32 = pyoracle.rel_l2(gpu_echos[:, :nsim], e32p)
    port_err = pyoracle.rel_l2(e32p, e64)

    # signal: a subset of readouts (Cartesian rows are independent), full voxel sum,
    # sized from a 2-readout calibration so that 5 runs take about two thirds of the budget
    def rows_traj(ridx):
        return W.CartesianTraj(len(ridx), ns, wl.trajectory.delta_t, wl.trajectory.k_start[ridx],
                               wl.trajectory.delta_k)

    cal = np.array([0, nr - 1])
    t0 = time.perf_counter()
    pyoracle.signal_cartesian(wl.tissue, rows_traj(cal), wl.coils, gpu_echos[cal].astype(np.complex64), "f32",
                              threads, native=native)
    t_cal = max(time.perf_counter() - t0, 1e-6)
    rows = int(min(nr, max(2, 2 * (2 * seconds_budget / 15) / t_cal)))
    ridx = np.linspace(0, nr - 1, rows).astype(int)
    traj = rows_traj(ridx)
    echos_rows = gpu_echos[ridx].astype(np.complex64)
    s32, t_sig, sig_runs = median_time(
        lambda: pyoracle.signal_cartesian(wl.tissue, traj, wl.coils, echos_rows, "f32", threads, native=native))
    s64 = pyoracle.signal_cartesian(wl.tissue, traj, wl.coils, echos_rows.astype(np.complex128), "f64", threads)
    sig_err64 = pyoracle.rel_l2(gpu_signal[:, ridx, :], s64)
    other_err = {f"signal_{k}_rel_l2_vs_fp64": pyoracle.rel_l2(v[:, ridx, :], s64)
                 for k, v in (other_signals or {}).items()}

    sim_rate = nsim * nTR / t_sim  # voxel.TR/s
    sig_rate = nc * rows * ns / t_sig  # k-space samples/s
    t_step = nv * nTR / sim_rate + nc * nr * ns / sig_rate
    build = "-O3 -march=native (built on this host)" if native else "-O3 -march=x86-64-v3 (native build failed)"
    base = {
        "value": nv * nTR / t_step,
        "unit": "voxel*TR/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"FP32 port of the reference algorithm, gcc {build}, OpenMP: FISP on voxels [0, {nsim}) of "
                   f"{nv} (median {t_sim:.2f} s of 5 runs), Cartesian signal on {rows} of {nr} readouts x all "
                   f"{nv} voxels (median {t_sig:.2f} s of 5 runs); step time = whole workload at these "
                   f"size-independent rates (voxel-TR/s, samples/s)"),
        "sim_voxel_tr_per_s": sim_rate,
        "signal_samples_per_s": sig_rate,
        "runs_s": {"sim": sim_runs, "signal": sig_runs},
    }
    accuracy = {
        "magnetization_rel_l2_vs_fp64": mag_err64,
        "magnetization_rel_l2_vs_fp32_port": mag_err32,
        "fp32_port_rel_l2_vs_fp64": port_err,
        "signal_rel_l2_vs_fp64": sig_err64,
        **other_err,
        "sample": f"voxels [0, {nsim}) for magnetization; readouts {ridx.tolist()} for signal",
    }
    return base, accuracy


PRODUCT_METHODS = {"matmul_fast": "gemm_fast", "matmul_low": "gemm_low", "direct": "direct"}
MFMA_PASSES = {"matmul_fast": 3, "matmul_low": 1, "direct": 0}


def products_bench(args):
    """signal + J.v + J^H.y at the 4-coil Cartesian shapes (1120 readouts x 224
    samples x 50176 voxels): one step is the three products on resident inputs
    (FISP echos and their finite-difference T1/T2 derivatives, a seeded [4][nv]
    direction and a seeded [4][1120][224] residual).  value = k-space samples/s
    summed over the three outputs; the roofline covers the contraction kernels
    (useful complex-MAC flops; the hi+lo split issues 3x that on MFMA)."""
    sig_method = args.method
    if sig_method not in PRODUCT_METHODS:
        sys.exit(f"--products takes --method in {sorted(PRODUCT_METHODS)}")
    jac_method = PRODUCT_METHODS[sig_method]
    wl = W.config(2)
    nv = wl.tissue.nvoxels
    nr, ns = wl.trajectory.nreadouts, wl.trajectory.samples_per_readout
    nc = wl.coils.shape[0]
    ctx = ca.init_context(0)
    params = ca.TissueParameters.from_tissue(wl.tissue)
    seq = ca.FispSequence.from_sequence(wl.sequence)
    traj = ca.CartesianTrajectory.from_trajectory(wl.trajectory)
    coils = ca.CompasArray.from_numpy(wl.coils)
    echos = ca.simulate_magnetization(params, seq)
    d = ca.simulate_magnetization_derivatives(echos, params, seq)
    rng = W.SplitMix64(0xC0A50004)
    v = ca.CompasArray.from_numpy((rng.normal(4 * nv) + 1j * rng.normal(4 * nv)).reshape(4, nv).astype(np.complex64))
    y = ca.CompasArray.from_numpy(((rng.normal(nc * nr * ns) + 1j * rng.normal(nc * nr * ns)) / np.sqrt(2))
                                  .reshape(nc, nr, ns).astype(np.complex64))
    ctx.synchronize()

    def step():
        s = ca.magnetization_to_signal(echos, params, traj, coils, method=sig_method)
        jv = ca.compute_jacobian(echos, d, params, traj, coils, v, method=jac_method)
        jh = ca.compute_jacobian_hermitian(echos, d, params, traj, coils, y, method=jac_method)
        return s, jv, jh

    for _ in range(args.warmup):
        out = step()
    ctx.synchronize()
    del out
    ctx.read_kernel_timings()
    ctx.set_kernel_timing(True)
    lat = []
    ctx.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        out = step()
        ctx.synchronize()
        lat.append(time.perf_counter() - t0)
    t_total = time.perf_counter() - t_start
    ctx.set_kernel_timing(False)
    timings = ctx.read_kernel_timings()
    kernels = {k: {"launches": c, "avg_ms": tot / max(c, 1), "total_ms": tot} for k, (c, tot) in timings.items()}
    contraction = [k for k in kernels if k in ("contraction_mfma", "hermitian_mfma", "contraction_fp32")]
    t_con = sum(kernels[k]["total_ms"] for k in contraction) / args.steps * 1e-3
    useful = 8.0 * nc * nr * ns * nv * (1 + 2 + 2)  # signal, J.v (two terms), J^H.y (two terms)
    passes = MFMA_PASSES[sig_method]
    if passes:
        roof = {"kernel": "+".join(sorted(contraction)), "bound": "mfma", "achieved": useful / t_con / 1e12,
                "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s", "frac": useful / t_con / 1e12 / PEAK_BF16_TFLOPS,
                "issued_frac": passes * useful / t_con / 1e12 / PEAK_BF16_TFLOPS, "traffic": None}
    else:
        roof = {"kernel": "+".join(sorted(contraction)), "bound": "valu", "achieved": useful / t_con / 1e12,
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": useful / t_con / 1e12 / PEAK_FP32_TFLOPS,
                "traffic": None}
    line = {
        "metric": "signal + J.v + J^H.y k-space samples/s (4 coils x 1120 x 224, 50176 voxels)",
        "value": 3 * nc * nr * ns * args.steps / t_total,
        "unit": "k-space samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if not passes else f"f32 operands as bf16 x{passes} on MFMA, f32 accumulate",
        "data": "synthetic (seeded tissue, FISP echos, finite-difference derivatives, random direction/residual)",
        "config": {"workload": f"Cartesian signal + J.v + J^H.y, {nc} coils x {nr} readouts x {ns} samples x "
                               f"{nv} voxels", "signal_method": sig_method, "jacobian_method": jac_method},
        "latency_ms": {"p50": 1e3 * percentile(lat, 50), "p99": 1e3 * percentile(lat, 99)},
        "kernels": kernels,
        "contraction_tuning": ca.contraction_tuning_report().strip().splitlines(),
        "roofline": roof,
        "derivatives": fisp_derivative_timing(ctx, params, seq, echos),
    }
    print(json.dumps(line))


def fisp_derivative_timing(ctx, params, seq, echos):
    """Wall time of the FISP echo derivatives at these shapes: the fused tangent
    pass (T1 + T2, and all four fields), one tangent pass per field, and the
    reference's forward differences (one more simulation per field)."""
    def timed(fn):
        fn()  # warm-up
        ctx.synchronize()
        t0 = time.perf_counter()
        res = fn()
        ctx.synchronize()
        return res, 1e3 * (time.perf_counter() - t0)

    two, four = (ca.T1, ca.T2), (ca.T1, ca.T2, ca.B1, ca.B0)
    fused2, t_f2 = timed(lambda: ca.simulate_magnetization_tangents(two, params, seq))
    _, t_f2e = timed(lambda: ca.simulate_magnetization_tangents(two, params, seq, echos=True))
    _, t_f4 = timed(lambda: ca.simulate_magnetization_tangents(four, params, seq))
    per, t_per = timed(lambda: [ca.simulate_magnetization_tangent(f, params, seq) for f in two])
    _, t_fd = timed(lambda: [ca.simulate_magnetization_derivative(f, echos, params, seq) for f in two])
    _, t_sim = timed(lambda: ca.simulate_magnetization(params, seq))
    diff = max(float(np.linalg.norm(a.collect() - b.collect()) / np.linalg.norm(b.collect()))
               for a, b in zip(fused2, per))
    return {"T1T2_fused_ms": t_f2, "T1T2_fused_with_echos_ms": t_f2e, "T1T2B1B0_fused_ms": t_f4,
            "T1T2_per_field_ms": t_per, "T1T2_finite_difference_ms": t_fd, "echos_ms": t_sim,
            "fused_vs_per_field_rel_l2": diff,
            "note": "FISP at the products shapes; finite difference: the reference's method (delta 1e-4), one "
                    "more simulation per field, the echoes given"}


def spiral_bench(args):
    """pSSFP + spiral workload: 50176 voxels x 1120 TRs x 35 isochromats x 25 RF
    sub-steps, golden-angle spiral 4 coils x 1120 x 512, six-parameter J.v and
    J^H.y.  One step = simulate_magnetization (pSSFP) + magnetization_to_signal
    (spiral) + J.v + J^H.y; the finite-difference derivatives of the echoes
    (T1, T2, B1, B0) are computed once before timing and held fixed.  value =
    voxel.TR/s of the step; the CPU baseline runs the FP32 port of the same four
    operations on bounded subsets (voxels for the simulation and J^H.y, readouts
    for the signal and J.v) and extrapolates linearly."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    wl = W.config(3)
    nv = wl.tissue.nvoxels
    nTR = int(wl.sequence.RF.shape[0])
    nz, nRF = int(wl.sequence.z.shape[0]), int(wl.sequence.gamma_dt_RF.shape[0])
    nr, ns = wl.trajectory.nreadouts, wl.trajectory.samples_per_readout
    nc = wl.coils.shape[0]
    ctx = ca.init_context(0)
    params = ca.TissueParameters.from_tissue(wl.tissue)
    seq = ca.pSSFPSequence.from_sequence(wl.sequence)
    traj = ca.SpiralTrajectory.from_trajectory(wl.trajectory)
    coils = ca.CompasArray.from_numpy(wl.coils)
    v = ca.CompasArray.from_numpy(wl.extra["v"])
    y = ca.CompasArray.from_numpy(wl.extra["y"])
    echos0 = ca.simulate_magnetization(params, seq)
    ds = [ca.simulate_magnetization_derivative(f, echos0, params, seq) for f in (ca.T1, ca.T2, ca.B1, ca.B0)]
    ctx.synchronize()

    # --method direct: the plain FP32 VALU sums; otherwise the signal method given and
    # the mixed-precision products (MFMA)
    jac_method = "direct" if args.method in ("direct", "naive") else ("gemm_low" if args.method == "matmul_low"
                                                                       else "gemm_fast")

    def step():
        e = ca.simulate_magnetization(params, seq)
        s = ca.magnetization_to_signal(e, params, traj, coils, method=args.method)
        jv = ca.compute_jacobian_spiral(e, ds, params, traj, coils, v, method=jac_method)
        jh = ca.compute_jacobian_hermitian_spiral(e, ds, params, traj, coils, y, method=jac_method)
        return e, s, jv, jh

    for _ in range(args.warmup):
        out = step()
    ctx.synchronize()
    del out
    ctx.read_kernel_timings()
    ctx.set_kernel_timing(True)
    lat = []
    ctx.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        out = step()
        ctx.synchronize()
        lat.append(time.perf_counter() - t0)
    t_total = time.perf_counter() - t_start
    ctx.set_kernel_timing(False)
    timings = ctx.read_kernel_timings()
    kernels = {k: {"launches": c, "avg_ms": tot / max(c, 1), "total_ms": tot} for k, (c, tot) in timings.items()}

    # algorithmic work per launch: pSSFP (57 nRF + 28) flop per isochromat.TR (FP64 state); spiral
    # sums 8 nc flop per (readout, sample, voxel) for the coil MACs + 6 for the sample recurrence
    # (J.v: two amplitudes, 16 nc + 6); J^H.y: 8 nc (y conj(Phi)) + 8 nc (the (s - ns/2) moment) + 6.
    # The spiral kernels run the contraction on MFMA (BF16 hi + lo: three v_mfma_f32_16x16x32_bf16
    # per 16 voxels x 16 samples x 16 rows) and generate the sample factor on VALU, which binds:
    # each entry gives the algorithmic rate against the FP32 vector peak and the issued MFMA
    # fraction.
    rsv = float(nr) * ns * nv
    rooflines = {}
    if "pssfp_iso" in kernels:
        ach = (57.0 * nRF + 28) * nz * nTR * nv / (kernels["pssfp_iso"]["avg_ms"] * 1e-3) / 1e12
        rooflines["pssfp_iso"] = {"bound": "valu", "achieved": ach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                                  "frac": ach / PEAK_FP64_TFLOPS,
                                  "note": "FP64 isochromat state; the FP32 vector peak is twice this"}
    issued = 2.0 * 16 * 2 * rsv * 3  # 16 real rows x 2 nv real K x ns x nr, three products
    mfma_kernels = {"spiral_sum_mfma": ((8.0 * nc + 6) * rsv + (16.0 * nc + 6) * rsv, 2 * issued),  # signal + J.v
                    "spiral_hermitian_mfma": ((16.0 * nc + 6) * rsv, issued)}
    for k, (flop, iss) in mfma_kernels.items():
        if k in kernels:
            step_s = kernels[k]["total_ms"] / args.steps * 1e-3
            ach = flop / step_s / 1e12
            rooflines[k] = {"bound": "valu", "achieved": ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                            "frac": ach / PEAK_FP32_TFLOPS,
                            "mfma_issued_frac": iss / step_s / 1e12 / PEAK_BF16_TFLOPS}
    for k, flop in (("signal_spiral_direct", (8.0 * nc + 6) * rsv + (16.0 * nc + 6) * rsv),
                    ("spiral_hermitian", (16.0 * nc + 6) * rsv)):
        if k in kernels:  # the FP32 VALU forms (plain methods)
            ach = flop / (kernels[k]["total_ms"] / args.steps * 1e-3) / 1e12
            rooflines[k] = {"bound": "valu", "achieved": ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                            "frac": ach / PEAK_FP32_TFLOPS}
    dominant = max(kernels, key=lambda k: kernels[k]["total_ms"])
    roof = dict(rooflines.get(dominant, {}), kernel=dominant, traffic=None) if dominant in rooflines else None

    # derivatives of the echoes w.r.t. T1, T2, B1, B0: the fused tangent pass
    # (all four fields, optionally the echoes, one recursion), one tangent pass
    # per field, and the reference's forward differences (one extra simulation
    # per field beside the echoes)
    fields = (ca.T1, ca.T2, ca.B1, ca.B0)

    def timed(fn):
        ctx.synchronize()
        t0 = time.perf_counter()
        res = fn()
        ctx.synchronize()
        return res, 1e3 * (time.perf_counter() - t0)

    timed(lambda: ca.simulate_magnetization_tangents(fields, params, seq, echos=True))  # warm-up
    fused, t_fused = timed(lambda: ca.simulate_magnetization_tangents(fields, params, seq))
    fused_e, t_fused_e = timed(lambda: ca.simulate_magnetization_tangents(fields, params, seq, echos=True))
    del fused_e
    timed(lambda: [ca.simulate_magnetization_tangent(f, params, seq) for f in fields])  # warm-up
    tangents, t_tan = timed(lambda: [ca.simulate_magnetization_tangent(f, params, seq) for f in fields])
    fds, t_fd = timed(lambda: [ca.simulate_magnetization_derivative(f, echos0, params, seq) for f in fields])
    derivatives = {"fields": "T1 T2 B1 B0", "fused_ms": t_fused, "fused_with_echos_ms": t_fused_e,
                   "tangent_per_field_ms": t_tan, "finite_difference_ms": t_fd,
                   "note": "fused: exact derivatives of the recursion, all fields in one pass (with_echos: and "
                           "the echoes); per field: one tangent pass each; finite difference: the reference's "
                           "method (delta 1e-4), one more simulation per field, the echoes given"}
    derivatives["fused_vs_per_field_rel_l2"] = max(
        float(np.linalg.norm(a.collect() - b.collect()) / np.linalg.norm(b.collect())) for a, b in zip(fused, tangents))
    del tangents

    e, s, jv, jh = (a.collect() for a in out)
    dsn = [d.collect() for d in ds]
    line = {
        "metric": "voxel*TR/s sim + k-space samples/s (pSSFP + spiral signal + 6-parameter J.v and J^H.y)",
        "value": nv * nTR * args.steps / t_total,
        "unit": "voxel*TR/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 tissue maps, sinc-Gaussian RF, golden-angle spiral, Gaussian coils)",
        "config": {"workload": f"pSSFP {nv} voxels x {nTR} TRs x {nz} isochromats x {nRF} RF steps + spiral signal "
                               f"{nc} coils x {nr} readouts x {ns} samples + J.v [6][{nv}] + J^H.y",
                   "signal_method": args.method, "jacobian_method": jac_method},
        "isochromat_tr_per_s": nv * nTR * nz * args.steps / t_total,
        "kspace_samples_per_s": 3 * nc * nr * ns * args.steps / t_total,
        "latency_ms": {"p50": 1e3 * percentile(lat, 50), "p99": 1e3 * percentile(lat, 99)},
        "kernels": kernels,
        "rooflines": rooflines,
        "roofline": roof,
        "derivatives": derivatives,
    }
    if not args.no_cpu_baseline:
        base, acc = spiral_cpu_baseline(wl, e, s, jv, jh, dsn)
        line["cpu_baseline"] = base
        line["accuracy"] = acc
        derivatives.update(derivative_accuracy(wl, [t.collect() for t in fused], [f.collect() for f in fds]))
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line))


def spiral_cpu_baseline(wl, e, s, jv, jh, ds, seconds_budget=16.0):
    """FP32 port (oracle/) of the pSSFP + spiral step on bounded subsets, and the
    error of the GPU outputs against the FP64 oracle on the same subsets.  Each
    leg is sized by a calibration run (the warm-up) and timed as the median of
    5 runs; the FP32 port is the oracle built for this host when that builds."""
    pyoracle, native, threads = oracle_for_timing()
    nv = wl.tissue.nvoxels
    nTR = int(wl.sequence.RF.shape[0])
    nr, ns = wl.trajectory.nreadouts, wl.trajectory.samples_per_readout
    nc = wl.coils.shape[0]
    budget = seconds_budget / 4 / 5  # per run: four legs of five runs

    def calibrated(run, unit_max, first):
        t0 = time.perf_counter()
        run(first)
        t_cal = max(time.perf_counter() - t0, 1e-6)
        return int(min(unit_max, max(first, first * budget / t_cal)))

    def tissue_sub(idx):
        return wl.tissue.replace(**{f: getattr(wl.tissue, f)[idx] for f in ("T1", "T2", "B1", "B0", "rho", "x", "y",
                                                                            "z")})

    def traj_sub(ridx):
        t = wl.trajectory
        return W.SpiralTraj(len(ridx), ns, t.delta_t, t.k_start[ridx], t.delta_k[ridx])

    # simulation: contiguous voxel slice
    nsim = calibrated(lambda n: pyoracle.pssfp(wl.tissue.slice(0, n), wl.sequence, "f32", threads, native=native),
                      nv, 16)
    _, t_sim, _ = median_time(lambda: pyoracle.pssfp(wl.tissue.slice(0, nsim), wl.sequence, "f32", threads,
                                                     native=native))
    e64 = pyoracle.pssfp(wl.tissue.slice(0, nsim), wl.sequence, "f64", threads)
    # signal and J.v: readout subsets (rows are independent)
    def rows(n):
        return np.linspace(0, nr - 1, n).astype(int)

    nsig = calibrated(lambda n: pyoracle.signal_spiral(wl.tissue, traj_sub(rows(n)), wl.coils, e[rows(n)], "f32",
                                                       threads, native=native), nr, 2)
    ri = rows(nsig)
    _, t_sig, _ = median_time(lambda: pyoracle.signal_spiral(wl.tissue, traj_sub(ri), wl.coils, e[ri], "f32",
                                                             threads, native=native))
    s64 = pyoracle.signal_spiral(wl.tissue, traj_sub(ri), wl.coils, e[ri].astype(np.complex128), "f64", threads)
    v = wl.extra["v"]
    njv = max(2, nsig // 2)
    rj = rows(njv)
    _, t_jv, _ = median_time(lambda: pyoracle.jacobian_spiral(wl.tissue, traj_sub(rj), wl.coils, e[rj],
                                                              *[d[rj] for d in ds], v, "f32", threads,
                                                              native=native))
    jv64 = pyoracle.jacobian_spiral(wl.tissue, traj_sub(rj), wl.coils, e[rj], *[d[rj] for d in ds], v, "f64",
                                    threads)
    # J^H.y: voxel subset (outputs per voxel are independent)
    nvh = calibrated(lambda n: pyoracle.jacobian_hermitian_spiral(
        tissue_sub(np.arange(n)), wl.trajectory, wl.coils[:, :n], e[:, :n], *[d[:, :n] for d in ds], wl.extra["y"],
        "f32", threads, native=native), nv, 8)
    vi = np.linspace(0, nv - 1, nvh).astype(int)
    ts = tissue_sub(vi)
    _, t_jh, _ = median_time(lambda: pyoracle.jacobian_hermitian_spiral(
        ts, wl.trajectory, wl.coils[:, vi], e[:, vi], *[d[:, vi] for d in ds], wl.extra["y"], "f32", threads,
        native=native))
    jh64 = pyoracle.jacobian_hermitian_spiral(ts, wl.trajectory, wl.coils[:, vi], e[:, vi], *[d[:, vi] for d in ds],
                                              wl.extra["y"], "f64", threads)
    t_step = (nv / nsim) * t_sim + (nr / nsig) * t_sig + (nr / njv) * t_jv + (nv / nvh) * t_jh
    base = {
        "value": nv * nTR / t_step,
        "unit": "voxel*TR/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"FP32 port, gcc -O3 {'-march=native' if native else '-march=x86-64-v3'}, OpenMP, medians of 5 "
                   f"runs: pSSFP on voxels [0, {nsim}) ({t_sim:.2f} s), spiral signal on {nsig} of {nr} readouts "
                   f"({t_sig:.2f} s), J.v on {njv} readouts ({t_jv:.2f} s), J^H.y on {nvh} voxels ({t_jh:.2f} s); "
                   f"step time = whole workload at these size-independent rates ({t_step:.1f} s)"),
    }
    acc = {
        "magnetization_rel_l2_vs_fp64": pyoracle.rel_l2(e[:, :nsim], e64),
        "signal_rel_l2_vs_fp64": pyoracle.rel_l2(s[:, ri, :], s64),
        "jv_rel_l2_vs_fp64": pyoracle.rel_l2(jv[:, rj, :], jv64),
        "jhy_rel_l2_vs_fp64": pyoracle.rel_l2(jh[:, vi], jh64),
        "note": "signal, J.v and J^H.y are checked against the FP64 oracle on the GPU echoes and derivatives",
    }
    return base, acc


def derivative_accuracy(wl, tangents, fds, nsample=48):
    """rel-L2 of the GPU derivatives (tangent and forward difference) against the
    FP64 oracle's central difference, on a strided sample of voxels.  Steps are
    exact in float32: 2^-16 of the value's binade for T1/T2, 2^-16 for B1,
    2^-14 Hz for B0."""
    import pyoracle

    threads = int(os.environ.get("ORACLE_THREADS", os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    nv = wl.tissue.nvoxels
    vi = np.linspace(0, nv - 1, nsample).astype(int)
    sub = wl.tissue.replace(**{f: getattr(wl.tissue, f)[vi] for f in ("T1", "T2", "B1", "B0", "rho", "x", "y", "z")})
    out = {"sample": f"{nsample} voxels (strided), FP64 central differences"}
    for k, name in enumerate(("T1", "T2", "B1", "B0")):
        v = getattr(sub, name).astype(np.float32)
        if name in ("T1", "T2"):
            h = np.exp2(np.floor(np.log2(v)) - 16).astype(np.float32)
        else:
            h = np.full_like(v, 2.0 ** (-16 if name == "B1" else -14))
        vp, vm = (v + h).astype(np.float32), (v - h).astype(np.float32)
        fp = pyoracle.pssfp(sub.replace(**{name: vp}), wl.sequence, "f64", threads)
        fm = pyoracle.pssfp(sub.replace(**{name: vm}), wl.sequence, "f64", threads)
        want = (fp - fm) / (vp.astype(np.float64) - vm.astype(np.float64))[None, :]
        out[f"{name}_tangent_rel_l2"] = pyoracle.rel_l2(tangents[k][:, vi], want)
        out[f"{name}_finite_difference_rel_l2"] = pyoracle.rel_l2(fds[k][:, vi], want)
    return out


def serve_wire_bench(args):
    """The serving loop through the wire format: every request is a frame
    (pre-encoded on the client), sent over the TCP front end on --connections
    connections in arrival order (round robin) at Poisson arrivals of --load x
    the capacity measured the same way; latency = scheduled arrival -> result
    frame received by the client (so it includes parsing, uploads, result
    read-back, encoding and the socket).  A sample of results is checked
    bitwise against the same requests submitted as library objects."""
    import threading

    from compas_amd import wire

    ctx = ca.init_context(0)
    n_req = args.requests
    specs = [serve_request_spec(i) for i in range(n_req)]
    cat = ServeCatalog()
    t0 = time.perf_counter()
    frames = [cat.frame(sp) for sp in specs]
    t_encode = time.perf_counter() - t0
    frame_bytes = sum(len(f) for f in frames)
    budget = int(args.budget_gib * (1 << 30))
    server = ca.Server(slots=args.slots, memory_budget=budget)
    fe = wire.TcpFrontEnd(server)
    nconn = max(1, args.connections)

    def run(ids, arrivals):
        """Send frames ids[j] at arrivals[j] (seconds from now, None = at once);
        returns {id: (t_send_scheduled, t_received, result)} on the client clock."""
        clients = [wire.TcpClient(fe.address) for _ in range(nconn)]
        per_conn = [[j for j in range(len(ids)) if j % nconn == k] for k in range(nconn)]
        received = [[] for _ in range(nconn)]

        def reader(k):
            for _ in per_conn[k]:
                r = clients[k].recv()
                received[k].append((time.perf_counter(), r))

        readers = [threading.Thread(target=reader, args=(k,)) for k in range(nconn)]
        for th in readers:
            th.start()
        base = time.perf_counter() + 0.05
        sched = {}
        for j, i in enumerate(ids):
            when = base + (arrivals[j] if arrivals is not None else 0.0)
            delay = when - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            sched[i] = when if arrivals is not None else time.perf_counter()
            clients[j % nconn].send(frames[i])
        for th in readers:
            th.join()
        for c in clients:
            c.close()
        got = {}
        for k in range(nconn):  # results come back per connection in sending order
            for j, (t, r) in zip(per_conn[k], received[k]):
                got[ids[j]] = (sched[ids[j]], t, r)
        return got

    # capacity: a closed-loop burst of a sample, through the same path
    cal_ids = np.random.default_rng(6).choice(n_req, size=min(calibration_size(args, n_req), n_req),
                                              replace=False).tolist()
    cal = run(cal_ids, None)
    cal_span = max(t for _, t, _ in cal.values()) - min(s for s, _, _ in cal.values())
    capacity = len(cal_ids) / cal_span
    lam = args.load * capacity
    arrivals = np.cumsum(np.random.default_rng(7).exponential(1.0 / lam, size=n_req))
    got = run(list(range(n_req)), arrivals)
    fe.close()
    errors = [f"request {i}: {r['error']}" for i, (_, _, r) in got.items() if not r["ok"]]
    lat = np.array([t - s for s, t, _ in got.values()])
    span = max(t for _, t, _ in got.values()) - min(s for s, _, _ in got.values())

    # a sample submitted as library objects must give the same bits
    check_ids = sorted(np.random.default_rng(5).choice(n_req, size=min(args.check, n_req), replace=False).tolist())
    worst = 0.0
    for i in check_ids:
        r = cat.request(specs[i])
        t = serve_submit(server, r)
        server.wait(t)
        want = server.result(t).collect()
        server.release(t)
        res = got[i][2]
        key = "jhy" if specs[i]["kind"] == "jhy" else "signal"
        worst = max(worst, 0.0 if np.array_equal(res[key], want) else rel_l2(res[key], want))
    server.close()
    vox_tr = sum(sp["nv"] * sp["nTR"] for sp in specs)
    counts = {k: sum(1 for sp in specs if sp["kind"] == k) for k in ("fisp", "pssfp", "jhy")}
    print(json.dumps({
        "metric": "serving over the wire format: requests/s and p50/p99 latency (arrival -> result frame at the client)",
        "value": len(got) / span,
        "unit": "requests/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f32 (signal and J^H.y contractions BF16 hi+lo on MFMA)",
        "data": "synthetic (the seeded request mix of --serve, every request a self-contained frame)",
        "config": {"workload": f"{n_req} requests as frames over TCP ({nconn} connections), Poisson arrivals at "
                               f"{args.load:.0%} of measured capacity, {args.slots} slots", "mix": counts},
        "voxel_tr_per_s": vox_tr / span,
        "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99)),
                       "mean": 1e3 * float(lat.mean())},
        "arrival_rate_per_s": lam,
        "capacity_per_s": capacity,
        "frame_bytes": {"total": frame_bytes, "mean": frame_bytes / n_req},
        "encode_s": t_encode,
        "wire_vs_objects_max_rel_l2": worst,
        "wire_vs_objects_checked": len(check_ids),
        "errors": errors[:5],
    }))


def reconstruct_bench(args):
    """Gauss-Newton fits on the HIP path (compas_amd.reconstruct): noise-free
    k-space data of a seeded phantom, a perturbed start, a fixed number of
    outer iterations; each run three ways: the normal operator J^H(J v)
    replayed from a captured HIP graph with CG on the device, the same with
    CG on the host, and the operator called eagerly (CG on the host).
      cartesian: 4096 voxels, FISP 256 TRs, 16 states, Cartesian 256 x 64,
                 1 coil; T1, T2, rho; 12 iterations, CG <= 100
      spiral:    the pSSFP spiral workload at 1/4 scale (3136 voxels, 280 TRs x 35
                 isochromats, spiral 280 x 128, 4 coils), T1, T2, B1, rho;
                 8 iterations, CG <= 100
    value = Gauss-Newton iterations/s of the cartesian fit with graphs and
    CG on the device."""
    from compas_amd import reconstruct as R

    ca.init_context(0)
    out = {}
    for name, wl, iters, fit_b1 in (("cartesian", W.config(1), 12, False), ("spiral", W.config(3, scale=4), 8, True)):
        t = wl.tissue
        params, seq, traj = _objects(wl)
        truth = R.Estimate(t.T1.copy(), t.T2.copy(), t.B1.copy(), t.B0.copy(), t.rho.copy(), t.x, t.y, t.z)
        spiral = isinstance(wl.trajectory, W.SpiralTraj)
        m = ca.simulate_magnetization(params, seq)
        if not spiral:
            m = ca.phase_encoding(m, params, traj)
        y = ca.magnetization_to_signal(m, params, traj, wl.coils, method="direct").collect()
        start = truth.copy()
        start.T1 = (t.T1 * 1.2).astype(np.float32)
        start.T2 = (t.T2 * 0.85).astype(np.float32)
        start.rho = (t.rho * 0.9).astype(np.complex64)
        if fit_b1:
            start.B1 = (t.B1 + 0.04).astype(np.float32)
        runs = {}
        # the first run warms up (autotuning, torch's allocator) and is not kept
        for label, use_graph, cg_dev in (("warmup", True, True), ("graph_device_cg", True, True),
                                         ("graph_host_cg", True, False), ("eager", False, False)):
            progress(f"reconstruct {name}: {label}")
            t0 = time.perf_counter()
            res = R.gauss_newton(y, start, seq, traj, wl.coils, iterations=iters, cg_iterations=100, cg_tol=1e-9,
                                 fit_b1=fit_b1, use_graph=use_graph, cg_on_device=cg_dev)
            wall = time.perf_counter() - t0
            if label == "warmup":
                continue
            runs[label] = {
                "seconds": wall, "iterations": len(res.cg_iterations), "ms_per_iteration": 1e3 * wall / max(1, len(res.cg_iterations)),
                "normal_calls": res.normal_calls, "normal_ms_per_call": 1e3 * res.normal_seconds / max(1, res.normal_calls),
                "objective_first": res.objective[0], "objective_last": res.objective[-1],
                "cg_iterations": res.cg_iterations,
                **{f"{f}_rel_err_p{q}": float(np.percentile(np.abs(getattr(res.estimate, f) / getattr(t, f) - 1), q))
                   for f in ("T1", "T2") for q in (50, 95, 100)}}
        out[name] = {"voxels": t.nvoxels, "fit": "T1 T2 B1 rho" if fit_b1 else "T1 T2 rho", **runs}
    c = out["cartesian"]["graph_device_cg"]
    print(json.dumps({
        "metric": "Gauss-Newton iterations/s (Cartesian FISP fit, 4096 voxels x 256 TRs, normal operator from a captured graph, "
                  "CG on the device)",
        "value": c["iterations"] / c["seconds"],
        "unit": "iterations/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f32 (products BF16 hi+lo on MFMA; derivatives by the fused tangent passes)",
        "data": "synthetic (seeded phantoms, noise-free k-space data, perturbed starts)",
        "config": {"workload": "Gauss-Newton: Cartesian FISP 4096 voxels x 256 TRs (12 iterations, CG <= 100) and spiral pSSFP "
                               "3136 voxels x 280 TRs x 35 isochromats with B1 (8 iterations, CG <= 100)"},
        "fits": out,
    }))


def _objects(wl):
    params = ca.TissueParameters.from_tissue(wl.tissue)
    seq = (ca.FispSequence.from_sequence(wl.sequence) if isinstance(wl.sequence, W.FispSeq)
           else ca.pSSFPSequence.from_sequence(wl.sequence))
    traj = (ca.CartesianTrajectory.from_trajectory(wl.trajectory) if isinstance(wl.trajectory, W.CartesianTraj)
            else ca.SpiralTrajectory.from_trajectory(wl.trajectory))
    return params, seq, traj


def calibration_size(args, n_req):
    if args.calibrate is not None:
        return args.calibrate
    return max(60, min(600, n_req // 10))


def progress(msg):
    """A progress line on stderr (long serving runs would otherwise be silent)."""
    print(f"[serve] {msg}", file=sys.stderr, flush=True)


def serve_bench(args):
    """Serving loop: requests with Poisson arrivals at `--load` x the measured
    capacity, continuous batching on `--slots` slots, p50/p99 latency (arrival
    -> result ready in HBM) and aggregate throughput; a seeded sample of the
    requests is re-run alone and compared with its result in the mix."""
    import threading

    ctx = ca.init_context(0)
    n_req = args.requests
    specs = [serve_request_spec(i) for i in range(n_req)]
    t0 = time.perf_counter()
    cat = ServeCatalog()
    reqs = []
    for i, sp in enumerate(specs):
        reqs.append(cat.request(sp))
        if (i + 1) % 1000 == 0:
            progress(f"set-up: {i + 1} of {n_req} requests resident")
    ctx.synchronize()
    t_setup = time.perf_counter() - t0
    progress(f"set-up done in {t_setup:.1f} s")
    check_ids = set(np.random.default_rng(5).choice(n_req, size=min(args.check, n_req), replace=False).tolist())
    budget = int(args.budget_gib * (1 << 30))

    # capacity: a closed-loop burst of a seeded sample of the requests (also warms up the
    # slot pools); request costs are heavy-tailed, so the sample is 10% of the run
    cal_ids = np.random.default_rng(6).choice(n_req, size=min(calibration_size(args, n_req), n_req),
                                              replace=False).tolist()
    server = ca.Server(slots=args.slots, memory_budget=budget)
    tickets = [serve_submit(server, reqs[i]) for i in cal_ids]
    # results are taken in completion order: a finished request's results
    # count against the server's memory budget until released
    times = []
    outstanding = list(tickets)
    while outstanding:
        still = []
        for t in outstanding:
            tm = server.wait(t, timeout=0.0)
            if tm is None:
                still.append(t)
            else:
                times.append(tm)
                server.release(t)
        if len(still) == len(outstanding):
            time.sleep(0.001)
        outstanding = still
    cal_span = max(d for _, _, d in times) - min(s for _, s, _ in times)
    capacity = len(cal_ids) / cal_span
    lam = args.rate if args.rate else args.load * capacity
    progress(f"capacity {capacity:.1f} requests/s; arrivals at {lam:.1f}/s")

    # open loop: Poisson arrivals; a collector thread takes results in
    # completion order (held results count against the memory budget)
    gaps = np.random.default_rng(7).exponential(1.0 / lam, size=n_req)
    arrivals = np.cumsum(gaps)
    results = {}
    stamps = [None] * n_req
    tickets = [None] * n_req
    submitted = threading.Semaphore(0)
    errors = []

    def collector():
        outstanding, received = [], 0
        while received < n_req or outstanding:
            while submitted.acquire(blocking=not outstanding):
                outstanding.append(received)
                received += 1
                if received == n_req:
                    break
            still = []
            for i in outstanding:
                try:
                    tm = server.wait(tickets[i], timeout=0.0)
                    if tm is None:
                        still.append(i)
                        continue
                    stamps[i] = tm
                    if i in check_ids:
                        results[i] = server.result(tickets[i]).collect()
                    server.release(tickets[i])
                except Exception as e:  # reported in the JSON line
                    errors.append(f"request {i}: {e}")
            if still and len(still) == len(outstanding):
                time.sleep(0.001)
            outstanding = still

    th = threading.Thread(target=collector)
    th.start()
    base = server.now() + 0.05
    for i in range(n_req):
        delay = base + arrivals[i] - server.now()
        if delay > 0:
            time.sleep(delay)
        tickets[i] = serve_submit(server, reqs[i])
        submitted.release()
        if (i + 1) % 500 == 0:
            progress(f"submitted {i + 1} of {n_req}")
    th.join()
    progress("all requests done; re-running the checked ones alone")
    batching = server.batch_stats()  # calibration burst + open loop
    server.close()

    ok = [s for s in stamps if s is not None]
    lat = np.array([d - a for a, _, d in ok])
    queue = np.array([s - a for a, s, _ in ok])
    span = max(d for _, _, d in ok) - min(a for a, _, _ in ok)
    vox_tr = sum(sp["nv"] * sp["nTR"] for sp in specs)
    kspace = sum(sp["nc"] * sp["nTR"] * SERVE_NS for sp in specs)

    # the same requests alone: one slot, one request in flight at a time
    alone = ca.Server(slots=1, memory_budget=budget)
    worst = 0.0
    for i in sorted(check_ids):
        t = serve_submit(alone, reqs[i])
        alone.wait(t)
        out = alone.result(t).collect()
        alone.release(t)
        if i in results:
            worst = max(worst, rel_l2(results[i], out))
    alone.close()

    counts = {k: sum(1 for sp in specs if sp["kind"] == k) for k in ("fisp", "pssfp", "jhy")}

    def pct(idx):
        la = np.array([stamps[i][2] - stamps[i][0] for i in idx])
        qu = np.array([stamps[i][1] - stamps[i][0] for i in idx])
        return {"n": len(idx), "p50_ms": 1e3 * float(np.percentile(la, 50)),
                "p99_ms": 1e3 * float(np.percentile(la, 99)), "queue_p99_ms": 1e3 * float(np.percentile(qu, 99))}

    done_ids = [i for i in range(n_req) if stamps[i] is not None]
    by_kind = {k: pct([i for i in done_ids if specs[i]["kind"] == k]) for k in counts if counts[k]}
    work = np.array([specs[i]["nv"] * specs[i]["nTR"] for i in done_ids], dtype=np.float64)
    edges = np.percentile(work, [25, 50, 75]) if len(work) else []
    quart = np.searchsorted(edges, work, side="right") if len(work) else []
    by_size = {f"q{q + 1}": pct([i for i, qq in zip(done_ids, quart) if qq == q]) for q in range(4)
               if any(qq == q for qq in quart)}
    line = {
        "metric": "serving: aggregate voxel*TR/s and k-space samples/s with p50/p99 latency (arrival -> result in HBM)",
        "value": vox_tr / span,
        "unit": "voxel*TR/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f32 (signal and J^H.y contractions BF16 hi+lo on MFMA)",
        "data": "synthetic (seeded request mix; resident tissue maps, sequences, trajectories, coils)",
        "config": {"workload": f"{n_req} requests, Poisson arrivals at "
                               + (f"{lam:.1f}/s" if args.rate else f"{args.load:.0%} of measured capacity")
                               + f", {args.slots} slots, budget {args.budget_gib:g} GiB",
                   "mix": counts, "samples_per_readout": SERVE_NS},
        "requests_per_s": len(ok) / span,
        "kspace_samples_per_s": kspace / span,
        "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99)),
                       "mean": 1e3 * float(lat.mean())},
        "queue_ms": {"p50": 1e3 * float(np.percentile(queue, 50)), "p99": 1e3 * float(np.percentile(queue, 99))},
        "latency_by_kind": by_kind,
        "latency_by_size": dict(by_size, note="quartiles of voxels x TRs per request"),
        "arrival_rate_per_s": lam,
        "capacity_per_s": capacity,
        "calibration": f"{len(cal_ids)} requests submitted at once: {cal_span:.2f} s",
        "alone_vs_mix_max_rel_l2": worst,
        "alone_vs_mix_checked": len(check_ids),
        "batching": dict(batching, requests=len(cal_ids) + n_req,
                         note="work-table simulation launches (FISP voxel blocks, pSSFP voxel-block x "
                              "slice-group items) over the calibration burst and the run; shared_launches "
                              "served two or more requests; queue_ms is arrival -> admission"),
        "errors": errors[:5],
        "setup_s": t_setup,
    }
    print(json.dumps(line))


def rel_l2(x, ref):
    x = np.asarray(x, dtype=np.complex128)
    ref = np.asarray(ref, dtype=np.complex128)
    den = np.linalg.norm(ref.ravel())
    return float(np.linalg.norm((x - ref).ravel()) / (den if den > 0 else 1.0))


def precision_study(args):
    """Contraction precisions side by side: FP32 VALU, BF16 hi+lo, FP16 hi+lo
    (per-row power-of-two scaling), BF16 x1 and FP16 x1.  For the 4-coil
    Cartesian shapes (1120 readouts x 224 samples x 50176 voxels) it times
    signal, J.v and J^H.y per precision (HIP events around the library's
    launches) and gives rel-L2 against the FP64 oracle on independent subsets
    (readouts for signal and J.v, voxels for J^H.y); the 4096-voxel
    single-coil signal (256 readouts x 64 samples) is checked in full."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    pyoracle.build()
    threads = int(os.environ.get("ORACLE_THREADS", os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    ctx = ca.init_context(0)

    def timed(fn, reps=3):
        fn()  # first call (autotuning)
        ctx.synchronize()
        ctx.read_kernel_timings()
        ctx.set_kernel_timing(True)
        for _ in range(reps):
            out = fn()
        ctx.synchronize()
        ctx.set_kernel_timing(False)
        k = ctx.read_kernel_timings()
        return out.collect(), sum(tot for c, tot in k.values()) / reps

    # small case: 4096 voxels, 1 coil, 256 x 64 samples, full output checked
    w1 = W.config(1)
    p1 = ca.TissueParameters.from_tissue(w1.tissue)
    s1 = ca.FispSequence.from_sequence(w1.sequence)
    t1 = ca.CartesianTrajectory.from_trajectory(w1.trajectory)
    c1 = ca.CompasArray.from_numpy(w1.coils)
    e1 = ca.simulate_magnetization(p1, s1)
    ref1 = pyoracle.signal_cartesian(w1.tissue, w1.trajectory, w1.coils, e1.collect().astype(np.complex128), "f64",
                                     threads)

    wl = W.config(2)
    nv = wl.tissue.nvoxels
    nr, ns = wl.trajectory.nreadouts, wl.trajectory.samples_per_readout
    nc = wl.coils.shape[0]
    params = ca.TissueParameters.from_tissue(wl.tissue)
    seq = ca.FispSequence.from_sequence(wl.sequence)
    traj = ca.CartesianTrajectory.from_trajectory(wl.trajectory)
    coils = ca.CompasArray.from_numpy(wl.coils)
    echos = ca.simulate_magnetization(params, seq)
    d = ca.simulate_magnetization_derivatives(echos, params, seq)
    rng = W.SplitMix64(0xC0A50004)
    v = (rng.normal(4 * nv) + 1j * rng.normal(4 * nv)).reshape(4, nv).astype(np.complex64)
    y = ((rng.normal(nc * nr * ns) + 1j * rng.normal(nc * nr * ns)) / np.sqrt(2)).reshape(nc, nr, ns).astype(
        np.complex64)
    va, ya = ca.CompasArray.from_numpy(v), ca.CompasArray.from_numpy(y)
    e = echos.collect()
    d1, d2 = d["T1"].collect(), d["T2"].collect()

    ridx = np.linspace(0, nr - 1, 6).astype(int)
    sub = W.CartesianTraj(len(ridx), ns, wl.trajectory.delta_t, wl.trajectory.k_start[ridx], wl.trajectory.delta_k)
    ref_sig = pyoracle.signal_cartesian(wl.tissue, sub, wl.coils, e[ridx].astype(np.complex128), "f64", threads)
    ref_jv = pyoracle.jacobian(wl.tissue, sub, wl.coils, e[ridx], d1[ridx], d2[ridx], v, "f64", threads)
    vsub = np.arange(0, nv, nv // 96)[:96]
    tsub = wl.tissue.replace(**{f: getattr(wl.tissue, f)[vsub] for f in ("T1", "T2", "B1", "B0", "rho", "x", "y", "z")})
    ref_jh = pyoracle.jacobian_hermitian(tsub, wl.trajectory, wl.coils[:, vsub], e[:, vsub], d1[:, vsub],
                                         d2[:, vsub], y, "f64", threads)

    flops = 8.0 * nc * nr * ns * nv
    res = {}
    try:
        for prec in ca.CONTRACTION_PRECISIONS:
            ca.set_contraction_precision(prec)
            small, _ = timed(lambda: ca.magnetization_to_signal(e1, p1, t1, c1, method="matmul_fast"), reps=1)
            s, ts = timed(lambda: ca.magnetization_to_signal(echos, params, traj, coils, method="matmul_fast"))
            jv, tj = timed(lambda: ca.compute_jacobian(echos, d, params, traj, coils, va, method="gemm_fast"))
            jh, th = timed(lambda: ca.compute_jacobian_hermitian(echos, d, params, traj, coils, ya,
                                                                 method="gemm_fast"))
            res[prec] = {
                "signal_4096vox_1coil": {"rel_l2_vs_fp64": pyoracle.rel_l2(small, ref1)},
                "signal": {"ms": ts, "useful_TFLOPs": flops / (ts * 1e-3) / 1e12,
                           "rel_l2_vs_fp64": pyoracle.rel_l2(s[:, ridx, :], ref_sig)},
                "jacobian": {"ms": tj, "useful_TFLOPs": 2 * flops / (tj * 1e-3) / 1e12,
                             "rel_l2_vs_fp64": pyoracle.rel_l2(jv[:, ridx, :], ref_jv)},
                "jacobian_hermitian": {"ms": th, "useful_TFLOPs": 2 * flops / (th * 1e-3) / 1e12,
                                       "rel_l2_vs_fp64": pyoracle.rel_l2(jh[:, vsub], ref_jh)},
            }
            print(f"# {prec}: " + json.dumps(res[prec]), file=sys.stderr, flush=True)
    finally:
        ca.set_contraction_precision(None)

    def total(p):
        return sum(res[p][k]["ms"] for k in ("signal", "jacobian", "jacobian_hermitian"))

    within = [p for p in res if all(res[p][k]["rel_l2_vs_fp64"] <= 1e-4
                                    for k in ("signal", "jacobian", "jacobian_hermitian"))]
    best = min((p for p in within if p != "fp32"), key=total, default=None)
    line = {
        "metric": "signal + J.v + J^H.y speed-up of the mixed-precision MFMA path over the plain FP32 VALU path "
                  "at <=1e-4 rel-L2 vs FP64",
        "value": total("fp32") / total(best) if best else None,
        "unit": "x",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": f"f32 operands split into {best} on MFMA, f32 accumulate" if best else "f32",
        "data": "synthetic",
        "config": {"workload": f"Cartesian contractions {nc} coils x {nr} readouts x {ns} samples x {nv} voxels"},
        "best_within_1e-4": best,
        "ms_signal_jv_jhy": {p: total(p) for p in res},
        "precisions": res,
        "contraction_tuning": ca.contraction_tuning_report().strip().splitlines(),
    }
    print(json.dumps(line))


SIGNAL_DTYPE = {
    "matmul_fast": "f32 sim; bf16x3 (hi+lo split) MFMA, fp32-acc contraction",
    "matmul_pedantic": "f32 sim; fp16x3 (hi+lo split, per-row scaling) MFMA, fp32-acc contraction",
    "matmul_low": "f32 sim; bf16x1 MFMA, fp32-acc contraction",
    "matmul": "f32 sim; bf16x3 (hi+lo split) MFMA, fp32-acc contraction",
    "direct": "f32 sim; f32 VALU contraction",
    "naive": "f32 sim; f32 VALU contraction",
}


def mixed_vs_plain(ctx, wl, params, seq, traj, coils, echos, reps=5, threads=None):
    """signal + J.v + J^H.y at this workload's shapes, the default mixed-precision
    methods (BF16 hi+lo on MFMA) against the plain FP32 VALU methods (Direct):
    wall time per step (median of `reps` synchronised steps after one warm-up)
    and each output's rel-L2 against the FP64 oracle on independent subsets
    (6 readouts for signal and J.v, 96 voxels for J^H.y).  Inputs: the step's
    echoes, their finite-difference T1/T2 derivatives, a seeded [4][nv] direction
    and a seeded [nc][nr][ns] residual."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    pyoracle.build()
    threads = threads or int(os.environ.get("ORACLE_THREADS", os.environ.get("OMP_NUM_THREADS",
                                                                               os.cpu_count() or 1)))
    nv = wl.tissue.nvoxels
    nr, ns = wl.trajectory.nreadouts, wl.trajectory.samples_per_readout
    nc = wl.coils.shape[0]
    d = ca.simulate_magnetization_derivatives(echos, params, seq)
    rng = W.SplitMix64(0xC0A50004)
    v = (rng.normal(4 * nv) + 1j * rng.normal(4 * nv)).reshape(4, nv).astype(np.complex64)
    y = ((rng.normal(nc * nr * ns) + 1j * rng.normal(nc * nr * ns)) / np.sqrt(2)).reshape(nc, nr, ns).astype(
        np.complex64)
    va, ya = ca.CompasArray.from_numpy(v), ca.CompasArray.from_numpy(y)

    def run(sig_m, jac_m):
        def step():
            return (ca.magnetization_to_signal(echos, params, traj, coils, method=sig_m),
                    ca.compute_jacobian(echos, d, params, traj, coils, va, method=jac_m),
                    ca.compute_jacobian_hermitian(echos, d, params, traj, coils, ya, method=jac_m))

        step()
        ctx.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            out = step()
            ctx.synchronize()
            ts.append(time.perf_counter() - t0)
        return out, 1e3 * float(np.median(ts))

    mixed, t_mixed = run("matmul_fast", "gemm_fast")
    plain, t_plain = run("direct", "direct")
    e = echos.collect()
    d1, d2 = d["T1"].collect(), d["T2"].collect()
    ridx = np.linspace(0, nr - 1, 6).astype(int)
    sub = W.CartesianTraj(len(ridx), ns, wl.trajectory.delta_t, wl.trajectory.k_start[ridx], wl.trajectory.delta_k)
    vsub = np.linspace(0, nv - 1, 96).astype(int)
    tsub = wl.tissue.replace(**{f: getattr(wl.tissue, f)[vsub] for f in ("T1", "T2", "B1", "B0", "rho", "x", "y",
                                                                          "z")})
    ref = {
        "signal": pyoracle.signal_cartesian(wl.tissue, sub, wl.coils, e[ridx].astype(np.complex128), "f64", threads),
        "jv": pyoracle.jacobian(wl.tissue, sub, wl.coils, e[ridx], d1[ridx], d2[ridx], v, "f64", threads),
        "jhy": pyoracle.jacobian_hermitian(tsub, wl.trajectory, wl.coils[:, vsub], e[:, vsub], d1[:, vsub],
                                           d2[:, vsub], y, "f64", threads),
    }

    def errors(outs):
        s_, jv_, jh_ = (o.collect() for o in outs)
        return {"signal": pyoracle.rel_l2(s_[:, ridx, :], ref["signal"]),
                "jv": pyoracle.rel_l2(jv_[:, ridx, :], ref["jv"]),
                "jhy": pyoracle.rel_l2(jh_[:, vsub], ref["jhy"])}

    return {
        "mixed_ms": t_mixed, "plain_fp32_ms": t_plain, "speedup": t_plain / t_mixed,
        "mixed_rel_l2_vs_fp64": errors(mixed), "plain_fp32_rel_l2_vs_fp64": errors(plain),
        "methods": {"mixed": "matmul_fast / gemm_fast (bf16x3 MFMA, fp32 accumulate)",
                    "plain_fp32": "direct (fp32 VALU, no MFMA)"},
        "kspace_samples_per_s": {"mixed": 3 * nc * nr * ns / (t_mixed * 1e-3),
                                 "plain_fp32": 3 * nc * nr * ns / (t_plain * 1e-3)},
        "note": f"one step = signal + J.v + J^H.y ({nc} coils x {nr} readouts x {ns} samples x {nv} voxels), "
                f"median of {reps} synchronised steps; errors on readouts {ridx.tolist()} (signal, J.v) and "
                f"96 strided voxels (J^H.y)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", type=int, default=2, choices=[1, 2, 3])
    ap.add_argument("--method", default="matmul_fast", choices=list(ca.SIGNAL_METHODS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-products", action="store_true",
                    help="skip the signal + J.v + J^H.y mixed-vs-plain comparison of the default run")
    ap.add_argument("--pipeline", type=int, default=3,
                    help="request-server slots the timed steps are spread over (1 = one stream, no overlap)")
    ap.add_argument("--serve", action="store_true", help="serving loop with Poisson arrivals")
    ap.add_argument("--requests", type=int, default=600)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--load", type=float, default=0.7)
    ap.add_argument("--rate", type=float, default=None,
                    help="with --serve: Poisson arrivals at this many requests/s instead of --load x capacity")
    ap.add_argument("--calibrate", type=int, default=None,
                    help="requests in the capacity burst (default: 10%% of --requests, 60 to 600)")
    ap.add_argument("--check", type=int, default=100)
    ap.add_argument("--budget-gib", type=float, default=96.0)
    ap.add_argument("--wire", action="store_true", help="with --serve: requests as wire frames over TCP")
    ap.add_argument("--connections", type=int, default=8, help="TCP connections of --serve --wire")
    ap.add_argument("--products", action="store_true",
                    help="time signal + J.v + J^H.y at the 4-coil Cartesian shapes (--method picks the precision)")
    ap.add_argument("--reconstruct", action="store_true",
                    help="Gauss-Newton fits (Cartesian FISP, spiral pSSFP), graph replay vs eager")
    ap.add_argument("--precision-study", action="store_true",
                    help="time signal, J.v and J^H.y per contraction method with their error vs FP64")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.precision_study:
        precision_study(args)
        return
    if args.reconstruct:
        reconstruct_bench(args)
        return
    if args.products:
        products_bench(args)
        return
    if args.serve:
        (serve_wire_bench if args.wire else serve_bench)(args)
        return
    if args.workload == 3:
        spiral_bench(args)
        return

    wl = W.config(args.workload)
    nv = wl.tissue.nvoxels
    nTR = int(wl.sequence.RF.shape[0])
    nr, ns = wl.trajectory.nreadouts, wl.trajectory.samples_per_readout
    nc = wl.coils.shape[0]

    ctx = ca.init_context(0)
    params = ca.TissueParameters.from_tissue(wl.tissue)
    seq = ca.FispSequence.from_sequence(wl.sequence)
    traj = ca.CartesianTrajectory.from_trajectory(wl.trajectory)
    coils = ca.CompasArray.from_numpy(wl.coils)
    ctx.synchronize()

    def step():
        echos = ca.simulate_magnetization(params, seq)
        signal = ca.magnetization_to_signal(echos, params, traj, coils, method=args.method)
        return echos, signal

    for _ in range(args.warmup):
        out = step()
    ctx.synchronize()
    del out

    def kernel_table(timings):
        return {k: {"launches": c, "avg_ms": tot / max(c, 1), "total_ms": tot} for k, (c, tot) in timings.items()}

    def roofline_of(kernels):
        dominant = max(kernels, key=lambda k: kernels[k]["total_ms"]) if kernels else None
        if dominant is None:
            return None
        work = algorithmic_work(dominant, wl, args.method)
        if work is None:
            return None
        amount, unit, bound, peak = work
        achieved = amount / (kernels[dominant]["avg_ms"] / 1e3)
        return {"kernel": dominant, "bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
                "frac": achieved / peak, "traffic": measured_traffic(dominant, args.workload)}

    # ---- sequential steps on the library's stream: per-step latency of one step on an
    # otherwise idle GPU, and kernel durations without overlap (HIP events per launch)
    ctx.read_kernel_timings()  # reset
    ctx.set_kernel_timing(True)
    lat = []
    ctx.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        out = step()
        ctx.synchronize()
        lat.append(time.perf_counter() - t0)
    t_seq = time.perf_counter() - t_start
    ctx.set_kernel_timing(False)
    kernels_seq = kernel_table(ctx.read_kernel_timings())
    sequential = {"ms_per_step": 1e3 * t_seq / args.steps, "value": nv * nTR * args.steps / t_seq,
                  "kernels": kernels_seq, "roofline": roofline_of(kernels_seq)}

    # ---- the same K steps as independent requests on the request server: consecutive
    # steps overlap on the device (the VALU-bound EPG of one step beside the MFMA
    # contraction of another).  This is the timed region of `value`.
    pipeline = None
    t_total, kernels, roofline = t_seq, kernels_seq, sequential["roofline"]
    if args.pipeline > 1:
        srv = ca.Server(slots=args.pipeline, memory_budget=64 << 30)
        for _ in range(max(args.warmup, 2 * args.pipeline)):
            t = srv.submit_simulate_signal(params, seq, traj, coils, method=args.method)
            srv.wait(t)
            srv.release(t)
        srv.read_kernel_timings()  # reset
        srv.set_kernel_timing(True)
        t_start = time.perf_counter()
        tickets = [srv.submit_simulate_signal(params, seq, traj, coils, method=args.method)
                   for _ in range(args.steps)]
        stamps = [srv.wait(t) for t in tickets]
        t_total = time.perf_counter() - t_start
        srv.set_kernel_timing(False)
        kernels = kernel_table(srv.read_kernel_timings())
        for t in tickets:
            srv.release(t)
        srv.close()
        service = [d - st for _, st, d in stamps]
        pipeline = {"slots": args.pipeline, "ms_per_step": 1e3 * t_total / args.steps,
                    "service_ms": {"p50": 1e3 * percentile(service, 50), "p99": 1e3 * percentile(service, 99)},
                    "roofline_overlapped": roofline_of(kernels),
                    "note": "kernel durations here include the time each launch shares the GPU with the "
                            "kernels of other steps"}
    # the dominant kernel's roofline uses its durations from the sequential phase:
    # one launch at a time, so the duration is the kernel's own
    if roofline is not None:
        roofline = dict(roofline, phase="sequential")

    ms_per_step = 1e3 * t_total / args.steps
    value = nv * nTR * args.steps / t_total

    # the same sequential step with the plain FP32 contraction (no MFMA)
    plain_step = None
    if args.method != "direct":
        def plain():
            e_ = ca.simulate_magnetization(params, seq)
            return ca.magnetization_to_signal(e_, params, traj, coils, method="direct")

        plain()
        ctx.synchronize()
        ts = []
        for _ in range(5):
            t0 = time.perf_counter()
            plain_out = plain()
            ctx.synchronize()
            ts.append(time.perf_counter() - t0)
        plain_step = {"ms_per_step": 1e3 * float(np.median(ts)), "method": "direct (fp32 VALU contraction)",
                      "value": nv * nTR / float(np.median(ts)), "note": "median of 5 sequential steps"}

    # plain streaming copy of known size: achievable HBM bandwidth, and the
    # reference dispatch that scales FETCH_SIZE / WRITE_SIZE counters to bytes
    probe_bytes = 1 << 30
    probe_ms = ctx.bandwidth_probe(probe_bytes)

    echos, signal = out
    line = {
        "metric": "voxel·TR/s sim + k-space samples/s; p50/p99 latency at ≤1e-4 rel-L2 error",
        "value": value,
        "unit": "voxel*TR/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": SIGNAL_DTYPE.get(args.method, "f32"),
        "data": "synthetic (seeded SplitMix64 tissue maps, sinusoidal flip train, Gaussian coil maps)",
        "config": {
            "workload": f"FISP {nv} voxels x {nTR} TRs x {wl.sequence.max_state} states + Cartesian signal "
                        f"{nc} coils x {nr} readouts x {ns} samples",
            "signal_method": args.method,
        },
        "kspace_samples_per_s": nc * nr * ns * args.steps / t_total,
        "latency_ms": {"p50": 1e3 * percentile(lat, 50), "p99": 1e3 * percentile(lat, 99),
                       "note": "one step alone on an idle GPU (sequential phase)"},
        "pipeline": pipeline,
        "sequential": sequential,
        "kernels": kernels,
        "hbm_copy_GBs": 2 * probe_bytes / (probe_ms * 1e-3) / 1e9,
        "contraction_tuning": ca.contraction_tuning_report().strip().splitlines(),
        "roofline": roofline,
        "plain_fp32_step": plain_step,
    }
    if args.method != "direct" and not args.no_products:
        line["products_mixed_vs_plain"] = mixed_vs_plain(ctx, wl, params, seq, traj, coils, echos)
    if not args.no_cpu_baseline:
        e = echos.collect()
        s = signal.collect()
        others = {"plain_fp32": plain_out.collect()} if plain_step is not None else None
        base, acc = cpu_baseline(wl, e, s, other_signals=others)
        line["cpu_baseline"] = base
        line["accuracy"] = acc
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line))


if __name__ == "__main__":
    main()
