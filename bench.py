#!/usr/bin/env python
"""Benchmark of the randomisation-method summation on one MI355X.

Workload (the headline configuration): unstructured 3D Exponential model,
var 2, len_scale [50, 20, 5], angles [0.5, 0.2, 0.1], mode_no 1024 (modes
drawn by the reference sampler restated in gstools_amd, seed 20170519),
4,000,000 points uniform in [0, 1000]^3 (RandomState(19970221)), FP64
coordinates isometrized on the host.  A step is one pass of the summation
over all points; ``value`` is mode.points/s of the FP32 path (FP64 phase,
FP32 cosine, blocked FP32->FP64 accumulation), with inputs resident in HBM.

Also reported: the FP64 plain path on the same inputs, the FP32 error
against it, the FP64 path's parity against the CPU oracle on a point
subsample, the dominant kernel's roofline from HIP events, and the CPU
baseline (the oracle's C/OpenMP restatement of the reference loop) timed on
a bounded sample of the same workload.

Extra sections (``--extras``, on by default): the 100 x 100 2D Gaussian
grid (N 1000, FP64), structured 512^3 Matern grid on the matrix cores, the
incompressible 128^3 grid, the periodic Fourier method on a 4096^2 grid,
ordinary kriging sums (1000 conditioning points onto 1e6 targets), variogram
pair binning (1e6 points), forward + adjoint on the 2D Stable configuration
and the 256-realization batch; each prints its own JSON object on stderr, and
the sections of the reference's configurations carry the CPU restatement
timed on a bounded sample (scaled linearly: the work is linear in the points).
"""

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "mode·points/s FP32/MFMA vs FP64 plain at ≤1e-5 err vs FP64; fwd+bwd time"
P_HEADLINE = 4_000_000
N_HEADLINE = 1024
FP32_VALU_PEAK_TFLOPS = 157.3
FP64_VALU_PEAK_TFLOPS = 78.6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def headline_inputs():
    from gstools_amd import Exponential, RandMeth

    model = Exponential(dim=3, var=2.0, len_scale=[50, 20, 5], angles=[0.5, 0.2, 0.1])
    rm = RandMeth(model, mode_no=N_HEADLINE, seed=20170519)
    pos = np.random.RandomState(19970221).uniform(0.0, 1000.0, size=(3, P_HEADLINE))
    iso = model.isometrize(pos)
    return rm, iso


def time_kernel(fn, steps, warmup, torch):
    """Warm up, then time `steps` calls with wall clock and per-call HIP events."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        starts[i].record(stream)
        fn()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    return wall / steps, float(np.mean(ev)) / 1e3


def host_cpu():
    """Model name and clock of the host CPU (what ``lscpu`` reports), from /proc/cpuinfo."""
    model, mhz = None, []
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                key, _, val = line.partition(":")
                key = key.strip()
                if key == "model name" and model is None:
                    model = val.strip()
                elif key == "cpu MHz":
                    mhz.append(float(val))
    except OSError:
        pass
    return {"model": model, "mhz_mean": float(np.mean(mhz)) if mhz else None,
            "logical_cpus_on_host": os.cpu_count()}


def _oracle_rate(cov, z1, z2, iso, nthreads, target_s):
    """mode.points/s of the C oracle on the first points of ``iso`` that take ~target_s."""
    import oracle

    n_modes = cov.shape[1]
    probe = 20000 if nthreads > 1 else 2000
    t0 = time.perf_counter()
    oracle.summate_c(cov, z1, z2, np.ascontiguousarray(iso[:, :probe]), nthreads=nthreads)
    dt = time.perf_counter() - t0
    rate = probe * n_modes / dt
    npts = int(min(iso.shape[1], max(probe, rate * target_s / n_modes)))
    t0 = time.perf_counter()
    oracle.summate_c(cov, z1, z2, np.ascontiguousarray(iso[:, :npts]), nthreads=nthreads)
    dt = time.perf_counter() - t0
    return npts * n_modes / dt, npts, dt


def cpu_baseline(rm, iso, target_s=10.0, single_core_s=5.0):
    """The oracle's C/OpenMP restatement of the reference loop on the host:
    all cores of this process's affinity mask, then one core."""
    cores = len(os.sched_getaffinity(0))
    cov, z1, z2 = rm._cov_sample, rm._z_1, rm._z_2
    rate, npts, dt = _oracle_rate(cov, z1, z2, iso, cores, target_s)
    rate1, npts1, dt1 = _oracle_rate(cov, z1, z2, iso, 1, single_core_s)
    cpu = host_cpu()
    return {
        "value": rate,
        "unit": "mode·points/s",
        "cores": cores,
        "kind": "port",
        "sample": f"first {npts} of {P_HEADLINE} points x {N_HEADLINE} modes, FP64, "
                  f"C/OpenMP restatement of the reference loop, {cores} threads, {dt:.1f} s",
        "single_core_value": rate1,
        "single_core_sample": f"first {npts1} points x {N_HEADLINE} modes, 1 thread, {dt1:.1f} s",
        "cpu_model": cpu["model"],
        "cpu_mhz_mean": cpu["mhz_mean"],
        "logical_cpus_on_host": cpu["logical_cpus_on_host"],
    }


def _cpu_restatement(fn, cov, z1, z2, pos, *extra, target_s=2.0):
    """The oracle's C/OpenMP restatement of the reference loop (``fn``: summate_c,
    summate_incompr_c or summate_adjoint_c) on all cores of this process's
    affinity mask, over the first points of ``pos`` that take about target_s;
    returns (mode.points/s, points timed, seconds, threads)."""
    cores = len(os.sched_getaffinity(0))
    n_modes = cov.shape[1]
    cut = lambda n: [np.ascontiguousarray(a[..., :n]) for a in (pos,) + extra]
    probe = min(pos.shape[1], 4000)
    t0 = time.perf_counter()
    fn(cov, z1, z2, *cut(probe), nthreads=cores)
    rate = probe * n_modes / (time.perf_counter() - t0)
    npts = int(min(pos.shape[1], max(probe, rate * target_s / n_modes)))
    args = cut(npts)
    t0 = time.perf_counter()
    fn(cov, z1, z2, *args, nthreads=cores)
    dt = time.perf_counter() - t0
    return npts * n_modes / dt, npts, dt, cores


MFMA_PEAK_TFLOPS = 2500.0  # dense bf16 / fp16


def bench_config1(torch, steps, warmup):
    """2D Gaussian var 1 len 10, mode_no 1000 (inversion sampler, seed 20170519) on the
    100 x 100 grid x, y = 0..99, FP64: the device sum, the numpy drop-in and the
    CPU restatement of the reference loop."""
    import oracle
    from gstools_amd import Gaussian, RandMeth, summate_device

    log("bench: 100 x 100 grid, N 1000, fp64")
    model = Gaussian(dim=2, var=1.0, len_scale=10.0)
    rm = RandMeth(model, mode_no=1000, seed=20170519)
    x = np.arange(100.0)
    pos = np.ascontiguousarray(np.asarray(np.meshgrid(x, x, indexing="ij")).reshape(2, -1))
    dev = torch.device("cuda", 0)
    cov, z1, z2, post = (torch.from_numpy(a).to(dev) for a in (rm._cov_sample, rm._z_1, rm._z_2, pos))
    out = torch.empty(pos.shape[1], dtype=torch.float64, device=dev)
    work = float(pos.shape[1]) * 1000
    wall, ker = time_kernel(lambda: summate_device(cov, z1, z2, post, "fp64", out=out), steps, warmup, torch)
    rm(pos)  # numpy in, numpy out (the reference's call), after a warm-up
    t0 = time.perf_counter()
    for _ in range(steps):
        field = rm(pos)
    wall_np = (time.perf_counter() - t0) / steps
    cores = len(os.sched_getaffinity(0))
    c = (rm._cov_sample, rm._z_1, rm._z_2, pos)
    ref = oracle.summate_c(*c, nthreads=cores)
    t0 = time.perf_counter()
    for _ in range(3):
        oracle.summate_c(*c, nthreads=cores)
    t_all = (time.perf_counter() - t0) / 3
    t0 = time.perf_counter()
    oracle.summate_c(*c, nthreads=1)
    t_one = time.perf_counter() - t0
    scale = float(np.sqrt(model.var / 1000))
    return {
        "workload": "2D Gaussian var 1 len 10 (len_rescaled 11.284), mode_no 1000, inversion sampler, seed "
                    "20170519, structured 100 x 100 (x, y = 0..99), fp64",
        "mode_points": work,
        "device_ms": wall * 1e3, "device_value": work / wall,
        "numpy_dropin_ms": wall_np * 1e3,
        "parity_vs_oracle_rel_max": oracle.rel_errors(out.cpu().numpy(), ref)[0],
        "dropin_vs_oracle_rel_max": oracle.rel_errors(field, scale * ref)[0],
        "cpu_restatement": {"value": work / t_all, "unit": "mode·points/s", "cores": cores,
                            "ms": t_all * 1e3, "single_core_value": work / t_one, "single_core_ms": t_one * 1e3,
                            "sample": "the whole grid, FP64"},
    }


def bench_mfma_grid(torch, steps, warmup):
    """512^3 Matern(nu=1.5, len 8) grid at spacing 0.5, N 4096: split bf16 / fp16 on MFMA."""
    import oracle
    from gstools_amd import Matern, RandMeth, summate_grid_device

    log("bench: mfma grid (512^3, N 4096)")
    model = Matern(dim=3, var=1.0, len_scale=8.0, nu=1.5)
    rm = RandMeth(model, mode_no=4096, seed=20170519)
    dev = torch.device("cuda", 0)
    cov = torch.from_numpy(rm._cov_sample).to(dev)
    z1 = torch.from_numpy(rm._z_1).to(dev)
    z2 = torch.from_numpy(rm._z_2).to(dev)
    ax = np.arange(512) * 0.5
    axes = [torch.from_numpy(ax).to(dev)] * 3
    n = 512**3
    work = float(n) * 4096
    out = torch.empty(n, dtype=torch.float32, device=dev)
    scale = float(np.sqrt(model.var / 4096))
    zsub = np.array([0, 101, 257, 511])
    ref_axes = [axes[0], axes[1], torch.from_numpy(ax[zsub]).to(dev)]
    ref = summate_grid_device(cov, z1, z2, ref_axes, "fp64").reshape(512, 512, len(zsub)).cpu().numpy() * scale
    res = {"workload": "structured 512^3 Matern nu=1.5 len 8, spacing 0.5, mode_no 4096 (MCMC modes)",
           "points": n, "modes": 4096}
    for prec in ("bf16x3", "fp16x3"):
        fn = lambda: summate_grid_device(cov, z1, z2, axes, prec, alpha=scale, out=out)
        wall, ker = time_kernel(fn, steps, warmup, torch)
        got = out.reshape(512, 512, 512)[:, :, zsub].cpu().numpy().astype(np.float64)
        rel_max, rel_rms = oracle.rel_errors(got, ref)
        tflops = 12.0 * work / ker / 1e12
        res[prec] = {"value": work / wall, "ms_per_step": wall * 1e3, "rel_max_vs_fp64": rel_max,
                     "rel_rms_vs_fp64": rel_rms, "mfma_tflops": tflops, "roofline_frac": tflops / MFMA_PEAK_TFLOPS}
    # FP64 direct sum on a 64-plane slab (the work is exactly linear in the points)
    slab = [axes[0][:64], axes[1], axes[2]]
    f64 = lambda: summate_grid_device(cov, z1, z2, slab, "fp64")
    wall64, ker64 = time_kernel(f64, 2, 1, torch)
    res["fp64_direct"] = {"value": 64 * 512 * 512 * 4096 / wall64, "measured_on": "64 of 512 x-planes"}
    res["speedup_bf16x3_vs_fp64"] = res["bf16x3"]["value"] / res["fp64_direct"]["value"]
    res["speedup_fp16x3_vs_fp64"] = res["fp16x3"]["value"] / res["fp64_direct"]["value"]
    # CPU restatement on the x = 0 plane (the work is exactly linear in the points)
    plane = np.asarray(np.meshgrid(ax[:1], ax, ax, indexing="ij")).reshape(3, -1)
    rate, npts, dt, cores = _cpu_restatement(oracle.summate_c, rm._cov_sample, rm._z_1, rm._z_2, plane)
    res["cpu_restatement"] = {"value": rate, "unit": "mode·points/s", "cores": cores,
                              "sample": f"first {npts} points of the x = 0 plane x 4096 modes, FP64, {dt:.1f} s",
                              "full_field_estimate_s": work / rate,
                              "speedup_bf16x3_vs_cpu": res["bf16x3"]["value"] / rate}
    # FP64 GEMM form (precision "fp64gemm"): 32 column chunks of 8192 columns
    out64 = torch.empty(n, dtype=torch.float64, device=dev)
    fg = lambda: summate_grid_device(cov, z1, z2, axes, "fp64gemm", out=out64)
    wallg, kerg = time_kernel(fg, max(2, steps // 4), 1, torch)
    got = out64.reshape(512, 512, 512)[:, :, zsub].cpu().numpy() * scale
    res["fp64gemm"] = {"value": work / wallg, "ms_per_step": wallg * 1e3,
                       "rel_max_vs_fp64": oracle.rel_errors(got, ref)[0],
                       "gemm_tflops": 4.0 * work / kerg / 1e12,
                       "speedup_vs_fp64_direct": (work / wallg) / res["fp64_direct"]["value"]}
    return res


def bench_incompr_grid(torch, steps, warmup):
    """One 128^3 incompressible Gaussian grid (a batch item of the mixed workload), N 4096:
    the split-bf16 grid path (one matrix-core sum per component) against the FP32
    incompressible sum on the materialised grid."""
    import oracle
    from gstools_amd import Gaussian, IncomprRandMeth, summate_incompr_device

    log("bench: incompressible grid (128^3, N 4096)")
    model = Gaussian(dim=3, var=1.0, len_scale=10.0)
    gen = IncomprRandMeth(model, mode_no=4096, seed=20170519, mean_velocity=1.0)
    dev = torch.device("cuda", 0)
    ax = np.arange(128, dtype=np.double)
    axes = [ax] * 3
    n = 128**3
    work = float(n) * 4096
    grid = _grid_pos(torch, [torch.from_numpy(ax).to(dev)] * 3)
    cov, z1, z2 = (torch.from_numpy(a).to(dev) for a in (gen._cov_sample, gen._z_1, gen._z_2))
    scale = float(np.sqrt(model.var / 4096))
    res = {"workload": "incompressible 3D Gaussian var 1 len 10, grid 128^3 spacing 1, mode_no 4096 (MCMC modes)",
           "points": n, "modes": 4096, "components": 3}
    wall_bf, _ = time_kernel(lambda: gen.structured(axes, precision="bf16x3"), steps, warmup, torch)
    out_bf = gen.structured(axes, precision="bf16x3").cpu().numpy().astype(np.float64)
    wall_32, _ = time_kernel(lambda: summate_incompr_device(cov, z1, z2, grid, "fp32"), steps, warmup, torch)
    out_64 = summate_incompr_device(cov, z1, z2, grid, "fp64").cpu().numpy() * scale
    out_bf[0] -= 1.0
    rel = [oracle.rel_errors(out_bf[c].ravel(), out_64[c]) for c in range(3)]
    res["bf16x3_structured"] = {"value": work / wall_bf, "ms_per_step": wall_bf * 1e3,
                                "rel_max_vs_fp64": max(r[0] for r in rel), "rel_rms_vs_fp64": max(r[1] for r in rel)}
    res["fp32_points"] = {"value": work / wall_32, "ms_per_step": wall_32 * 1e3}
    res["speedup"] = wall_32 / wall_bf
    return res


def bench_fourier(torch, steps, warmup):
    """Periodic 2D Gaussian (var 1, len 200) on a 4096^2 grid, period 4096, mode_no 256^2 (Fourier method)."""
    import oracle
    from gstools_amd import Fourier, Gaussian, summate_fourier_grid_device, summate_fourier_tensor_device

    log("bench: fourier grid (4096^2, 256^2 modes)")
    side, m = 4096, 256
    gen = Fourier(Gaussian(dim=2, var=1.0, len_scale=200.0), period=float(side), mode_no=m, seed=1681903)
    dev = torch.device("cuda", 0)
    fac, modes, z1, z2 = (torch.from_numpy(np.ascontiguousarray(a)).to(dev)
                          for a in (gen._spectrum_factor, gen._modes, gen._z_1, gen._z_2))
    ax = np.linspace(0.0, float(side), side, endpoint=False)
    axes = [torch.from_numpy(ax).to(dev)] * 2
    n, nm = side * side, m * m
    work = float(n) * nm
    out = torch.empty(n, dtype=torch.float32, device=dev)
    rows = np.array([0, 1, 1000, 2047, 4095])
    ref = summate_fourier_grid_device(fac, modes, z1, z2, [torch.from_numpy(ax[rows]).to(dev), axes[1]], "fp64")
    ref = ref.reshape(len(rows), side).cpu().numpy()
    # the oracle on a few points of those rows
    cols = np.array([0, 17, 2048, 4095])
    pts = np.stack(np.meshgrid(ax[rows], ax[cols], indexing="ij")).reshape(2, -1)
    ora = oracle.summate_fourier_c(gen._spectrum_factor, gen._modes, gen._z_1, gen._z_2, pts)
    res = {"workload": "periodic 2D Gaussian var 1 len 200, grid 4096^2 spacing 1, period 4096, "
                       "mode_no [256, 256] (Fourier method, seed 1681903)",
           "points": n, "modes": nm,
           "fp64_parity_vs_oracle_rel_max": oracle.rel_errors(ref[:, cols].ravel(), ora)[0]}
    for prec in ("bf16x3", "fp16x3"):
        fn = lambda: summate_fourier_grid_device(fac, modes, z1, z2, axes, prec, out=out)
        wall, ker = time_kernel(fn, steps, warmup, torch)
        got = out.reshape(side, side)[rows].cpu().numpy().astype(np.float64)
        rel_max, rel_rms = oracle.rel_errors(got, ref)
        tflops = 12.0 * work / ker / 1e12
        res[prec] = {"value": work / wall, "ms_per_step": wall * 1e3, "rel_max_vs_fp64": rel_max,
                     "rel_rms_vs_fp64": rel_rms, "mfma_tflops": tflops, "roofline_frac": tflops / MFMA_PEAK_TFLOPS}
    f32 = lambda: summate_fourier_grid_device(fac, modes, z1, z2, axes, "fp32", out=out)
    wall32, ker32 = time_kernel(f32, max(2, steps // 2), 1, torch)
    got = out.reshape(side, side)[rows].cpu().numpy().astype(np.float64)
    res["fp32_direct"] = {"value": work / wall32, "ms_per_step": wall32 * 1e3,
                          "rel_max_vs_fp64": oracle.rel_errors(got, ref)[0],
                          "roofline_frac": (2 * 2 + 6) * work / ker32 / 1e12 / FP32_VALU_PEAK_TFLOPS}
    slab = [axes[0][:128], axes[1]]
    f64 = lambda: summate_fourier_grid_device(fac, modes, z1, z2, slab, "fp64")
    wall64, _ = time_kernel(f64, 2, 1, torch)
    res["fp64_direct"] = {"value": 128 * side * float(nm) / wall64, "measured_on": "128 of 4096 x-rows"}
    res["speedup_bf16x3_vs_fp64"] = res["bf16x3"]["value"] / res["fp64_direct"]["value"]
    # the modes are a tensor grid: one FP64 GEMM contraction per axis (same sum, FP64)
    kax = [torch.from_numpy(k).to(dev) for k in gen._kaxes]
    out64 = torch.empty(n, dtype=torch.float64, device=dev)
    ften = lambda: summate_fourier_tensor_device(fac, kax, z1, z2, axes, out=out64)
    wallt, kert = time_kernel(ften, steps, warmup, torch)
    got = out64.reshape(side, side)[rows].cpu().numpy()
    # GEMM flops: first stage (2 side x 2 m) . (2 m x m), last stage (side x 2 m) . (2 m x side)
    flops = 8.0 * side * m * m + 4.0 * side * side * m
    res["tensor_fp64"] = {"value": work / wallt, "ms_per_step": wallt * 1e3,
                          "rel_max_vs_fp64_direct": oracle.rel_errors(got, ref)[0],
                          "rel_max_vs_oracle": oracle.rel_errors(got[:, cols].ravel(), ora)[0],
                          "gemm_flops": flops, "fp64_tflops": flops / kert / 1e12,
                          "note": "per-axis contraction (modes on a tensor grid), FP64 matrix cores"}
    res["speedup_tensor_vs_bf16x3"] = res["tensor_fp64"]["value"] / res["bf16x3"]["value"]
    # 3D: periodic Gaussian on a 256^3 grid with 64^3 modes (first, batched middle and last stage)
    side3, m3 = 256, 64
    gen3 = Fourier(Gaussian(dim=3, var=1.0, len_scale=20.0), period=float(side3), mode_no=m3, seed=1681903)
    f3, z13, z23 = (torch.from_numpy(np.ascontiguousarray(a)).to(dev)
                    for a in (gen3._spectrum_factor, gen3._z_1, gen3._z_2))
    kax3 = [torch.from_numpy(k).to(dev) for k in gen3._kaxes]
    ax3 = np.linspace(0.0, float(side3), side3, endpoint=False)
    axes3 = [torch.from_numpy(ax3).to(dev)] * 3
    out3 = torch.empty(side3 ** 3, dtype=torch.float64, device=dev)
    f3d = lambda: summate_fourier_tensor_device(f3, kax3, z13, z23, axes3, out=out3)
    wall3, ker3 = time_kernel(f3d, steps, warmup, torch)
    idx = np.array([0, 1, 77, 128, 255])
    sel = np.stack(np.meshgrid(idx, idx, idx, indexing="ij")).reshape(3, -1)
    ora3 = oracle.summate_fourier_c(gen3._spectrum_factor, gen3._modes, gen3._z_1, gen3._z_2, ax3[sel])
    got3 = out3.reshape(side3, side3, side3)[tuple(torch.from_numpy(sel).to(dev))].cpu().numpy()
    work3 = float(side3) ** 3 * m3 ** 3
    flops3 = 2.0 * (2 * side3 * m3 * m3 * 2 * m3 + side3 * 2 * side3 * m3 * 2 * m3 + side3 * side3 * side3 * 2 * m3)
    res["tensor_fp64_3d"] = {"workload": "periodic 3D Gaussian var 1 len 20, grid 256^3, period 256, mode_no 64^3",
                             "value": work3 / wall3, "ms_per_step": wall3 * 1e3,
                             "rel_max_vs_oracle_125_points": oracle.rel_errors(got3, ora3)[0],
                             "gemm_flops": flops3, "fp64_tflops": flops3 / ker3 / 1e12,
                             "per_point_bf16x3_estimate_ms": 12.0 * work3 / (res["bf16x3"]["mfma_tflops"] * 1e12) * 1e3}
    return res


def bench_krige(torch, steps, warmup):
    """Ordinary kriging sums: 3D Exponential, 1000 conditioning points onto 1e6 targets, FP64."""
    import oracle
    from gstools_amd import Exponential, krige as gkrige

    log("bench: kriging sums (K 1001, 1e6 targets)")
    model = Exponential(dim=3, var=2.0, len_scale=50.0)
    rs = np.random.RandomState(20170519)
    k_cond, n_pts = 1000, 1_000_000
    cond_pos = rs.uniform(0, 1000, size=(3, k_cond))
    cond_val = rs.normal(size=k_cond)
    krig = gkrige.Ordinary(model, cond_pos, cond_val)
    system = krig.system
    dev = torch.device("cuda", 0)
    mat_h = system.inverse
    cond_h = krig._residual()
    mat = torch.from_numpy(mat_h).to(dev)
    cond = torch.from_numpy(cond_h).to(dev)
    # right-hand sides (krige_size, P) built on the device by the front end's own
    # path: covariance rows from the device covariance kernel plus the unbiased row
    pts = rs.uniform(0, 1000, size=(3, n_pts))
    vecs = system.rhs_device(model.isometrize(pts), np.ones((1, n_pts)), krig.exact)
    kk = k_cond + 1
    flops = 2.0 * kk * kk * n_pts  # the reference's formulation, W = M V in full
    # matrix-core flops the kernel executes: row tile r (128 rows) runs its K steps
    # (16 deep) from the last one down to its diagonal block, on 128-point tiles
    jlast = ((kk - 1) // 16) * 16
    ksteps = sum((jlast - 128 * r) // 16 + 1 for r in range(-(-kk // 128)))
    mfma_flops = 2.0 * 128 * 16 * ksteps * 128 * (-(-n_pts // 128))
    both = lambda: gkrige.krige_sum_device(mat, vecs, cond, variance=True)
    wall_v, ker_v = time_kernel(both, steps, warmup, torch)
    f, e = both()
    only = lambda: gkrige.krige_sum_device(mat, vecs, cond, variance=False)
    wall_f, ker_f = time_kernel(only, steps, warmup, torch)
    sub = np.random.RandomState(4).choice(n_pts, 2000, replace=False)
    vs = vecs[:, torch.from_numpy(sub).to(dev)].cpu().numpy()
    rhs_diff = float(np.max(np.abs(vs - system.rhs_host(model.isometrize(pts[:, sub]), np.ones((1, sub.size)),
                                                             krig.exact))))
    t0 = time.perf_counter()
    f_ref, e_ref = oracle.krige_c(mat_h, vs, cond_h)
    t_cpu = time.perf_counter() - t0
    fe = f.cpu().numpy()[sub]
    ee = e.cpu().numpy()[sub]
    scale = np.abs(mat_h) @ np.abs(vs)
    tflops = mfma_flops / ker_v / 1e12  # lower bound: ker_v also holds the field pass
    return {
        "workload": "ordinary kriging sums, 3D Exponential var 2 len 50, 1000 conditioning points (+1 unbiased row) "
                    "onto 1e6 targets uniform in [0,1000]^3, fp64, krig_vecs resident in HBM",
        "cond_points": kk, "targets": n_pts,
        "field_and_variance_ms": wall_v * 1e3,
        "variance_form": "v^T M v = v^T S v, S = (M + M^T)/2: block upper triangle of S (off-diagonal blocks doubled) "
                         "on the matrix cores; field = (M^T cond) . v, one pass over V",
        "mfma_flops_executed": mfma_flops, "gemm_tflops": tflops,
        "roofline_frac_fp64_mfma": tflops / FP64_VALU_PEAK_TFLOPS,
        "reference_2k2p_equivalent_tflops": flops / ker_v / 1e12,
        "field_only_ms": wall_f * 1e3,
        "field_only_hbm_gbs": 8.0 * kk * n_pts / ker_f / 1e9,
        "max_err_over_bound_field": float(np.max(np.abs(fe - f_ref) / (np.abs(cond_h) @ scale))),
        "max_err_over_bound_error": float(np.max(np.abs(ee - e_ref) / np.sum(np.abs(vs) * scale, axis=0))),
        "device_rhs_max_abs_diff": rhs_diff,
        "cpu_oracle_points_per_s": 2000 / t_cpu,
        "gpu_points_per_s": n_pts / wall_v,
    }


def bench_variogram(torch, steps, warmup):
    """Unstructured 3D variogram of 1e6 points (20 bins), plus directional and axis-lag kernels."""
    import oracle
    from gstools_amd import variogram as gv

    log("bench: variogram pair binning (1e6 points, 20 bins)")
    rs = np.random.RandomState(19970221)
    n_pts = 1_000_000
    pos = rs.uniform(0, 1000, size=(3, n_pts))
    field = rs.normal(size=(1, n_pts))
    edges = np.linspace(0, 577.0, 21)
    dev = torch.device("cuda", 0)
    pt = torch.from_numpy(pos).to(dev)
    ft = torch.from_numpy(field).to(dev)
    et = torch.from_numpy(edges).to(dev)
    pairs = n_pts * (n_pts - 1) / 2
    run = lambda: gv._pairs(ft, et, pt, "m")
    wall, ker = time_kernel(run, max(2, steps // 2), 1, torch)
    g, c = run()
    # size-independent check: with one bin over [0, inf) every pair is counted once
    _, call = gv._pairs(ft, torch.tensor([0.0, np.inf], dtype=torch.float64, device=dev), pt, "m")
    # parity on a 3000-point subset against the oracle (the reference loop over bins x pairs)
    sub = rs.choice(n_pts, 3000, replace=False)
    gs_, cs_ = gv.unstructured(field[:, sub], edges, pos[:, sub])
    t0 = time.perf_counter()
    g_ref, c_ref = oracle.vario_unstructured_c(field[:, sub], edges, pos[:, sub])
    t_cpu = time.perf_counter() - t0
    dirs = np.array([[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]])
    dt = torch.from_numpy(dirs).to(dev)
    drun = lambda: gv._pairs(ft[:, :200_000], et, pt[:, :200_000], "m", dt, np.pi / 8, 20.0, True)
    dwall, _ = time_kernel(drun, 2, 1, torch)
    lag_np = rs.normal(size=(1024, 100_000))
    lag_f = torch.from_numpy(lag_np).to(dev)
    lrun = lambda: gv._lags(lag_f, None, "m")
    lwall, _ = time_kernel(lrun, 3, 1, torch)
    # spot lags against numpy (sum of squared differences / (2 count), counts closed-form)
    lg, lc = (t.cpu().numpy() for t in lrun())
    lag_err = 0.0
    for k in (1, 2, 7, 31, 32, 33, 500, 1023):
        d = lag_np[k:] - lag_np[:-k]
        ref = np.sum(d * d) / (2.0 * d.size)
        lag_err = max(lag_err, abs(lg[k] - ref) / ref)
    lag_counts_ok = bool(np.array_equal(lc[1:], (1024 - np.arange(1, 1024)) * 100_000))
    sub_pairs = 3000 * 2999 / 2
    return {
        "workload": "unstructured 3D variogram, 1e6 points uniform in [0,1000]^3, one field, 20 bins to 577, "
                    "Matheron, fp64",
        "pairs": pairs, "ms": wall * 1e3, "pairs_per_s": pairs / wall,
        "all_pairs_counted_once": bool(int(call.sum().item()) == int(pairs)),
        "subset_counts_equal_oracle": bool(np.array_equal(cs_, c_ref)),
        "subset_rel_max_vs_oracle": float(np.max(np.abs(gs_ - g_ref) / np.maximum(np.abs(g_ref), 1e-300))),
        "cpu_oracle_pairs_per_s": sub_pairs / t_cpu,
        "cpu_oracle_note": "the reference loop visits every pair once per bin (20 bins); C/OpenMP restatement, "
                           "all host threads",
        "directional_3dirs_200k_points_ms": dwall * 1e3,
        "directional_pairs_per_s": 200_000 * 199_999 / 2 / dwall,
        "axis_lags_1024x100000_ms": lwall * 1e3,
        "axis_lag_pairs_per_s": (1024 * 1023 / 2) * 100_000 / lwall,
        "axis_lags_rel_max_vs_numpy_8_lags": lag_err,
        "axis_lag_counts_exact": lag_counts_ok,
    }


# Modes of the 2D Stable workload: the opt-in series density.  The default
# (the reference's N = 200, h = 0.001 quadrature) truncates the transform
# beyond k L ~ 20; its radial pdf is then not normalisable and the MCMC radii
# of this model reach |k| ~ 1e10 (median ~21 against ~0.1), phases of ~1e13
# rad that no precision resolves.
C4_STABLE_DENSITY = "series"
C4_DENSITY_NOTE = ("modes from Stable(density='series'), not the reference's truncated quadrature (default), "
                   "whose MCMC radii for this model diverge (median |k| ~21, max ~8e10)")


def bench_adjoint(torch, steps, warmup):
    """2D Stable(alpha 1.5, len 15), N 2048, 1e6 points in [0,500]^2, FP32: forward vs forward+adjoint."""
    import oracle
    from gstools_amd import Stable, RandMeth, summate_adjoint_device, summate_device

    log("bench: forward + adjoint (2D Stable, 1e6 points, N 2048)")
    model = Stable(dim=2, var=1.0, len_scale=15.0, alpha=1.5, density=C4_STABLE_DENSITY)
    rm = RandMeth(model, mode_no=2048, seed=20170519)
    dev = torch.device("cuda", 0)
    cov = torch.from_numpy(rm._cov_sample).to(dev)
    z1 = torch.from_numpy(rm._z_1).to(dev)
    z2 = torch.from_numpy(rm._z_2).to(dev)
    npts = 1_000_000
    pos = torch.from_numpy(np.random.RandomState(19970222).uniform(0, 500, size=(2, npts))).to(dev)
    g = torch.from_numpy(np.random.RandomState(3).normal(size=npts).astype(np.float32)).to(dev)
    out = torch.empty(npts, dtype=torch.float32, device=dev)
    work = float(npts) * 2048
    fwd = lambda: summate_device(cov, z1, z2, pos, "fp32", out=out)
    wall_f, ker_f = time_kernel(fwd, steps, warmup, torch)

    def fwd_adj():
        summate_device(cov, z1, z2, pos, "fp32", out=out)
        summate_adjoint_device(cov, z1, z2, pos, g, "fp32")

    wall_fa, _ = time_kernel(fwd_adj, steps, warmup, torch)
    # gradient accuracy on 2000 observation points: FP32 adjoint vs FP64 adjoint vs the oracle
    obs = np.random.RandomState(4).choice(npts, 2000, replace=False)
    pobs = pos[:, obs].contiguous()
    gobs = np.random.RandomState(5).normal(size=2000)
    a32 = [t.cpu().numpy() for t in summate_adjoint_device(cov, z1, z2, pobs, torch.tensor(gobs, device=dev), "fp32")]
    a64 = [t.cpu().numpy() for t in summate_adjoint_device(cov, z1, z2, pobs, torch.tensor(gobs, device=dev), "fp64")]
    ora = oracle.summate_adjoint_c(rm._cov_sample, rm._z_1, rm._z_2, pobs.cpu().numpy(), gobs)
    rel = lambda a, b: float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
    # CPU restatement (forward and adjoint loops) on a prefix of the points
    pos_h, g_h = pos.cpu().numpy(), g.cpu().numpy().astype(np.float64)
    c = (rm._cov_sample, rm._z_1, rm._z_2)
    rate_f, n_f, dt_f, cores = _cpu_restatement(oracle.summate_c, *c, pos_h)
    rate_a, n_a, dt_a, _ = _cpu_restatement(oracle.summate_adjoint_c, *c, pos_h, g_h)
    cpu = {"forward_value": rate_f, "adjoint_value": rate_a, "unit": "mode·points/s", "cores": cores,
           "sample": f"forward: first {n_f} points x 2048 modes ({dt_f:.1f} s); adjoint: first {n_a} points "
                     f"({dt_a:.1f} s); FP64",
           "forward_plus_adjoint_estimate_ms": (work / rate_f + work / rate_a) * 1e3}
    return {
        "workload": "unstructured 2D Stable alpha 1.5 len 15, mode_no 2048 (MCMC modes), 1e6 points in [0,500]^2, fp32",
        "stable_density": C4_DENSITY_NOTE,
        "forward_ms": wall_f * 1e3,
        "forward_plus_adjoint_ms": wall_fa * 1e3,
        "ratio_fwd_adj_over_fwd": wall_fa / wall_f,
        "forward_value": work / wall_f,
        "grad_rel_err_fp32_vs_fp64": [rel(a32[i], a64[i]) for i in range(3)],
        "grad_rel_err_fp64_vs_oracle": [rel(a64[i], ora[i]) for i in range(3)],
        "cpu_restatement": cpu,
        "speedup_fwd_adj_vs_cpu": cpu["forward_plus_adjoint_estimate_ms"] / (wall_fa * 1e3),
    }


def _batch_specs():
    """The 256 mixed realizations: (kind, model args, N, seed, size)."""
    from gstools_amd import MasterRNG

    rs = np.random.RandomState(20170519)
    seeds = MasterRNG(20170519)
    modes = [1000, 2048, 4096, 8192]
    specs = []
    for _ in range(128):
        side = int(round(np.exp(rs.uniform(np.log(64), np.log(2048)))))
        specs.append(("grid2d", int(rs.choice(modes)), seeds(), side))
    for _ in range(96):
        npts = int(round(np.exp(rs.uniform(np.log(1e4), np.log(1e6)))))
        specs.append(("points3d", int(rs.choice(modes)), seeds(), npts))
    for _ in range(32):
        specs.append(("incompr3d", int(rs.choice(modes)), seeds(), 128))
    return specs


def _sample_modes(spec):
    """One realization's modes through ``RandMeth`` (the one-at-a-time sampler)."""
    from gstools_amd import RandMeth

    kind, n, seed, _ = spec
    rm = RandMeth(_model_of(kind), mode_no=n, seed=seed)
    return rm._cov_sample, rm._z_1, rm._z_2


def _model_of(kind):
    from gstools_amd import Exponential, Gaussian

    if kind == "grid2d":
        return Gaussian(dim=2, var=1.0, len_scale=10.0)
    if kind == "points3d":
        return Exponential(dim=3, var=2.0, len_scale=[50, 20, 5], angles=[0.5, 0.2, 0.1])
    return Gaussian(dim=3, var=1.0, len_scale=10.0)


def batch_modes():
    """Sample the modes of all 256 realizations.

    Each kind shares one model, so its realizations are sampled together
    (``sample_modes_many``: every numpy draw made natively for all of them,
    the MCMC chains of both MCMC kinds in one GPU launch, identical arrays to
    one ``RandMeth`` per realization).  Timed twice (the first call also loads the kernels); the
    host lockstep sampler and the one-at-a-time sampler are timed for
    comparison, the latter on a few MCMC realizations scaled by their share
    of the MCMC modes.
    """
    import torch

    from gstools_amd.sampling import sample_modes_many

    specs = _batch_specs()

    kinds = ("grid2d", "points3d", "incompr3d")
    groups_idx = [[i for i, sp in enumerate(specs) if sp[0] == kind] for kind in kinds]

    def sample_all(device):
        groups = [(_model_of(kind), [specs[i][1] for i in idx], [specs[i][2] for i in idx])
                  for kind, idx in zip(kinds, groups_idx)]
        modes = [None] * len(specs)
        for idx, got in zip(groups_idx, sample_modes_many(groups, device=device)):
            for i, g in zip(idx, got):
                modes[i] = g
        return modes

    times = []
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        modes = sample_all(None)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    t0 = time.perf_counter()
    host = sample_all(False)
    t_host = time.perf_counter() - t0
    same = all(all(np.array_equal(a, b) for a, b in zip(m, h)) for m, h in zip(modes, host))
    mcmc = [sp for sp in specs if sp[0] != "grid2d"]
    sub = mcmc[::16]
    t0 = time.perf_counter()
    for sp in sub:
        _sample_modes(sp)
    t_sub = time.perf_counter() - t0
    share = sum(sp[1] for sp in sub) / sum(sp[1] for sp in mcmc)
    info = {"realizations": len(specs), "mcmc_realizations": len(mcmc),
            "sampling_s": times[-1], "sampling_s_first_call": times[0],
            "host_lockstep_s": t_host, "device_equals_host_bitwise": bool(same),
            "one_at_a_time_s_est": t_sub / share,
            "one_at_a_time_measured_on": f"{len(sub)} of {len(mcmc)} MCMC realizations, scaled by mode count",
            "speedup_vs_one_at_a_time": (t_sub / share) / times[-1]}
    log(f"bench: sampled modes of {len(specs)} realizations in {times[-1] * 1e3:.1f} ms (device chains); "
        f"host lockstep {t_host:.2f} s; one at a time ~{t_sub / share:.1f} s")
    return specs, modes, info


def bench_batch(torch, steps, warmup, specs, modes, sampling_info=None):
    """256 realizations: one persistent launch (HIP graph) vs per-realization launches."""
    import oracle
    from gstools_amd import Exponential, summate_device, summate_grid_device, summate_incompr_device
    from gstools_amd.batch import BatchItem, BatchPlan, autotune

    log("bench: batch of 256 realizations")
    dev = torch.device("cuda", 0)
    expo = Exponential(dim=3, var=2.0, len_scale=[50, 20, 5], angles=[0.5, 0.2, 0.1])
    items, singles = [], []
    work = 0.0
    inc_pos = {}  # meshgrid of each incompressible grid size, built once (outside the timed loops)
    for spec in specs:
        if spec[0] == "incompr3d" and spec[3] not in inc_pos:
            inc_pos[spec[3]] = _grid_pos(torch, [torch.arange(spec[3], dtype=torch.float64, device=dev)] * 3)
    for r, (spec, (cov, z1, z2)) in enumerate(zip(specs, modes)):
        kind, n, seed, size = spec
        covt, z1t, z2t = (torch.from_numpy(a).to(dev) for a in (cov, z1, z2))
        if kind == "grid2d":
            axes = [torch.arange(size, dtype=torch.float64, device=dev)] * 2
            it = BatchItem(covt, z1t, z2t, axes=axes)
            singles.append(lambda c=covt, a=z1t, b=z2t, ax=axes, o=it.out: summate_grid_device(
                c, a, b, ax, "fp32", out=o.view(-1)))
        elif kind == "points3d":
            pos = np.random.RandomState(seed).uniform(0.0, 1000.0, size=(3, size))
            post = torch.from_numpy(np.ascontiguousarray(expo.isometrize(pos))).to(dev)
            it = BatchItem(covt, z1t, z2t, pos=post)
            singles.append(lambda c=covt, a=z1t, b=z2t, p=post, o=it.out: summate_device(
                c, a, b, p, "fp32", out=o.view(-1)))
        else:
            axes = [torch.arange(size, dtype=torch.float64, device=dev)] * 3
            it = BatchItem(covt, z1t, z2t, axes=axes, incompr=True)
            singles.append(lambda c=covt, a=z1t, b=z2t, p=inc_pos[size], o=it.out: summate_incompr_device(
                c, a, b, p, "fp32", out=o))
        items.append(it)
        work += float(it.P) * n
    best, tune_ms = autotune(items, reps=1)
    plan = BatchPlan(items, *best)
    graph_wall, _ = time_kernel(lambda: plan.launch(graph=True), steps, warmup, torch)
    eager_wall, _ = time_kernel(plan.launch, steps, warmup, torch)
    batched = [it.out.clone() for it in items]
    ntiles = plan.tiles

    def per_item():
        for f in singles:
            f()
    per_wall, _ = time_kernel(per_item, max(1, steps // 2), 1, torch)
    same = all(torch.equal(it.out, b) for it, b in zip(items, batched))
    # spot checks against the oracle on 2000 points of a few items
    errs = []
    for r in (0, 130, 250):
        cov, z1, z2 = modes[r]
        it = items[r]
        idx = np.arange(min(2000, it.P))
        if it.axes is not None:
            ax = [a.cpu().numpy() for a in torch.split(it.axes, list(it.shape[:it.d]))]
            full = np.asarray(np.meshgrid(*ax, indexing="ij")).reshape(it.d, -1)
            p = full[:, idx]
        else:
            p = it.pos.cpu().numpy()[:, idx]
        ref = oracle.summate_incompr_c(cov, z1, z2, p) if it.incompr else oracle.summate_c(cov, z1, z2, p)[None]
        errs.append(oracle.rel_errors(batched[r][:, idx].cpu().numpy().astype(np.float64), ref)[0])
    plan.close()
    # the same batch with the grid items (2D and incompressible 3D) on the matrix cores
    from gstools_amd.batch import MixedBatch
    mb_default = MixedBatch(items, "bf16x3", *best, tune=False)
    default_wall, _ = time_kernel(lambda: mb_default.launch(graph=True), steps, warmup, torch)
    t0 = time.perf_counter()
    mb = MixedBatch(items, "bf16x3", *best, tune=True)  # tunes each grid class on first use (cached)
    tune_s = time.perf_counter() - t0
    mixed_wall, _ = time_kernel(lambda: mb.launch(graph=True), steps, warmup, torch)
    mixed_errs = []
    for r in (0, 130, 250):
        cov, z1, z2 = modes[r]
        it = items[r]
        idx = np.arange(min(2000, it.P))
        if it.axes is not None:
            ax = [a.cpu().numpy() for a in torch.split(it.axes, list(it.shape[:it.d]))]
            p = np.asarray(np.meshgrid(*ax, indexing="ij")).reshape(it.d, -1)[:, idx]
        else:
            p = it.pos.cpu().numpy()[:, idx]
        ref = oracle.summate_incompr_c(cov, z1, z2, p) if it.incompr else oracle.summate_c(cov, z1, z2, p)[None]
        got = it.out[:, idx].cpu().numpy().astype(np.float64)
        mixed_errs.append(max(oracle.rel_errors(got[c], ref[c])[1] for c in range(got.shape[0])))
    # CPU restatement on a stratified sample: every 16th realization, its first 20000 points
    cores = len(os.sched_getaffinity(0))
    sample_work, sample_s = 0.0, 0.0
    for r in range(0, len(items), 16):
        cov, z1, z2 = modes[r]
        it = items[r]
        idx = np.arange(min(20000, it.P))
        if it.axes is not None:
            ax = [a.cpu().numpy() for a in torch.split(it.axes, list(it.shape[:it.d]))]
            p = np.ascontiguousarray(np.asarray(np.meshgrid(*ax, indexing="ij")).reshape(it.d, -1)[:, idx])
        else:
            p = np.ascontiguousarray(it.pos.cpu().numpy()[:, idx])
        fn = oracle.summate_incompr_c if it.incompr else oracle.summate_c
        t0 = time.perf_counter()
        fn(cov, z1, z2, p, nthreads=cores)
        sample_s += time.perf_counter() - t0
        sample_work += float(idx.size) * cov.shape[1] * (it.d if it.incompr else 1)
    cpu_rate = sample_work / sample_s
    work_weighted = sum(float(it.P) * it.N * (it.d if it.incompr else 1) for it in items)
    cpu_batch = {"value": cpu_rate, "unit": "mode·points/s (incompressible: per component)", "cores": cores,
                 "sample": f"every 16th realization ({len(range(0, len(items), 16))}), first 20000 points of each, "
                           f"FP64, {sample_s:.1f} s",
                 "full_batch_estimate_s": work_weighted / cpu_rate}
    # the same launches on one stream (no graph branches): the same outputs bit for bit
    out4 = [it.out.clone() for it in items]
    mb1 = MixedBatch(items, "bf16x3", *best, tune=True, branches=1)
    one_wall, _ = time_kernel(lambda: mb1.launch(graph=True), steps, warmup, torch)
    branches_equal = all(torch.equal(it.out, o) for it, o in zip(items, out4))
    mixed = {"precision": "bf16x3 grids (128 2D + 32 incompressible 3D) on 4 graph branches beside the fp32 point "
                          "sets' persistent launch, one HIP graph",
             "graph_ms": mixed_wall * 1e3, "value": work / mixed_wall, "speedup_vs_fp32_graph": graph_wall / mixed_wall,
             "default_plans_graph_ms": default_wall * 1e3, "rel_rms_vs_oracle_spot": mixed_errs,
             "mfma_tuning": mb.tuning_summary(), "mfma_tuning_s_first_use": tune_s,
             "branches": mb.branches, "one_stream_graph_ms": one_wall * 1e3,
             "branches_equal_one_stream": bool(branches_equal)}
    return {
        "workload": "256 realizations: 128 2D Gaussian grids (side log-U[64,2048]), 96 3D Exponential point "
                    "sets (log-U[1e4,1e6]), 32 incompressible 3D Gaussian 128^3 grids; mode_no in "
                    "{1000,2048,4096,8192}; seeds from MasterRNG(20170519); fp32",
        "mode_points": work,
        "tiles": ntiles,
        "autotuned_tile_points_x_mode_chunk": list(best),
        "autotune_ms": tune_ms,
        "graph_ms": graph_wall * 1e3,
        "persistent_eager_ms": eager_wall * 1e3,
        "per_realization_ms": per_wall * 1e3,
        "value_graph": work / graph_wall,
        "value_per_realization": work / per_wall,
        "speedup_vs_per_realization": per_wall / graph_wall,
        "batched_equals_single_launches": bool(same),
        "rel_max_vs_oracle_spot": errs,
        "mixed_matrix_cores": mixed,
        "cpu_restatement": cpu_batch,
        "speedup_mixed_vs_cpu": cpu_batch["full_batch_estimate_s"] / mixed_wall,
        "mode_sampling": sampling_info,
        "sampling_and_sums": _end_to_end(torch, specs),
    }


def _end_to_end(torch, specs, reps=3):
    """Sampling and sums of the 256 realizations, end to end: sampling then the
    sums, against :func:`gstools_amd.batch.sample_and_sum` (each group's sums
    launched as soon as its modes are final, beside the device chains).
    Positions, axes and the output buffers (one set per run kind) are on the
    device beforehand, as in a loop that reuses them; both runs build their
    items and batches inside the timing and launch eagerly."""
    from gstools_amd import Exponential
    from gstools_amd.batch import BatchItem, MixedBatch, sample_and_sum
    from gstools_amd.sampling import sample_modes_many

    dev = torch.device("cuda", 0)
    expo = Exponential(dim=3, var=2.0, len_scale=[50, 20, 5], angles=[0.5, 0.2, 0.1])
    kinds = ("grid2d", "points3d", "incompr3d")
    idx = {k: [i for i, sp in enumerate(specs) if sp[0] == k] for k in kinds}
    groups = [(_model_of(k), [specs[i][1] for i in idx[k]], [specs[i][2] for i in idx[k]]) for k in kinds]
    geo = {}
    for i, (kind, n, seed, size) in enumerate(specs):
        if kind == "points3d":
            pos = np.random.RandomState(seed).uniform(0.0, 1000.0, size=(3, size))
            geo[i] = torch.from_numpy(np.ascontiguousarray(expo.isometrize(pos))).to(dev)
        else:
            host = np.arange(size, dtype=np.float64)
            nd = 2 if kind == "grid2d" else 3
            geo[i] = ([torch.from_numpy(host).to(dev)] * nd, [host] * nd)  # device axes and their host copies

    outs = {}
    for run in ("seq", "ovl"):
        for i, (kind, n, seed, size) in enumerate(specs):
            npts = size if kind == "points3d" else size ** (2 if kind == "grid2d" else 3)
            outs[run, i] = torch.empty((3 if kind == "incompr3d" else 1, npts), dtype=torch.float32, device=dev)
    run_kind = ["seq"]

    def items_of(g, modes):
        # one upload per group: each realization's cov, z1, z2 as contiguous views of three flat buffers
        k = kinds[g]
        # staged through pinned memory: a pageable copy would wait for the kernels of the other groups
        flat = []
        for c in range(3):
            host = torch.empty(sum(m[c].size for m in modes), dtype=torch.float64, pin_memory=True)
            np.concatenate([np.ravel(m[c]) for m in modes], out=host.numpy())
            flat.append(host.to(dev, non_blocking=True))
        out, o, oz = [], 0, 0
        for i, (cov, z1, z2) in zip(idx[k], modes):
            d, n = cov.shape
            args = (flat[0][o:o + d * n].view(d, n), flat[1][oz:oz + n], flat[2][oz:oz + n])
            o, oz = o + d * n, oz + n
            res = outs[run_kind[0], i]
            if k == "points3d":
                out.append(BatchItem(*args, pos=geo[i], out=res))
            else:
                out.append(BatchItem(*args, axes=geo[i][0], axes_host=geo[i][1], incompr=k == "incompr3d",
                                     out=res))
        return out

    def sequential():
        run_kind[0] = "seq"
        modes = sample_modes_many(groups)
        its = [it for g in range(len(kinds)) for it in items_of(g, modes[g])]
        mb = MixedBatch(its, "bf16x3", tune=True)
        mb.launch(graph=False)
        torch.cuda.synchronize()
        return its

    def overlapped():
        run_kind[0] = "ovl"
        _, its, keep = sample_and_sum(groups, items_of)
        torch.cuda.synchronize()
        return [it for g in its for it in g]

    seq_s, ovl_s = [], []
    for rep in range(reps + 1):  # the first round warms the sampler and the tuning cache
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        a = sequential()
        seq_s.append(time.perf_counter() - t0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b = overlapped()
        ovl_s.append(time.perf_counter() - t0)
    same = len(a) == len(b) and all(torch.equal(x.out, y.out) for x, y in zip(a, b))
    return {"sequential_ms": 1e3 * float(np.median(seq_s[1:])), "overlapped_ms": 1e3 * float(np.median(ovl_s[1:])),
            "outputs_equal": bool(same),
            "note": "sampling (native draws, MCMC chains on the device) and the MixedBatch sums, eager launches, "
                    "outputs preallocated; overlapped: each group's sums start as soon as its modes are final "
                    "(the device groups first once their chains have run)"}


def _grid_pos(torch, axes):
    g = torch.meshgrid(*axes, indexing="ij")
    return torch.stack([a.reshape(-1) for a in g]).contiguous()


def bench_mfma_ceiling(torch):
    """Matrix-core ceilings on this chip: nothing but independent back-to-back
    MFMAs (``gsx_probe_mfma``, 16 accumulators per wave, two waves per SIMD),
    with near-constant and with random operands.  Random operands toggle the
    multipliers' inputs as real data does; the extra power lowers the clock,
    so that rate, not the published peak, is what a GEMM-shaped kernel on
    real data can approach."""
    import ctypes

    from gstools_amd import _native

    log("bench: matrix-core ceilings")
    dev = torch.device("cuda", 0)
    blocks = 2 * torch.cuda.get_device_properties(0).multi_processor_count
    out = torch.empty(blocks * 256, dtype=torch.float64, device=dev)
    res = {"blocks": blocks, "note": "16 independent accumulators per wave, two waves per SIMD, no memory traffic"}
    for name, kind, flops, iters, peak in (("bf16", _native.GSX_PROBE_BF16, 16384.0, 16000, 2500.0),
                                           ("f64", _native.GSX_PROBE_F64, 2048.0, 4000, 78.6)):
        for rnd in (0, 1):
            def call():
                stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
                _native.check(_native.lib().gsx_probe_mfma(kind, rnd, iters, blocks,
                                                           ctypes.c_void_p(out.data_ptr()), stream), "probe_mfma")

            _, ker = time_kernel(call, 5, 2, torch)
            tf = blocks * 4 * iters * 16 * flops / ker / 1e12
            tag = f"{name}_{'random' if rnd else 'constant'}"
            res[tag + "_tflops"] = tf
            res[tag + "_frac_of_peak"] = tf / peak
            res[tag + "_ms"] = ker * 1e3
    if not bool(torch.isfinite(out).all()):
        raise RuntimeError("probe_mfma: non-finite checksums")
    return res


def bench_copy(torch):
    """A plain 1 GiB device copy: the known-size reference for HBM byte counters."""
    n = 1 << 27  # 2^27 doubles = 1 GiB
    a = torch.empty(n, dtype=torch.float64, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    wall, ker = time_kernel(lambda: b.copy_(a), 5, 2, torch)
    return {"bytes_read": n * 8, "bytes_written": n * 8, "ms": ker * 1e3, "GBps": 2 * n * 8 / ker / 1e9}


def load_traffic():
    """HBM bytes per launch of the headline kernel, measured with rocprofv3 --pmc."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    try:
        with open(path) as fh:
            return json.load(fh).get("k_forward_f32_exponential3d")
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--extras", type=int, default=1, help="also run the other configurations")
    ap.add_argument("--only", choices=["fp32", "fp64", "config1", "mfma", "incompr", "fourier", "krige", "variogram",
                                       "adjoint", "batch", "copy", "ceiling"],
                    help="run one section only (for profiling) and print its JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the summation shards over points with no communication, see the design notes",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from gstools_amd import summate_device
    import oracle

    if not torch.cuda.is_available():
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    want_batch = args.only == "batch" or (args.only is None and args.extras)
    batch_inputs = batch_modes() if want_batch else None
    dev = torch.device("cuda", 0)
    k_ext = max(3, args.steps // 4)
    w_ext = max(1, min(args.warmup, 3))  # the extra sections run long kernels
    if args.only in ("config1", "mfma", "incompr", "fourier", "krige", "variogram", "adjoint", "batch", "copy",
                     "ceiling"):
        fn = {"config1": lambda: bench_config1(torch, k_ext, w_ext),
              "mfma": lambda: bench_mfma_grid(torch, k_ext, w_ext),
              "incompr": lambda: bench_incompr_grid(torch, k_ext, w_ext),
              "variogram": lambda: bench_variogram(torch, k_ext, w_ext),
              "krige": lambda: bench_krige(torch, k_ext, w_ext),
              "fourier": lambda: bench_fourier(torch, k_ext, w_ext),
              "adjoint": lambda: bench_adjoint(torch, k_ext, w_ext),
              "batch": lambda: bench_batch(torch, k_ext, w_ext, *batch_inputs),
              "copy": lambda: bench_copy(torch),
              "ceiling": lambda: bench_mfma_ceiling(torch)}[args.only]
        print(json.dumps({args.only: fn()}), flush=True)
        return

    log("bench: sampling modes and points")
    rm, iso = headline_inputs()
    cov = torch.from_numpy(rm._cov_sample).to(dev)
    z1 = torch.from_numpy(rm._z_1).to(dev)
    z2 = torch.from_numpy(rm._z_2).to(dev)
    pos = torch.from_numpy(np.ascontiguousarray(iso)).to(dev)
    out32 = torch.empty(P_HEADLINE, dtype=torch.float32, device=dev)
    out64 = torch.empty(P_HEADLINE, dtype=torch.float64, device=dev)
    work = float(P_HEADLINE) * N_HEADLINE
    f64 = lambda: summate_device(cov, z1, z2, pos, "fp64", out=out64)
    f32 = lambda: summate_device(cov, z1, z2, pos, "fp32", out=out32)
    if args.only == "fp64":
        wall64, ker64 = time_kernel(f64, args.steps, args.warmup, torch)
        print(json.dumps({"fp64": {"value": work / wall64, "kernel_ms": ker64 * 1e3}}), flush=True)
        return
    if args.only == "fp32":
        wall32, ker32 = time_kernel(f32, args.steps, args.warmup, torch)
        print(json.dumps({"fp32": {"value": work / wall32, "kernel_ms": ker32 * 1e3}}), flush=True)
        return

    log("bench: fp64 plain")
    wall64, ker64 = time_kernel(f64, max(3, args.steps // 4), args.warmup, torch)
    log("bench: fp32")
    wall32, ker32 = time_kernel(f32, args.steps, args.warmup, torch)

    ref64 = out64.cpu().numpy()
    got32 = out32.cpu().numpy().astype(np.float64)
    rel_max, rel_rms = oracle.rel_errors(got32, ref64)
    sub = np.random.RandomState(5).choice(P_HEADLINE, 4096, replace=False)
    ora = oracle.summate_c(rm._cov_sample, rm._z_1, rm._z_2, np.ascontiguousarray(iso[:, sub]))
    parity64 = oracle.rel_errors(ref64[sub], ora)[0]

    flops = 2 * 3 + 6  # per mode.point, d = 3
    achieved = flops * work / ker32 / 1e12
    line = {
        "metric": METRIC,
        "value": work / wall32,
        "unit": "mode·points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "unstructured 3D Exponential, var 2, len_scale [50,20,5], angles [0.5,0.2,0.1], "
                        "mode_no 1024 (MCMC modes, seed 20170519), 4e6 points uniform in [0,1000]^3",
            "points": P_HEADLINE,
            "modes": N_HEADLINE,
            "path": "fp32 (fp64 phase reduced by magic addition, fract form for modes beyond the bound, v_cos_f32, blocked fp32->fp64 sum)",
        },
        "roofline": {
            "bound": "valu",
            "kernel": "k_forward_f32<3,8,false,UnstructuredPoints<3>>",
            "achieved": achieved,
            "peak": FP32_VALU_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_VALU_PEAK_TFLOPS,
            "traffic": load_traffic(),
            "flops_per_mode_point": flops,
            "kernel_ms": ker32 * 1e3,
        },
        "fp64_plain": {
            "value": work / wall64,
            "ms_per_step": wall64 * 1e3,
            "kernel_ms": ker64 * 1e3,
            "roofline_frac": flops * work / ker64 / 1e12 / FP64_VALU_PEAK_TFLOPS,
            "parity_vs_oracle_rel_max": parity64,
        },
        "fp32_vs_fp64": {"rel_max": rel_max, "rel_rms": rel_rms, "speedup": wall64 / wall32},
    }
    if args.extras:
        extras = {}
        sections = (("config1", bench_config1),
                    ("mfma_grid", bench_mfma_grid), ("incompr_grid", bench_incompr_grid), ("fourier", bench_fourier),
                    ("krige", bench_krige),
                    ("variogram", bench_variogram),
                    ("adjoint", bench_adjoint),
                    ("batch", lambda t, k, w: bench_batch(t, k, w, *batch_inputs)),
                    ("mfma_ceiling", lambda t, k, w: bench_mfma_ceiling(t)))
        for name, fn in sections:
            try:
                extras[name] = fn(torch, k_ext, w_ext)
            except Exception as exc:  # report, keep the headline line
                extras[name] = {"error": repr(exc)}
            log(f"bench: {name}: {json.dumps(extras[name])}")
        line["extras"] = extras
    if not args.no_cpu:
        log("bench: cpu baseline")
        line["cpu_baseline"] = cpu_baseline(rm, iso)
        line["vs_cpu"] = line["value"] / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
