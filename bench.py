"""Throughput benchmark of the memetic phase-retrieval hot path on one MI355X.

Workload (the plain-method drop-in case): one 512 x 512 float32 pattern of 1-5
overlapping projected ellipsoids, 1e9 photons with Poisson noise, 12 px beamstop,
two 4 px detector-gap strips and 0.5 % random saturated pixels masked;
population 64. One step = one RAAR block of 25 iterations (beta = 0.87, the
block length of that schedule) over the whole population, with the population
resident in HBM. value = individual-iterations per second (iterations x
population / time).

Prints one JSON line with the metric, a roofline entry for the dominant kernel
(HIP-event timed on the backend's stream during the timed region) and the CPU
oracle timed on this host as the CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "individual-iters/s at 512² pop64; time-to-Rf<0.05, batched+grad vs plain GPU"


def algorithmic_bytes(n, P, alg):
    """HBM bytes per launch of the two iteration kernels (algorithmic, no re-reads):
    column pass 16 B/px (+ the amplitude, 4 B/px once per launch); row update pass
    33 B/px for HIO/RAAR, 17 B/px for ER."""
    npx = n * n
    col = 16 * npx * P + 4 * npx
    row = (33 if alg in ("HIO", "RAAR") else 17) * npx * P
    return col, row


def cpu_baseline(amp, rho, supp, alg, beta, iters, log):
    """The FP32 CPU oracle (own FFT, one individual per OpenMP thread) on a
    bounded sample of the same workload."""
    from oracle import oracle as orc
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    P = rho.shape[0]
    sample = rho
    orc.iterate(alg, sample[:1], supp[:1], amp, 1, beta, 0.0, nthreads=1)  # warm the plan cache
    t0 = time.perf_counter()
    orc.iterate(alg, sample, supp[:P], amp, iters, beta, 0.0, nthreads=cores)
    dt = time.perf_counter() - t0
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    log(f"cpu baseline: {P} individuals x {iters} {alg} iterations on {cores} threads ({cpu}): {dt:.2f} s")
    return {"value": P * iters / dt, "unit": "individual-iters/s", "cores": cores, "kind": "port",
            "sample": f"FP32 oracle, {P} individuals x {iters} {alg} iterations at {amp.shape[0]}^2 ({cpu})"}


def copy_calibration(mib, log):
    """Three device-to-device copies of `mib` MiB through the HIP runtime's plain
    copy kernel, so that a counter run (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)
    records what the counters read for a known number of bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    for _ in range(3):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3):   # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: 3 device-to-device copies of {mib} MiB")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--pop", type=int, default=64)
    ap.add_argument("--block", type=int, default=25, help="iterations per step")
    ap.add_argument("--alg", default="RAAR", choices=["RAAR", "HIO", "ER"])
    ap.add_argument("--beta", type=float, default=0.87)
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate", type=int, default=0, metavar="MIB",
                    help="also run device-to-device copies of MIB MiB (counter calibration)")
    args = ap.parse_args()

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported (the population "
              "shards by individuals across GPUs; only the one-GPU path is built).", file=sys.stderr)
        sys.exit(2)

    from s_spi_spring_amd import hipcore, synth

    n, P = args.n, args.pop
    log(f"generating pattern n={n}")
    I, truth = synth.ellipsoids_pattern(seed=0, n=n) if n >= 256 else synth.polygon_pattern(seed=0, n=n)
    amp = synth.corner_amplitude(I)
    supp = synth.disk_support(n, max(8, int(0.3 * n)))
    scale = float(np.sqrt(np.maximum(I, 0).sum()) / supp.sum())
    rho = synth.random_start(n, P, supp, seed=1, scale=scale)
    supps = np.broadcast_to(supp, rho.shape).astype(np.int32)

    be = hipcore.Backend(amp, capacity=P, device=0)
    be.set_population(rho, supps)
    if args.calibrate:
        copy_calibration(args.calibrate, log)
    log("warmup")
    for _ in range(args.warmup):
        be.iterate_async(args.alg, args.block, args.beta, 0.0)
    be.sync()

    log(f"timed: {args.steps} steps of {args.block} {args.alg} iterations, P={P}")
    be.timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        be.iterate_async(args.alg, args.block, args.beta, 0.0)
    be.sync()
    dt = time.perf_counter() - t0
    col_ms, col_cnt = be.timing_read(hipcore.KIND_COL_PROJ)
    row_ms, row_cnt = be.timing_read(hipcore.KIND_ROW_UPDATE)
    be.timing(False)
    ind_iters = args.steps * args.block * P
    value = ind_iters / dt
    col_b, row_b = algorithmic_bytes(n, P, args.alg)
    col_avg = col_ms / max(col_cnt, 1) * 1e-3
    row_avg = row_ms / max(row_cnt, 1) * 1e-3
    col_gbs = col_b / col_avg / 1e9
    row_gbs = row_b / row_avg / 1e9
    # dominant kernel = the one with the larger share of the timed region
    if row_ms >= col_ms:
        dom = {"kernel": f"k_row<{n},{args.alg}>", "achieved": row_gbs, "bytes_per_launch": row_b,
               "avg_launch_us": row_avg * 1e6}
    else:
        dom = {"kernel": f"k_col<{n},PROJ>", "achieved": col_gbs, "bytes_per_launch": col_b,
               "avg_launch_us": col_avg * 1e6}
    log(f"column pass: {col_cnt} launches, avg {col_avg * 1e6:.1f} us, {col_gbs:.0f} GB/s algorithmic")
    log(f"row pass:    {row_cnt} launches, avg {row_avg * 1e6:.1f} us, {row_gbs:.0f} GB/s algorithmic")
    iter_bytes = (col_b + row_b) / P
    log(f"whole iteration: {value * iter_bytes / 1e9:.0f} GB/s algorithmic "
        f"({value * iter_bytes / 1e9 / HBM_PEAK_GBS:.3f} of HBM peak)")

    roofline = {"bound": "hbm", "achieved": round(dom["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(dom["achieved"] / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": dom["kernel"], "algorithmic_bytes_per_launch": dom["bytes_per_launch"],
                "avg_launch_us": round(dom["avg_launch_us"], 2),
                "iteration_frac": round(value * iter_bytes / 1e9 / HBM_PEAK_GBS, 4)}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(amp, rho, supps, args.alg, args.beta, args.cpu_iters, log)

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "individual-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"ellipsoids {n}^2 pattern, population {P}, {args.alg}(beta={args.beta}) "
                               f"blocks of {args.block} iterations, population resident in HBM",
                   "n": n, "population": P, "alg": args.alg, "iterations_per_step": args.block},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    be.close()


if __name__ == "__main__":
    main()
