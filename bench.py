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
import csv
import hashlib
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
ROW_MODE = {"HIO": 2, "RAAR": 3, "ER": 4}       # RowMode of spr_kernels.hpp
TIMING_STRIDE = 8
DEVICE_SOURCES = ("spr_fft.hpp", "spr_kernels.hpp", "spr_fused.hpp", "spr_aux.hpp", "spr_grad.hpp", "spr_tuned.hpp",
                  "spr_backend.hip")
FP32_PEAK_TFLOPS = 157.3
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_iteration_512_pop64_raar.json")
COUNTER_CMD = ("cd /tmp && export TMPDIR=/tmp && timeout -s KILL 90 rocprofv3 --pmc {ctr} --kernel-trace "
               "--output-format csv -d <dir> -o run -- python3 -u /root/repo/bench.py --steps 2 --warmup 1 "
               "--no-cpu --prewarm 0 --ttr-seeds 0 --calibrate {mib}{extra}")


def algorithmic_bytes(n, P, alg, recomputed_state=False):
    """HBM bytes per launch of the two iteration kernels (algorithmic, no re-reads):
    column pass 16 B/px (+ the amplitude, 4 B/px once per launch); row update pass
    for HIO/RAAR 32.125 B/px with the state stored (the field read and written,
    rho_n read and written, the support as one bit per pixel) or 24.125 B/px with
    it recomputed from its row transform (the projected field and the state's row
    transform read, the new row transform written, the support bits); ER 16.125
    B/px in both forms."""
    npx = n * n
    col = 16 * npx * P + 4 * npx
    if alg in ("HIO", "RAAR"):
        row = (24.125 if recomputed_state else 32.125) * npx * P
    else:
        row = 16.125 * npx * P
    return col, row


def cpu_baseline(amp, rho, supp, alg, beta, iters, log):
    """The FP32 CPU oracle's throughput port (oracle/fast_port.hpp: the same
    driver and per-pixel arithmetic with a vectorised radix-8/4/2/3 transform;
    one individual per OpenMP thread, as the reference's Solver) on a bounded
    sample of the same workload."""
    from oracle import oracle as orc
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    P = rho.shape[0]
    sample = rho
    orc.iterate_fast(alg, sample[:1], supp[:1], amp, 1, beta, 0.0, nthreads=1)  # page in the library
    t0 = time.perf_counter()
    orc.iterate_fast(alg, sample, supp[:P], amp, iters, beta, 0.0, nthreads=cores)
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
    # For information, the plain FP32 oracle and the FP64 oracle on small samples.
    info = {}
    for name, arr, k, its in (("plain_fp32_oracle", sample, P, 5), ("fp64_oracle", sample.astype(np.complex128), 16, 5)):
        t0 = time.perf_counter()
        orc.iterate(alg, arr[:k], supp[:k], amp.astype(arr.real.dtype), its, beta, 0.0, nthreads=cores)
        t = time.perf_counter() - t0
        info[name] = {"value": k * its / t, "sample": f"{k} individuals x {its} {alg} iterations"}
        log(f"{name}: {k * its / t:.1f} individual-iterations/s")
    return {"value": P * iters / dt, "unit": "individual-iters/s", "cores": cores, "kind": "port",
            "sample": f"FP32 oracle (throughput port, AVX2), {P} individuals x {iters} {alg} iterations at "
                      f"{amp.shape[0]}^2 ({cpu})", **info}


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


def device_source_hash():
    """Hash of the device sources: a traffic summary applies to the kernels it was
    measured on only."""
    h = hashlib.sha256()
    for f in DEVICE_SOURCES:
        with open(os.path.join(ROOT, "s_spi_spring_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def fused_flops(n, P, block):
    """Nominal FFT flops per launch of the on-chip block kernel: 4(block+1) sets
    of N length-N transforms (5 N log2 N flops each) per individual; the
    projection and update arithmetic are not counted."""
    return 4 * (block + 1) * 5 * n * n * np.log2(n) * P


def rocprof_kernel_name(kernel, n, alg, variant):
    if kernel == "fused":
        return f"void spr::k_iter_fused<{n}, {ROW_MODE[alg]}>(spr::FusedArgs)"
    if kernel == "row":
        return f"void spr::k_row<{n}, {ROW_MODE[alg]}, {variant}>(spr::RowArgs)"
    return f"void spr::k_col<{n}, 0, {variant}>(spr::ColArgs)"


def _counter_values(run_dir, counter):
    vals = {}
    with open(os.path.join(run_dir, "run_counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                vals.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return vals


def summarize_traffic(fetch_dir, write_dir, mib, out_path, log, extra=""):
    """HBM bytes per dispatch of every spr:: kernel from a FETCH_SIZE and a
    WRITE_SIZE counter run of `bench.py --calibrate MIB`. Each counter is scaled
    so that it reads MIB MiB for the three MIB MiB device-to-device copies (the
    largest copy dispatches of the run)."""
    fetch = _counter_values(fetch_dir, "FETCH_SIZE")
    write = _counter_values(write_dir, "WRITE_SIZE")

    def factor(vals):
        copies = sorted(v for k, lst in vals.items() if k.startswith("__amd_rocclr_copyBuffer") for v in lst)[-3:]
        if len(copies) < 3:
            raise SystemExit("the counter runs hold no calibration copies (run them with --calibrate MIB)")
        return mib * 1024.0 / (sum(copies) / len(copies))

    ff, fw = factor(fetch), factor(write)
    kernels = {}
    for k in sorted(fetch):
        if not k.startswith(("void spr::", "spr::")) or k not in write:
            continue
        f_kib = sum(fetch[k]) / len(fetch[k])
        w_kib = sum(write[k]) / len(write[k])
        kernels[k] = {"dispatches": len(fetch[k]), "fetch_kib": round(f_kib, 1), "write_kib": round(w_kib, 1),
                      "hbm_bytes": int(round((f_kib * ff + w_kib * fw) * 1024))}
    out = {"what": "HBM bytes per dispatch (FETCH_SIZE x fetch_factor + WRITE_SIZE x write_factor, KiB x 1024)",
           "commands": [COUNTER_CMD.format(ctr="FETCH_SIZE", mib=mib, extra=extra),
                        COUNTER_CMD.format(ctr="WRITE_SIZE", mib=mib, extra=extra)],
           "device_source_hash": device_source_hash(),
           "calibration": {"copy_mib": mib, "fetch_factor": round(ff, 4), "write_factor": round(fw, 4)},
           "kernels": kernels}
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    for k, v in kernels.items():
        log(f"{k}: {v['hbm_bytes'] / 1e6:.1f} MB per dispatch over {v['dispatches']} dispatches")
    log(f"wrote {out_path}")


def measured_traffic(kernel_name, log):
    """HBM bytes per launch of `kernel_name` from the committed counter summary,
    or None when there is none or it was measured on other device sources."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except OSError:
        log("no HBM traffic summary: roofline.traffic is null")
        return None
    if t.get("device_source_hash") != device_source_hash():
        log("HBM traffic summary was measured on other device sources: roofline.traffic is null")
        return None
    k = t["kernels"].get(kernel_name)
    return None if k is None else k["hbm_bytes"]


def time_to_rf_runs(seeds, generations, target, log, threshold=None, groups=(), group_seeds=None):
    """Runs of the time-to-R_F study: the plain and the gradient schedules
    (whole population per launch) over `seeds` seeded 512^2 ellipsoid patterns
    (population 64), and for each k in `groups` the plain schedule run k
    individuals per launch (the reference's structure: each of its k GPU
    contexts runs one individual at a time) over the first `group_seeds` seeds."""
    from s_spi_spring_amd import workloads
    log("warm-up run (not recorded)")
    workloads.time_to_rf(10_000, "plain", target=target, generations=0)
    legs = [("plain", "plain", 0), ("grad", "grad", 0)] + [(f"plain_groups_of_{k}", "plain", k) for k in groups]
    runs = {name: [] for name, _, _ in legs}
    for seed in range(seeds):
        for name, form, k in legs:
            if k and group_seeds is not None and seed >= group_seeds:
                continue
            r = workloads.time_to_rf(seed, form, target=target, generations=generations, threshold=threshold,
                                     max_batch=k)
            runs[name].append(r)
            log(f"seed {seed:2d} {name:5s}: t={r['time_to_target']} s gen={r['generation']} "
                f"E_F={r['fourier_error']:.4f} (floor {r['noise_floor']:.4f}) real-space={r['real_space_error']:.4f} "
                f"wall={r['wall']:.1f} s history={r['history']}")
    return legs, runs


def ttr_summary(rs):
    t = [r["time_to_target"] for r in rs if r["time_to_target"] is not None]
    rse = [r["real_space_error"] for r in rs]
    return {"reached": len(t), "runs": len(rs),
            "median_time_s": float(np.median(t)) if t else None,
            "mean_time_s": float(np.mean(t)) if t else None,
            "median_generations": float(np.median([r["generation"] for r in rs])),
            "median_fourier_error": float(np.median([r["fourier_error"] for r in rs])),
            "median_real_space_error": float(np.median(rse)),
            "real_space_error_q25_q75": [float(np.percentile(rse, q)) for q in (25, 75)]}


def _ratio(a, b):
    return a / b if a and b else None


def ttr_paired(plain, grad):
    """Per-seed comparison of two legs run on the same seeds (the seeds both
    reach the target): the median of the per-seed time ratios plain / grad, the
    seeds where each leg is faster, and the two-sided Wilcoxon signed-rank test
    of the per-seed time differences (p near 1: no systematic difference)."""
    try:
        from scipy.stats import wilcoxon
    except ImportError:   # the test statistic is then left out
        wilcoxon = None
    tp = {r["seed"]: r["time_to_target"] for r in plain}
    tg = {r["seed"]: r["time_to_target"] for r in grad}
    both = sorted(s for s in tp if tp[s] is not None and tg.get(s) is not None)
    a = np.array([tp[s] for s in both])
    b = np.array([tg[s] for s in both])
    d = a - b
    p = float(wilcoxon(a, b).pvalue) if wilcoxon is not None and len(both) and np.any(d != 0) else None
    return {"seeds_both_reached": len(both),
            "median_ratio_plain_over_grad": float(np.median(a / b)) if len(both) else None,
            "grad_faster": int(np.sum(d > 0)), "plain_faster": int(np.sum(d < 0)), "wilcoxon_p": p}


def time_to_rf_fields(seeds, generations, target, log, group_seeds=8):
    """The time-to-R_F half of the headline metric for the default bench line:
    batched+gradient against batched plain over `seeds` seeds (the gradient
    operator's gain), and the plain schedule with 8 individuals per launch (the
    reference's default of 8 GPU contexts) over `group_seeds` seeds."""
    from s_spi_spring_amd import workloads
    legs, runs = time_to_rf_runs(seeds, generations, target, log, groups=(8,), group_seeds=group_seeds)
    sp, sg, s8 = ttr_summary(runs["plain"]), ttr_summary(runs["grad"]), ttr_summary(runs["plain_groups_of_8"])
    return {"target_fourier_error": target, "seeds": seeds, "generations_max": generations,
            "workload": "ellipsoids 512^2, 1e9 photons, population 64, RAAR400+HIO100+ER50 per generation in "
                        "25-iteration blocks with shrink-wrap; gradient form: 10 support-projected Adam steps "
                        "on the Poisson likelihood after each RAAR block",
            "batched_grad": sg, "batched_plain": sp, "plain_8_per_launch": dict(s8, seeds=group_seeds),
            "speedup_grad_vs_batched_plain": _ratio(sp["median_time_s"], sg["median_time_s"]),
            "paired_grad_vs_batched_plain": ttr_paired(runs["plain"], runs["grad"]),
            "speedup_grad_vs_8_per_launch": _ratio(s8["median_time_s"], sg["median_time_s"]),
            "median_noise_floor": float(np.median([r["noise_floor"] for r in runs["plain"]])),
            "sw_threshold": workloads.SW_THRESHOLD_512}


def time_to_rf_study(seeds, generations, target, log, threshold=None, groups=()):
    """Time to best E_F < target of the plain and the gradient schedules over
    `seeds` seeded 512^2 ellipsoid patterns (population 64); with `groups`, also
    the plain schedule run k individuals per launch. One JSON line."""
    from s_spi_spring_amd import workloads
    legs, runs = time_to_rf_runs(seeds, generations, target, log, threshold, groups)
    summary = ttr_summary

    sp, sg = summary(runs["plain"]), summary(runs["grad"])
    speedup = (sp["median_time_s"] / sg["median_time_s"]) if sp["median_time_s"] and sg["median_time_s"] else None
    grouped = {}
    for name, _, k in legs[2:]:
        sk = summary(runs[name])
        sk["individuals_per_launch"] = k
        sk["speedup_grad_vs_this"] = (sk["median_time_s"] / sg["median_time_s"]
                                      if sk["median_time_s"] and sg["median_time_s"] else None)
        sk["speedup_plain_vs_this"] = (sk["median_time_s"] / sp["median_time_s"]
                                       if sk["median_time_s"] and sp["median_time_s"] else None)
        grouped[name] = sk
    out = {"metric": f"time-to-Rf<{target}", "unit": "s", "higher_is_better": False, "seeds": seeds,
           "generations_max": generations, "config": {"workload": "ellipsoids 512^2 pattern, 1e9 photons, "
                                                      "population 64, RAAR400+HIO100+ER50 per generation in "
                                                      "25-iteration blocks with shrink-wrap"},
           "plain": sp, "grad": sg, "speedup_grad_vs_plain": speedup,
           "paired_grad_vs_plain": ttr_paired(runs["plain"], runs["grad"]), **grouped,
           "sw_threshold": threshold if threshold is not None else workloads.SW_THRESHOLD_512,
           "median_noise_floor": float(np.median([r["noise_floor"] for r in runs["plain"]])),
           "per_seed": {f: [{k: r[k] for k in ("seed", "time_to_target", "generation", "fourier_error",
                                               "real_space_error", "noise_floor", "history")} for r in rs]
                        for f, rs in runs.items()}}
    print(json.dumps(out), flush=True)


def differentiation_study(seeds, generations, log):
    """The low-count differentiation workload (1024^2 truncated octahedron,
    1e8 photons, population 32, RAAR blocks alternating with 20 Adam steps on
    the Poisson likelihood): best Fourier error and time per generation; one
    JSON line."""
    from s_spi_spring_amd import workloads
    runs = []
    for seed in range(seeds):
        r = workloads.differentiation_run(seed, generations=generations)
        runs.append(r)
        log(f"seed {seed}: E_F by generation {r['history']} at {r['times']} s; real-space {r['real_space_error']:.3f} "
            f"(noise floor {r['noise_floor']:.4f})")
    per_gen = [np.diff(r["times"]).tolist() for r in runs]
    out = {"metric": "s per memetic generation, low-count differentiation workload", "unit": "s",
           "higher_is_better": False, "seeds": seeds, "generations": generations,
           "median_s_per_generation": float(np.median([t for ts in per_gen for t in ts])),
           "config": {"workload": "truncated-octahedron 1024^2 pattern, 1e8 photons, population 32, per generation "
                                  "8 x [RAAR 25, 20 support-projected Adam steps on the Poisson likelihood (step 1e-2 "
                                  "of the RMS density), shrink-wrap] + ER 50"},
           "iterations_per_generation": runs[0]["iterations_per_generation"],
           "adam_steps_per_generation": runs[0]["adam_steps_per_generation"],
           "per_seed": runs}
    print(json.dumps(out), flush=True)


def refine_bench(n, P, steps, log):
    """Throughput of the Adam refinement steps (spr_refine, Poisson loss, no
    support mask) on a synthetic population: the time of `steps` more steps
    than a short run, per step, so that the set-up and the final loss cancel.
    One step is the column loss pass (field read and written, 16 B/px, and the
    amplitude once per launch) and the row Adam pass (field, density and both
    moments read and written, 64 B/px): 80 B/px algorithmic. One JSON line."""
    from s_spi_spring_amd import hipcore, synth
    I, _ = synth.octahedron_pattern(seed=0, n=n, photons=1e8) if n >= 512 else synth.polygon_pattern(seed=0, n=n)
    amp = synth.corner_amplitude(I)
    supp = synth.disk_support(n, max(8, int(0.3 * n)))
    scale = float(np.sqrt(np.maximum(I, 0).sum()) / supp.sum())
    rho = synth.random_start(n, P, supp, seed=1, scale=scale)
    supps = np.broadcast_to(supp, rho.shape).astype(np.int32)
    be = hipcore.Backend(amp, capacity=P, device=0)
    short = 5

    def timed(k):
        be.set_population(rho, supps)
        be.sync()
        t0 = time.perf_counter()
        be.refine(k, lr=1e-2)
        return time.perf_counter() - t0

    timed(short)   # warm-up (code load, clocks)
    reps = [(timed(short), timed(short + steps)) for _ in range(3)]
    be.close()
    per_step = float(np.median([(b - a) / steps for a, b in reps]))
    step_bytes = (16 + 64) * n * n * P + 4 * n * n
    gbs = step_bytes / per_step / 1e9
    log(f"refinement: {steps} steps over {P} individuals at {n}^2: {per_step * 1e3:.3f} ms per step, "
        f"{gbs:.0f} GB/s algorithmic ({gbs / HBM_PEAK_GBS:.3f} of HBM peak)")
    out = {"metric": "Adam refinement steps/s (masked-Poisson loss gradient)", "value": round(1.0 / per_step, 1),
           "unit": "steps/s", "higher_is_better": True, "n_gpus": 1, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"octahedron {n}^2 pattern, 1e8 photons, population {P}, spr_refine (Poisson)",
                      "n": n, "population": P, "steps": steps},
           "ms_per_step": round(per_step * 1e3, 4),
           "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes_per_step": step_bytes}}
    print(json.dumps(out), flush=True)


def batch_bench(args, log):
    """Mixed-size batch: `args.batch` independent polygon patterns with sizes
    drawn from {256, 384, 512, 768, 1024} and `args.batch_pop` individuals each,
    all resident in HBM; one step = one RAAR block of `args.batch_block`
    iterations over every individual of every pattern (one HIP graph launch per
    block, or direct launches with --no-graphs). One JSON line."""
    from s_spi_spring_amd import hipcore, synth
    P = args.batch_pop
    log(f"generating {args.batch} patterns")
    amps, starts = [], {}
    t0 = time.perf_counter()
    for I, _ in synth.batch_patterns(args.batch, seed=0):
        amps.append(synth.corner_amplitude(I))
    sizes = [a.shape[0] for a in amps]
    log(f"generated in {time.perf_counter() - t0:.1f} s; sizes: " +
        ", ".join(f"{n}^2 x {sizes.count(n)}" for n in sorted(set(sizes))))
    b = hipcore.Batch(amps, [P] * len(amps), device=0)
    b.graphs(not args.no_graphs)
    b.concurrent(args.batch_concurrent)
    for k, a in enumerate(amps):
        n = a.shape[0]
        if n not in starts:   # one random start per size, scaled per pattern
            supp = synth.disk_support(n, max(8, int(0.3 * n)))
            starts[n] = (synth.random_start(n, P, supp, seed=1), supp)
        rho, supp = starts[n]
        scale = float(np.sqrt((np.maximum(a, 0).astype(np.float64) ** 2).sum()) / supp.sum())
        b.set_population(k, rho * np.complex64(scale), supp)
    block = [("RAAR", [args.batch_block, args.beta])]
    log("warmup")
    for _ in range(args.warmup):
        b.run(block)
    log(f"timed: {args.steps} blocks of {args.batch_block} RAAR iterations over {b.total} individuals")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.run(block)
    dt = time.perf_counter() - t0
    ind_iters = args.steps * args.batch_block * b.total
    value = ind_iters / dt
    # algorithmic bytes of one iteration of the whole batch: per individual 16
    # B/px of column pass and 32.125 (state stored) or 24.125 (recomputed) of
    # row pass, by the size class's row variant; each pattern's amplitude once
    # per column pass
    per_px = {n: 16 + (24.125 if hipcore.row_variant_recomputes_state(n, hipcore.default_variants(n)[1]) else 32.125)
              for n in set(sizes)}
    it_bytes = sum(per_px[n] * n * n * P + 4.0 * n * n for n in sizes)
    frac = it_bytes * (value / b.total) / 1e9 / HBM_PEAK_GBS
    log(f"{value:.0f} individual-iterations/s; whole iteration {it_bytes * value / b.total / 1e9:.0f} GB/s "
        f"algorithmic ({frac:.3f} of HBM peak)")
    ev = b.evaluate(0.0)
    ef = np.concatenate(ev["error_fourier"])
    out = {"metric": "individual-iters/s, mixed-size pattern batch", "value": round(value, 1),
           "unit": "individual-iters/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"{len(amps)} polygon patterns of sizes drawn from 256^2..1024^2, population {P} "
                                  f"each, RAAR(beta={args.beta}) blocks of {args.batch_block} iterations, "
                                  f"{'one HIP graph launch' if not args.no_graphs else 'direct launches'} per block",
                      "patterns": len(amps), "sizes": {str(n): sizes.count(n) for n in sorted(set(sizes))},
                      "population": P, "individuals": b.total, "iterations_per_step": args.batch_block,
                      "graphs": not args.no_graphs, "size_classes": "concurrent" if args.batch_concurrent else "in turn"},
           "iteration_frac": round(frac, 4),
           "fourier_error_median_after": float(np.median(ef))}
    print(json.dumps(out), flush=True)
    b.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--pop", type=int, default=64)
    ap.add_argument("--block", type=int, default=25, help="iterations per step")
    ap.add_argument("--alg", default="RAAR", choices=["RAAR", "HIO", "ER"])
    ap.add_argument("--beta", type=float, default=0.87)
    ap.add_argument("--cpu-iters", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true",
                    help="time the steps by wall clock only (no per-launch events, no roofline entry)")
    ap.add_argument("--calibrate", type=int, default=0, metavar="MIB",
                    help="also run device-to-device copies of MIB MiB (counter calibration)")
    ap.add_argument("--time-to-rf", type=int, default=0, metavar="SEEDS",
                    help="instead of the throughput step: time to best E_F < --target of the plain and "
                         "gradient schedules over SEEDS patterns (one JSON line)")
    ap.add_argument("--target", type=float, default=0.05)
    ap.add_argument("--plain-groups", default="", metavar="K,K",
                    help="time-to-rf: also run the plain schedule K individuals per launch for each K")
    ap.add_argument("--generations", type=int, default=10)
    ap.add_argument("--prewarm", type=float, default=0.5, help="seconds of untimed iterations before the warmup")
    ap.add_argument("--traffic-out", default=TRAFFIC_FILE,
                    help="--summarize-traffic: output file (default: the headline workload's summary)")
    ap.add_argument("--traffic-args", default="",
                    help="--summarize-traffic: the workload arguments the counter runs used (recorded)")
    ap.add_argument("--ttr-seeds", type=int, default=32,
                    help="default run: seeds of the time-to-R_F fields (0 leaves them out)")
    ap.add_argument("--ttr-group-seeds", type=int, default=8,
                    help="default run: seeds of the 8-individuals-per-launch time-to-R_F leg")
    ap.add_argument("--sw-threshold", type=float, default=None,
                    help="shrink-wrap threshold of the time-to-R_F schedules (default 0.03)")
    ap.add_argument("--differentiation", type=int, default=0, metavar="SEEDS",
                    help="run the low-count 1024^2 differentiation workload on SEEDS patterns (--generations)")
    ap.add_argument("--refine-steps", type=int, default=0, metavar="K",
                    help="time K Adam refinement steps at --n/--pop instead (e.g. --n 1024 --pop 32)")
    ap.add_argument("--batch", type=int, default=0, metavar="PATTERNS",
                    help="mixed-size batch workload of PATTERNS patterns (256 in the batching configuration)")
    ap.add_argument("--batch-pop", type=int, default=32)
    ap.add_argument("--batch-block", type=int, default=50)
    ap.add_argument("--no-graphs", action="store_true", help="batch: direct launches instead of HIP graphs")
    ap.add_argument("--batch-concurrent", action="store_true",
                    help="batch: run the size classes of a block side by side instead of in turn")
    ap.add_argument("--summarize-traffic", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="write the HBM traffic summary from two rocprofv3 --pmc runs of "
                         "`bench.py --calibrate MIB` (FETCH_SIZE, WRITE_SIZE) and exit; no GPU needed")
    args = ap.parse_args()

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    if args.summarize_traffic:
        summarize_traffic(*args.summarize_traffic, args.calibrate or 256, args.traffic_out, log,
                          (" " + args.traffic_args) if args.traffic_args else "")
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported (the population "
              "shards by individuals across GPUs; only the one-GPU path is built).", file=sys.stderr)
        sys.exit(2)

    if args.time_to_rf:
        time_to_rf_study(args.time_to_rf, args.generations, args.target, log, args.sw_threshold,
                         groups=[int(k) for k in args.plain_groups.split(",") if k])
        return

    if args.batch:
        batch_bench(args, log)
        return
    if args.differentiation:
        differentiation_study(args.differentiation, args.generations, log)
        return
    if args.refine_steps:
        refine_bench(args.n, args.pop, args.refine_steps, log)
        return

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
    # Untimed pre-warm of about half a second, so that the clocks have ramped up
    # whatever --warmup is, then the requested warmup steps.
    t_pre = time.perf_counter()
    while time.perf_counter() - t_pre < args.prewarm:
        be.iterate_async(args.alg, args.block, args.beta, 0.0)
        be.sync()
    log("warmup")
    for _ in range(args.warmup):
        be.iterate_async(args.alg, args.block, args.beta, 0.0)
    be.sync()

    log(f"timed: {args.steps} steps of {args.block} {args.alg} iterations, P={P}")
    # HIP events bracket every TIMING_STRIDE-th column and row launch of the
    # timed region (an event pair around every launch would cost ~9 % of it).
    be.timing(not args.no_timing, stride=TIMING_STRIDE)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        be.iterate_async(args.alg, args.block, args.beta, 0.0)
    be.sync()
    dt = time.perf_counter() - t0
    col_ms, col_cnt = be.timing_read(hipcore.KIND_COL_PROJ)
    row_ms, row_cnt = be.timing_read(hipcore.KIND_ROW_UPDATE)
    blk_ms, blk_cnt = be.timing_read(hipcore.KIND_BLOCK)
    be.timing(False)
    ind_iters = args.steps * args.block * P
    value = ind_iters / dt
    chunk_bytes, chunk = be.chunk()
    fused = be.fused()[1]
    recomputed = hipcore.row_variant_recomputes_state(n, be.variants()[1])
    col_b, row_b = algorithmic_bytes(n, min(P, chunk), args.alg, recomputed)   # per launch (one chunk)
    iter_bytes = (col_b + row_b) / min(P, chunk)
    iteration_frac = value * iter_bytes / 1e9 / HBM_PEAK_GBS
    if not fused:
        log(f"whole iteration: {value * iter_bytes / 1e9:.0f} GB/s algorithmic ({iteration_frac:.3f} of HBM peak)")
    roofline = None
    if fused and blk_cnt:
        # One launch per block, the field in LDS: bounded by the vector ALUs.
        avg = blk_ms / blk_cnt * 1e-3
        tf = fused_flops(n, P, args.block) / avg / 1e12
        log(f"on-chip block: {blk_cnt} launches, avg {avg * 1e6:.1f} us, {tf:.1f} TFLOP/s nominal FFT flops")
        roofline = {"bound": "valu", "achieved": round(tf, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(tf / FP32_PEAK_TFLOPS, 4), "traffic": None,
                    "kernel": rocprof_kernel_name("fused", n, args.alg, 0),
                    "algorithmic_flops_per_launch": fused_flops(n, P, args.block),
                    "avg_launch_us": round(avg * 1e6, 2)}
    elif col_cnt and row_cnt:
        col_avg = col_ms / col_cnt * 1e-3
        row_avg = row_ms / row_cnt * 1e-3
        col_gbs = col_b / col_avg / 1e9
        row_gbs = row_b / row_avg / 1e9
        # dominant kernel = the one with the larger share of the timed region
        cvar, rvar = be.variants()
        if row_ms >= col_ms:
            dom = {"kernel": rocprof_kernel_name("row", n, args.alg, rvar), "achieved": row_gbs,
                   "bytes_per_launch": row_b, "avg_launch_us": row_avg * 1e6}
        else:
            dom = {"kernel": rocprof_kernel_name("col", n, args.alg, cvar), "achieved": col_gbs,
                   "bytes_per_launch": col_b, "avg_launch_us": col_avg * 1e6}
        log(f"column pass: {col_cnt} launches, avg {col_avg * 1e6:.1f} us, {col_gbs:.0f} GB/s algorithmic")
        log(f"row pass:    {row_cnt} launches, avg {row_avg * 1e6:.1f} us, {row_gbs:.0f} GB/s algorithmic")
        roofline = {"bound": "hbm", "achieved": round(dom["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(dom["achieved"] / HBM_PEAK_GBS, 4),
                    "traffic": measured_traffic(dom["kernel"], log),
                    "kernel": dom["kernel"], "algorithmic_bytes_per_launch": dom["bytes_per_launch"],
                    "avg_launch_us": round(dom["avg_launch_us"], 2),
                    "iteration_frac": round(iteration_frac, 4)}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(amp, rho, supps, args.alg, args.beta, args.cpu_iters, log)
    cvars = be.variants()
    be.close()
    ttr = None
    if args.ttr_seeds > 0 and n == 512 and P == 64:
        ttr = time_to_rf_fields(args.ttr_seeds, args.generations, args.target, log, args.ttr_group_seeds)

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
        "config": {"workload": f"{'ellipsoids' if n >= 256 else 'polygon'} {n}^2 pattern, population {P}, {args.alg}(beta={args.beta}) "
                               f"blocks of {args.block} iterations, population resident in HBM",
                   "n": n, "population": P, "alg": args.alg, "iterations_per_step": args.block,
                   "kernel_variants": dict(zip(("column", "row"), cvars)),
                   "state": "recomputed" if recomputed else "stored",
                   "algorithmic_bytes_per_px": round(iter_bytes / (n * n), 3),
                   "chunk_individuals": min(P, chunk), "on_chip_blocks": fused},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "time_to_rf": ttr,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
