"""Autotuning of the iteration-pass variants on the target GPU.

For every supported size from 256 up, a population of synthetic patterns of
about 2^24 pixels in total (64 individuals at 512^2) is made resident and
spr_autotune times every column and row variant on it in whole iterations (a
column pass then a row pass, `--reps` iterations each, `--rounds` rounds; a
variant's time is its median over the rounds): the column variants with the
handle's row variant, the row variants with the chosen column variant. The
fastest variant of each pass (the lowest-numbered one within 1.5 % of the
fastest) per size is written as the compiled default table (spr_tuned.hpp) and
the iteration times as JSON (`column_ms`, `row_ms`: ms per iteration with each
variant); rebuild the library afterwards.

    python -m s_spi_spring_amd.autotune --header OUT.hpp --json OUT.json
"""
from __future__ import annotations

import argparse
import datetime
import json
import sys

import numpy as np

SIZES = (256, 384, 512, 768, 1024, 2048)


TIE = 1.015


def pick(ms):
    """Lowest-numbered variant within TIE of the fastest (near-ties are timing
    noise; the lower index keeps the table stable between tuning runs)."""
    ms = np.asarray(ms)
    return int(np.nonzero(ms <= ms.min() * TIE)[0][0])


def tune_size(n, reps, rounds, log):
    from s_spi_spring_amd import hipcore, synth
    P = max(1, (1 << 24) // (n * n))
    I, _ = synth.ellipsoids_pattern(seed=0, n=n) if n <= 512 else synth.octahedron_pattern(seed=0, n=n)
    amp = synth.corner_amplitude(I)
    supp = synth.disk_support(n, max(8, int(0.3 * n)))
    rho = synth.random_start(n, P, supp, seed=1, scale=float(np.sqrt(np.maximum(I, 0).sum()) / supp.sum()))
    be = hipcore.Backend(amp, capacity=P)
    be.set_population(rho, np.broadcast_to(supp, rho.shape).astype(np.int32))
    cms, rms = [], []
    for _ in range(rounds):
        cm, rm = be.autotune(reps)
        cms.append(cm)
        rms.append(rm)
    be.close()
    cm, rm = np.median(cms, 0), np.median(rms, 0)
    best = (pick(cm), pick(rm))
    log(f"n={n} P={P}: column ms {np.round(cm, 4).tolist()} row ms {np.round(rm, 4).tolist()} -> {best}")
    return {"n": n, "population": P, "column_ms": cm.tolist(), "row_ms": rm.tolist(),
            "column_variant": best[0], "row_variant": best[1]}


def header_text(results, device):
    rows = ",\n".join(f"    {{{r['n']}, {r['column_variant']}, {r['row_variant']}}}" for r in results)
    return f"""// Variant choices of the iteration passes per grid size, written by
// s_spi_spring_amd/autotune.py from spr_autotune timings on {device};
// sizes not listed use variant 0.
#pragma once

namespace spr {{

struct TunedVariants {{
    int n, col, row;
}};
constexpr TunedVariants kTuned[] = {{
{rows},
}};

}}  // namespace spr
"""


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--rounds", type=int, default=3)
    ap.add_argument("--sizes", type=int, nargs="*", default=list(SIZES))
    ap.add_argument("--header", required=True, help="output path of the variant table header")
    ap.add_argument("--json", required=True, help="output path of the timings")
    ap.add_argument("--label", default=None, help="device label for the header (default: the HIP device name)")
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    from s_spi_spring_amd import hipcore
    lib = hipcore.load()
    import ctypes
    name = ctypes.create_string_buffer(256)
    device = "device 0"
    if args.label:
        device = args.label
    elif lib.spr_device_name(0, name, 256) == 0:
        device = name.value.decode()
    results = [tune_size(n, args.reps, args.rounds, log) for n in args.sizes]
    with open(args.header, "w") as f:
        f.write(header_text(results, device))
    with open(args.json, "w") as f:
        json.dump({"device": device, "date": datetime.date.today().isoformat(), "reps": args.reps,
                   "rounds": args.rounds, "sizes": results}, f, indent=1)
    log(f"wrote {args.header} and {args.json}")


if __name__ == "__main__":
    main()
