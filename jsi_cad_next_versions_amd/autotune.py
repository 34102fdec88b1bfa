"""Autotuner of the distance screens' launch and seed choices.

For each (degree, nuv) of a grid of workloads (the hard-wave pair built at
that degree, round-1 grid nuv, k = 4096, FP32SCREEN, polish 64: the
benchmark's query shape), it measures the served rate (a batching server with
4 workers, the same query queued n times) and picks, by coordinate descent
from the built-in defaults, the value of each knob of jsi_tune_set:

* ``seed_pairs``: seed cell pairs of the refinement rounds per unit of k (a
  tighter threshold costs a larger seed);
* ``filter_walk``: slots a refinement-filter workgroup walks (fewer, longer
  lived workgroups against more, shorter ones);
* ``r1_scale``: round-1 screen workgroups, percent of the default grid;
* ``r1_seeds``: round-1 seed tile pairs (fewer loosen the screen's threshold).

A change is kept when it is faster by more than ``--min-gain`` (run-to-run
noise).  The knobs change launch shapes and thresholds only, never a result.
Writes the table that ``runtime.Context`` loads (``runtime.TUNING_PATH``):

    python -m jsi_cad_next_versions_amd.autotune [--out PATH]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import surfaces as S
from ._abi import QUERY_DISTANCE, default_params
from .runtime import TUNING_PATH, Context, Server

CANDIDATES = {
    "seed_pairs": [8, 16, 32, 64, 128],
    "filter_walk": [2, 4, 8, 16],
    "r1_scale": [50, 100, 200, 400],
    "r1_seeds": [64, 128, 256],
}
DEFAULTS = {"seed_pairs": 32, "filter_walk": 8, "r1_scale": 100, "r1_seeds": 256}


WORKERS = 4


def served_rate(srv, q, n, reps):
    """Median queries/s over `reps` runs of n queued copies of q (after a
    warm-up that lets every worker build the batch plans of the current
    choices)."""
    for _ in range(2):
        for t in [srv.submit(q) for _ in range(64 * WORKERS)]:
            srv.poll(t)
    rates = []
    for _ in range(reps):
        t0 = time.perf_counter()
        for t in [srv.submit(q) for _ in range(n)]:
            srv.poll(t)
        rates.append(n / (time.perf_counter() - t0))
    return float(np.median(rates))


def tune_entry(ctx, srv, degree, nuv, n, reps, min_gain, log):
    a, b = S.hard_wave_pair(deg=degree)
    p = default_params(k=4096, precision="fp32screen", polish=64)
    q = ctx.query(QUERY_DISTANCE, a, S.full_region(a, nuv, 1), b, S.full_region(b, nuv, 2), p)
    choice = dict(DEFAULTS)

    def apply(ch):
        for knob, v in ch.items():
            ctx.tune_set(degree, nuv, knob, v)

    apply(choice)
    base = served_rate(srv, q, n, reps)
    best = base
    trials = [{"choice": dict(choice), "queries_per_s": base}]
    for _ in range(2):  # two passes of coordinate descent
        changed = False
        for knob, values in CANDIDATES.items():
            for v in values:
                if v == choice[knob]:
                    continue
                ch = dict(choice, **{knob: v})
                apply(ch)
                r = served_rate(srv, q, n, reps)
                trials.append({"choice": ch, "queries_per_s": r})
                log(f"  degree {degree} nuv {nuv} {ch}: {r:.0f} q/s")
                if r > best * (1.0 + min_gain):
                    best, choice, changed = r, ch, True
            apply(choice)
        if not changed:
            break
    # the reported rates: both forms measured again, side by side
    apply(DEFAULTS)
    base = served_rate(srv, q, n, reps)
    apply(choice)
    best = served_rate(srv, q, n, reps)
    return {"degree": degree, "nuv": nuv, "choice": choice, "default_queries_per_s": base,
            "tuned_queries_per_s": best, "gain": best / base, "trials": trials}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=TUNING_PATH)
    ap.add_argument("--degrees", default="2,3,5")
    ap.add_argument("--nuvs", default="128,256,512")
    ap.add_argument("--queries", type=int, default=320, help="queued copies per timed run")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--min-gain", type=float, default=0.02)
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    ctx = Context(0, tuning=False)
    entries = []
    with Server(ctx, slots=WORKERS, max_pending=1 << 20) as srv:
        for d in (int(x) for x in args.degrees.split(",")):
            for nuv in (int(x) for x in args.nuvs.split(",")):
                log(f"degree {d} nuv {nuv}")
                e = tune_entry(ctx, srv, d, nuv, args.queries, args.reps, args.min_gain, log)
                log(f"degree {d} nuv {nuv}: {e['choice']} {e['default_queries_per_s']:.0f} -> "
                    f"{e['tuned_queries_per_s']:.0f} q/s")
                entries.append(e)
    table = {"workload": "hard-wave pair at each degree, round-1 nuv, k=4096, FP32SCREEN, polish 64; served by "
                         "4 batching workers",
             "knobs": CANDIDATES, "defaults": DEFAULTS, "entries": entries}
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    log(f"wrote {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
