"""Accuracy of the mixed-precision step against the FP64 step.

Runs the same synthetic stratified box (workload.stratified_box) through
both forms of the library for N steps, each with its own adaptive dt, and
reports the per-field errors of the mixed state against the FP64 state
(workload.field_errors), the simulated times reached and the step times.

    python rast-3dmhd_amd/accuracy.py --npx 256 --npy 256 --npz 128 --steps 1000
"""

from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))


def run(npx, npy, npz, steps, precision, meta, st, chunk=100, log=None):
    pkg = importlib.import_module("rast-3dmhd_amd")
    with pkg.Box(meta, precision=precision) as box:
        box.upload(st)
        t0 = time.perf_counter()
        done = 0
        while done < steps:
            n = min(chunk, steps - done)
            dt, um = box.steps(n)
            done += n
            if log:
                log(f"{precision}: {done}/{steps} steps, dt {dt:.6e}, umach {um:.4e}")
        wall = time.perf_counter() - t0
        kt = box.time_steps(5)
        nit, timc, timt = box.clock()
        out = box.download()
    return out, {"wall_s": wall, "ms_step_events": kt["step"], "ms_stage": kt["stage"], "nit": nit,
                 "timc": timc, "last_dt": dt}


def compare(npx, npy, npz, steps, seed=20260, log=None):
    """Both forms on workload.stratified_box for `steps` steps."""
    wl = importlib.import_module("rast-3dmhd_amd.workload")
    meta, st = wl.stratified_box(npx, npy, npz, seed=seed)
    ref, t64 = run(npx, npy, npz, steps, "fp64", meta, st, log=log)
    mix, tmx = run(npx, npy, npz, steps, "mixed", meta, st, log=log)
    err = wl.field_errors(meta, mix, ref)
    return {"box": [npx, npy, npz], "steps": steps, "seed": seed,
            "workload": "contrast 10, Ra 1e5, Pr 1, Pm 1, Ta 1e4 vertical axis, tanh z grid, B seed |k|<=4 amp 1e-5",
            "fp64": t64, "mixed": tmx, "timc_rel_diff": abs(tmx["timc"] - t64["timc"]) / t64["timc"],
            "speedup_events": t64["ms_step_events"] / tmx["ms_step_events"], "errors": err}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--npx", type=int, default=256)
    ap.add_argument("--npy", type=int, default=256)
    ap.add_argument("--npz", type=int, default=128)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=20260)
    ap.add_argument("--out", default=None, help="write the JSON report here")
    a = ap.parse_args()

    def log(*x):
        print(*x, file=sys.stderr, flush=True)

    r = compare(a.npx, a.npy, a.npz, a.steps, seed=a.seed, log=log)
    s = json.dumps(r, indent=1)
    if a.out:
        with open(a.out, "w") as f:
            f.write(s + "\n")
    print(s)


if __name__ == "__main__":
    main()
