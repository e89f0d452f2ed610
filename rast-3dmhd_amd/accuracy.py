"""Accuracy of the mixed-precision step against the FP64 step.

Runs the same synthetic stratified box (workload.stratified_box) through
both forms of the library for N steps, at one fixed dt (default) or each
with its own CFL dt, and reports the per-field errors of the mixed state against the FP64 state
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


def run(npx, npy, npz, steps, precision, meta, st, chunk=100, log=None, fixed_dt=0.0):
    pkg = importlib.import_module("rast-3dmhd_amd")
    with pkg.Box(meta, precision=precision) as box:
        box.upload(st)
        if fixed_dt:
            box.set_fixed_dt(fixed_dt)
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


def compare(npx, npy, npz, steps, seed=20260, log=None, fixed=True):
    """Both forms on workload.stratified_box for `steps` steps, at the fixed
    dt of the FP64 form's first CFL step (fixed=True: equal simulated
    times) or each with its own CFL dt."""
    pkg = importlib.import_module("rast-3dmhd_amd")
    wl = importlib.import_module("rast-3dmhd_amd.workload")
    meta, st = wl.stratified_box(npx, npy, npz, seed=seed)
    dt0 = 0.0
    if fixed:
        with pkg.Box(meta) as box:
            box.upload(st)
            dt0 = box.step()[0]
    ref, t64 = run(npx, npy, npz, steps, "fp64", meta, st, log=log, fixed_dt=dt0)
    mix, tmx = run(npx, npy, npz, steps, "mixed", meta, st, log=log, fixed_dt=dt0)
    err = wl.field_errors(meta, mix, ref)
    return {"box": [npx, npy, npz], "steps": steps, "seed": seed, "dt": "fixed" if fixed else "adaptive",
            "fixed_dt": dt0 or None,
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
    ap.add_argument("--adaptive", action="store_true", help="each form with its own CFL dt")
    ap.add_argument("--out", default=None, help="write the JSON report here")
    a = ap.parse_args()

    def log(*x):
        print(*x, file=sys.stderr, flush=True)

    r = compare(a.npx, a.npy, a.npz, a.steps, seed=a.seed, log=log, fixed=not a.adaptive)
    s = json.dumps(r, indent=1)
    if a.out:
        with open(a.out, "w") as f:
            f.write(s + "\n")
    print(s)


if __name__ == "__main__":
    main()
