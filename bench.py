"""Throughput of the explicit compressible-MHD RK3 step on one MI355X.

Workload (default): 512x512x256 box, MHD with a 45-degree tilted rotation
axis (latitude f-plane), m=1 polytrope with density contrast 30, Ra=1e6,
Pr=1, Pm=2, Ta=1e6, arctan-stretched z grid (max/min spacing 3), seeded
uniform random T perturbation 1e-4 and a seed field from a random vector
potential in modes |k|<=8 at amplitude 1e-3.  Plain FP64 (the drop-in form,
bit-identical to the reference).  Synthetic initial data; adaptive dt.

A step is one full RK3 step (CFL reduction + 3 stages + boundary planes).
The timed region is K steps back to back with the inputs resident in HBM.
Prints one JSON line (see DESIGN.md for the field definitions).
"""

from __future__ import annotations

import argparse
import importlib
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK = 8.0e12            # MI355X HBM3E, bytes/s
BYTES_PER_ZONE_STEP = 640.0  # 8 fields x 10 accesses x 8 B (q, Z registers; one kernel per stage)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(npx, npy, npz, seed=20260):
    cfg = importlib.import_module("rast-3dmhd_amd.config")
    ics = importlib.import_module("rast-3dmhd_amd.ics")
    ph = cfg.physics(contrast=30.0, ra=1e6, pr=1.0, pm=2.0, ta=1e6, tilt_deg=45.0)
    par = {k: cfg.f32(v) for k, v in ph.items()}
    mt = cfg.metrics(npx, npy, npz, xmax=par[17], ymax=par[18], zmax=par[19], ngrid=1)
    t, r = cfg.polytrope(mt["zee"], par[3], par[4])
    meta = cfg.constants(npx, npy, npz, par, lmag=True, lrot=True, izcon=1, itcon=0, tb=float(t[-1]))
    meta.update(mt)
    st = ics.static_state(npx, npy, npz, t, r, ampt=1e-4, lmag=True, seed=seed, fast_noise=True)
    st["bx"], st["by"], st["bz"] = ics.vector_potential_field(meta, kmax=8, amp=1e-3, seed=seed)
    for k in ("ru", "rv", "rw", "ro", "tt"):
        ics.periodic_fill(st[k])
    return meta, st


def cpu_baseline_reference(exe, ranks, n1=1, n2=3, timeout=900):
    """Reference MPI build timed on this host: (wall(n2) - wall(n1)) / (n2 - n1)."""
    mpiexec = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
    walls = {}
    for n in (n1, n2):
        with tempfile.TemporaryDirectory() as d:
            env = dict(os.environ, RAST_NTOTAL=str(n), RAST_NSTEP0="100000")
            t0 = time.perf_counter()
            subprocess.run([mpiexec, "-launcher", "fork", "-n", str(ranks), exe], cwd=d, env=env,
                           check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                           timeout=timeout)
            walls[n] = time.perf_counter() - t0
            log(f"cpu baseline: reference NTOTAL={n} wall {walls[n]:.2f} s")
    return (walls[n2] - walls[n1]) / (n2 - n1)


def cpu_baseline_port(seconds=15.0):
    """The C oracle (single thread) stepping a 128x128x64 cut of the same
    physics for as many steps as fit in about `seconds`."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    O.build()
    meta, st = workload(128, 128, 64)
    o = O.Oracle(meta, st)
    n, t0 = 0, time.perf_counter()
    while True:
        o.step()
        n += 1
        el = time.perf_counter() - t0
        if el > seconds or n >= 50:
            break
    return el / n, n, 128 * 128 * 64


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--npx", type=int, default=512)
    ap.add_argument("--npy", type=int, default=512)
    ap.add_argument("--npz", type=int, default=256)
    ap.add_argument("--cpu-baseline", choices=["auto", "reference", "port", "none"], default="auto")
    ap.add_argument("--ranks", type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs the whole box resident on one GPU; "
              "multi-GPU runs are not part of it (see DESIGN.md, 'More than one GPU').", file=sys.stderr)
        sys.exit(2)

    pkg = importlib.import_module("rast-3dmhd_amd")
    pkg.load()
    npx, npy, npz = args.npx, args.npy, args.npz
    zones = npx * npy * npz
    log(f"building workload {npx}x{npy}x{npz}")
    meta, st = workload(npx, npy, npz)
    box = pkg.Box(meta)
    box.upload(st)
    del st
    log("warmup")
    box.steps(args.warmup)
    log(f"timing {args.steps} steps")
    t0 = time.perf_counter()
    dt, umach = box.steps(args.steps)
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / args.steps
    zu = zones * args.steps / wall
    # per-kernel durations from HIP events on the library's stream
    kt = box.time_steps(min(args.steps, 5))
    stage_bytes = zones * BYTES_PER_ZONE_STEP / 3.0
    achieved = stage_bytes / (kt["stage"] * 1e-3)
    box.close()

    cpu = None
    if args.cpu_baseline != "none":
        exe = os.path.join(REPO, "oracle", "_ref", "bench512", "3dmhd.exe")
        use_ref = args.cpu_baseline in ("auto", "reference") and os.path.exists(exe) and (npx, npy, npz) == (512, 512, 256)
        if use_ref:
            try:
                s_per_step = cpu_baseline_reference(exe, args.ranks)
                cpu = {"value": zones / s_per_step, "unit": "zone-updates/s", "cores": args.ranks,
                       "kind": "reference",
                       "sample": f"reference Fortran/MPI build, {npx}x{npy}x{npz}, {args.ranks} ranks "
                                 f"(NPEY=8,NPEZ=2), (wall(3 steps)-wall(1 step))/2 = {s_per_step:.2f} s/step"}
            except Exception as e:  # fall back to the port
                log(f"reference baseline failed: {e}")
        if cpu is None and args.cpu_baseline != "reference":
            s, n, z = cpu_baseline_port()
            cpu = {"value": z / s, "unit": "zone-updates/s", "cores": 1, "kind": "port",
                   "sample": f"C oracle (restated reference step), 128x128x64 cut of the workload, "
                             f"{n} steps, {s:.3f} s/step, 1 thread"}

    # HBM bytes per stage launch from the committed rocprofv3 --pmc summary of
    # this workload (FETCH_SIZE / WRITE_SIZE, calibrated on a plain copy)
    traffic = None
    tpath = os.path.join(REPO, "profiles", "k_stage_512x512x256_fp64_traffic.json")
    if (npx, npy, npz) == (512, 512, 256) and os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f)["bytes_per_launch_mean"]

    out = {
        "metric": "RK3 zone-updates/s (plain FP64 MHD step, 512x512x256) and %HBM peak",
        "value": zu,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (polytrope + seeded T noise + random-vector-potential seed field)",
        "config": {"workload": f"{npx}x{npy}x{npz} MHD + 45deg rotation, contrast 30, Ra 1e6, Pm 2, Ta 1e6, plain FP64",
                   "npx": npx, "npy": npy, "npz": npz, "parallelism": "single GPU, whole box resident"},
        "hbm_frac_step": zu * BYTES_PER_ZONE_STEP / HBM_PEAK,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "algorithmic_bytes_per_launch": stage_bytes,
                     "kernel": "k_stage (fused RHS + RK3 stage), mean of the 3 stage launches",
                     "ms_stage": kt["stage"], "ms_cfl": kt["cfl"], "ms_bcon": kt["bcon"],
                     "ms_step_events": kt["step"]},
        "cpu_baseline": cpu,
        "last_dt": dt,
        "last_umach": umach,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
