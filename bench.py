"""Throughput of the explicit compressible-MHD RK3 step on one MI355X.

Headline workload: 512x512x256 box, MHD with a 45-degree tilted rotation
axis (latitude f-plane), m=1 polytrope with density contrast 30, Ra=1e6,
Pr=1, Pm=2, Ta=1e6, arctan-stretched z grid (max/min spacing 3), seeded
uniform random T perturbation 1e-4 and a seed field from a random vector
potential in modes |k|<=8 at amplitude 1e-3, 500 RK3 steps with adaptive dt.
`value` is the mixed-precision form (FP32 perturbations about the FP64
hydrostatic background, FP64 accumulation); the plain FP64 form (the
drop-in, bit-identical to the reference) is timed on the same box beside it.

Ensemble: 64 independent boxes (sizes 32..96 x 32..96 x 32..64, random Ra,
Ta, contrast, own seeds and dt), mixed precision, autotuned tiling, HIP-graph
replay, against the same boxes stepped one at a time.

Accuracy of the mixed form: a 256x256x128 box (contrast 10, Ra 1e5, Pr 1,
Pm 1, Ta 1e4 vertical axis, tanh z grid, seed field |k|<=4 at 1e-5) is run
1000 steps in both forms and the per-field errors are reported.

A step is one full RK3 step (CFL reduction + 3 stages + boundary planes).
The timed region is K steps back to back with the inputs resident in HBM.
Prints one JSON line.
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

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
# algorithmic HBM bytes per zone per RK3 step (8 fields, one kernel per stage)
BYTES_PER_ZONE_STEP = {
    "fp64": 640.0,   # q read + Z read/write + q write: 10 accesses x 8 B
    # stage 1: h, R (BF16) read, h, D written; stage 2: h, D read and written;
    # stage 3: h, D read, h, R written: (14 + 16 + 14) B x 8 fields
    "mixed": 352.0,
}
KERNEL = {"fp64": "k_stage (fused RHS + RK3 stage, FP64)",
          "mixed": "k_stage_m (fused FP32 perturbation RHS + FP64 RK3 accumulation)"}
TRAFFIC = {"fp64": "k_stage_512x512x256_fp64_traffic.json", "mixed": "k_stage_m_512x512x256_mixed_traffic.json"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def wl():
    return importlib.import_module("rast-3dmhd_amd.workload")


def workload(npx, npy, npz, seed=20260):
    return wl().convection_box(npx, npy, npz, seed=seed)


def timed(pkg, meta, st, precision, steps, warmup):
    box = pkg.Box(meta, precision=precision)
    try:
        box.upload(st)
        log(f"{precision}: warmup {warmup}")
        box.steps(warmup)
        log(f"{precision}: timing {steps} steps")
        t0 = time.perf_counter()
        dt, umach = box.steps(steps)
        wall = time.perf_counter() - t0
        # per-kernel durations from HIP events on the library's stream
        kt = box.time_steps(min(steps, 5))
    finally:
        box.close()
    return {"wall": wall, "ms_per_step": 1e3 * wall / steps, "kt": kt, "dt": dt, "umach": umach}


def roofline(precision, zones, kt, shape):
    stage_bytes = zones * BYTES_PER_ZONE_STEP[precision] / 3.0
    achieved = stage_bytes / (kt["stage"] * 1e-3)
    traffic = None
    tpath = os.path.join(REPO, "profiles", TRAFFIC[precision])
    if shape == (512, 512, 256) and os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f)["bytes_per_launch_mean"]
    return {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": traffic, "algorithmic_bytes_per_launch": stage_bytes,
            "kernel": KERNEL[precision] + ", mean of the 3 stage launches",
            "ms_stage": kt["stage"], "ms_cfl": kt["cfl"], "ms_bcon": kt["bcon"], "ms_step_events": kt["step"]}


def accuracy(npx=256, npy=256, npz=128, steps=1000, fixed=True):
    acc = importlib.import_module("rast-3dmhd_amd.accuracy")
    r = acc.compare(npx, npy, npz, steps, log=log, fixed=fixed)
    e = r["errors"]
    return {"box": r["box"], "steps": steps, "workload": r["workload"], "dt": r["dt"], "fixed_dt": r["fixed_dt"],
            "rel_l2": e["rel_l2"], "max_rel_l2": e["max_rel_l2"], "velocity_rel_l2": e["velocity"],
            "entropy_rel_fluct_l2": e["entropy"], "rho_rel_fluct_l2": e["fluct_l2"]["ro"],
            "tt_rel_fluct_l2": e["fluct_l2"]["tt"], "b_rel_l2": e.get("b_field"),
            "divb_max_mixed": e["divb_max"]["test"], "divb_max_fp64": e["divb_max"]["ref"],
            "timc_rel_diff": r["timc_rel_diff"],
            "ms_step_fp64": r["fp64"]["ms_step_events"], "ms_step_mixed": r["mixed"]["ms_step_events"]}


def ensemble(pkg, nbox=64, steps=100, precision="mixed", tiling=None):
    """The random 64-box ensemble: autotuned tiling (or the given
    (tile_x, tile_y, kdepth, grid), with no autotuning), graph-replayed steps,
    against the same boxes stepped one at a time through the single-box
    library."""
    specs = wl().ensemble_specs(nbox)
    boxes = [wl().ensemble_box(s) for s in specs]
    with pkg.Ensemble([m for m, _ in boxes], precision=precision) as ens:
        for b, (_, s) in enumerate(boxes):
            ens.upload(b, s)
        report = None
        if tiling is not None:
            ens.set_tile(*tiling)
        ens.steps(2)
        if tiling is None:
            report = ens.autotune(nsteps=3)
            log("ensemble autotune: " + json.dumps(report))
        ms = ens.time_steps(steps)
        zones = ens.zones()
        tiling = ens.tile()
    ms_single = 0.0
    for m, s in boxes:
        with pkg.Box(m, precision=precision) as box:
            box.upload(s)
            box.steps(2)
            ms_single += box.time_steps(10)["step"]
    return {"boxes": nbox, "zones": zones, "precision": precision, "steps": steps,
            "ms_per_step": ms, "value": zones / (ms * 1e-3), "unit": "zone-updates/s",
            "hbm_frac_step": zones / (ms * 1e-3) * BYTES_PER_ZONE_STEP[precision] / HBM_PEAK,
            "tiling": {"tile_x": tiling[0], "tile_y": tiling[1], "kdepth": tiling[2], "grid": tiling[3]},
            "autotune_candidates": len(report) if report else 0,
            "autotune_ms_range": [min(r["ms_step"] for r in report), max(r["ms_step"] for r in report)]
                                 if report else None,
            "ms_per_step_boxes_one_at_a_time": ms_single, "speedup_vs_one_at_a_time": ms_single / ms}


def host_cores():
    """Cores this process may use: the physical cores (distinct physical id,
    core id) among the CPUs in its affinity mask, capped by the CPU share the
    environment gives (OMP_NUM_THREADS, when set); and the CPU model."""
    cpus = sorted(os.sched_getaffinity(0))
    phys, model = {}, "unknown"
    try:
        cur = {}
        with open("/proc/cpuinfo") as f:
            for line in f.read().split("\n") + [""]:
                if not line.strip():
                    if "processor" in cur:
                        phys[int(cur["processor"])] = (cur.get("physical id", "0"), cur.get("core id", cur["processor"]))
                        model = cur.get("model name", model)
                    cur = {}
                elif ":" in line:
                    k, v = line.split(":", 1)
                    cur[k.strip()] = v.strip()
    except OSError:
        pass
    npc = len({phys.get(c, ("?", str(c))) for c in cpus}) if phys else len(cpus)
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    cores = min(npc, share) if share > 0 else npc
    return {"cores": max(1, cores), "physical_cores_in_affinity": npc, "cpus_in_affinity": len(cpus),
            "cpu_share": share or None, "model": model}


def baseline_exe(cores):
    """The largest reference build (oracle/_ref/bench512_r<ranks>) with
    ranks <= cores; (path, ranks) or (None, 0)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    try:
        import refconfigs
    finally:
        sys.path.pop(0)
    best = (None, 0)
    for r in refconfigs.BASELINE_RANKS:
        exe = os.path.join(REPO, "oracle", "_ref", refconfigs.baseline_name(r), "3dmhd.exe")
        if r <= cores and os.path.exists(exe) and r > best[1]:
            best = (exe, r)
    return best


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
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--npx", type=int, default=512)
    ap.add_argument("--npy", type=int, default=512)
    ap.add_argument("--npz", type=int, default=256)
    ap.add_argument("--precision", choices=["both", "mixed", "fp64"], default="both")
    ap.add_argument("--accuracy-steps", type=int, default=1000, help="0 skips the mixed-vs-FP64 error run")
    ap.add_argument("--ensemble-boxes", type=int, default=64, help="0 skips the batched ensemble run")
    ap.add_argument("--ensemble-tiling", type=int, nargs=4, metavar=("TILE_X", "TILE_Y", "KDEPTH", "GRID"),
                    help="fixed ensemble tiling instead of autotuning (for profiling one tiling)")
    ap.add_argument("--cpu-baseline", choices=["auto", "reference", "port", "none"], default="auto")
    ap.add_argument("--ranks", type=int, default=0,
                    help="MPI ranks of the reference CPU baseline (default: the largest build that fits the "
                         "host's physical cores and CPU share)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus}: the whole box runs resident on one GPU; runs on more than "
              "one GPU are not part of this implementation (use --gpus 1).", file=sys.stderr)
        sys.exit(2)

    pkg = importlib.import_module("rast-3dmhd_amd")
    pkg.load()
    shape = (args.npx, args.npy, args.npz)
    zones = args.npx * args.npy * args.npz
    log(f"building workload {shape}")
    meta, st = workload(*shape)
    forms = ["mixed", "fp64"] if args.precision == "both" else [args.precision]
    res = {p: timed(pkg, meta, st, p, args.steps, args.warmup) for p in forms}
    del st
    head = forms[0]
    zu = {p: zones * args.steps / r["wall"] for p, r in res.items()}

    ens = None
    if args.ensemble_boxes > 0:
        log(f"ensemble: {args.ensemble_boxes} boxes")
        ens = ensemble(pkg, nbox=args.ensemble_boxes, tiling=args.ensemble_tiling)

    acc = None
    if args.accuracy_steps > 0 and args.precision == "both":
        log(f"accuracy: mixed vs fp64, 256x256x128, {args.accuracy_steps} steps")
        acc = accuracy(steps=args.accuracy_steps)
        acc["adaptive_dt"] = accuracy(steps=args.accuracy_steps, fixed=False)

    cpu = None
    if args.cpu_baseline != "none":
        host = host_cores()
        exe, ranks = baseline_exe(args.ranks or host["cores"])
        use_ref = args.cpu_baseline in ("auto", "reference") and exe is not None and shape == (512, 512, 256)
        if use_ref:
            try:
                s_per_step = cpu_baseline_reference(exe, ranks)
                cpu = {"value": zones / s_per_step, "unit": "zone-updates/s", "cores": ranks,
                       "kind": "reference",
                       "sample": f"reference Fortran/MPI build (plain FP64), {args.npx}x{args.npy}x{args.npz}, "
                                 f"{ranks} ranks (NPEY={ranks // 2},NPEZ=2), (wall(3 steps)-wall(1 step))/2 "
                                 f"= {s_per_step:.2f} s/step",
                       "host": host}
            except Exception as e:  # fall back to the port
                log(f"reference baseline failed: {e}")
        if cpu is None and args.cpu_baseline != "reference":
            s, n, z = cpu_baseline_port()
            cpu = {"value": z / s, "unit": "zone-updates/s", "cores": 1, "kind": "port",
                   "sample": f"C oracle (restated reference step), 128x128x64 cut of the workload, "
                             f"{n} steps, {s:.3f} s/step, 1 thread"}

    forms_out = {}
    for p, r in res.items():
        forms_out[p] = {"value": zu[p], "unit": "zone-updates/s", "ms_per_step": r["ms_per_step"],
                        "hbm_frac_step": zu[p] * BYTES_PER_ZONE_STEP[p] / HBM_PEAK,
                        "bytes_per_zone_step": BYTES_PER_ZONE_STEP[p],
                        "roofline": roofline(p, zones, r["kt"], shape), "last_dt": r["dt"], "last_umach": r["umach"]}
    out = {
        "metric": "RK3 zone-updates/s + %HBM peak; mixed vs FP64 at stated error",
        "value": zu[head],
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res[head]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head == "mixed" else "f64",
        "data": "synthetic (polytrope + seeded T noise + random-vector-potential seed field)",
        "config": {"workload": f"{args.npx}x{args.npy}x{args.npz} MHD + 45deg rotation, contrast 30, Ra 1e6, "
                               f"Pm 2, Ta 1e6, {args.steps} RK3 steps, "
                               + ("mixed precision" if head == "mixed" else "plain FP64"),
                   "npx": args.npx, "npy": args.npy, "npz": args.npz,
                   "precision": head, "parallelism": "single GPU, whole box resident"},
        "hbm_frac_step": forms_out[head]["hbm_frac_step"],
        "roofline": forms_out[head]["roofline"],
        "cpu_baseline": cpu,
        "forms": forms_out,
        "mixed_speedup_vs_fp64": (zu["mixed"] / zu["fp64"]) if len(res) == 2 else None,
        "mixed_accuracy": acc,
        "ensemble": ens,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
