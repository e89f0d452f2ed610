"""Writes the driver fixtures in this directory from runs of the reference's
own main program.

Runs only in the build container, where /root/reference and amdflang exist.
For each box of oracle/refconfigs.DRIVER_CONFIGS it compiles the reference's
3dmhd.f + 3dmhdsub.f with a generated parameter block and setup routine
(oracle/build_ref.py) and runs it under MPI twice:

  <name>/          a fresh start (STATIC) for NTOTAL steps with a dump after
                   every step: the per-rank rast3d.dat0.NNNN direct-access
                   files, rast3d.par and the 80-byte-record rast3d.lis log,
                   exactly as the reference wrote them;
  <name>_restart/  a restart from dump 1 of that run (NSTART = 1) for
                   NTOTAL - 1 steps: the same three kinds of file.

These are the reference's output files, kept as data.

Usage: python tests/golden/make_driver_golden.py [name ...]
"""

from __future__ import annotations

import importlib.util
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))


def _load(path, name):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


build_ref = _load(os.path.join(REPO, "oracle", "build_ref.py"), "build_ref")
CONFIGS = build_ref.refconfigs.DRIVER_CONFIGS
OUTPUTS = ("rast3d.par", "rast3d.lis")


def outputs(run_dir: str, npe: int):
    return list(OUTPUTS) + [f"rast3d.dat0.{r:04d}" for r in range(npe)]


def run_reference(exe: str, run_dir: str, npe: int, ntotal: int, nstart: int = 0) -> None:
    env = dict(os.environ, RAST_NTOTAL=str(ntotal), RAST_NSTEP0="1", RAST_NSTART=str(nstart))
    r = subprocess.run([build_ref.MPIEXEC, "-launcher", "fork", "-n", str(npe), exe], cwd=run_dir, env=env,
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{exe} failed:\n{r.stdout}\n{r.stderr}")


def main(names):
    if not build_ref.available():
        sys.exit("reference sources or amdflang not available; fixtures are committed")
    for name in names or list(CONFIGS):
        cfg = CONFIGS[name]
        d = build_ref.build(name, cfg, main_exe=True)
        exe = os.path.join(d, "3dmhd.exe")
        p = cfg["params"]
        npe = int(p["NPEY"]) * int(p["NPEZ"])
        n = int(cfg["nsteps"])
        run = os.path.join(d, "run")
        shutil.rmtree(run, ignore_errors=True)
        os.makedirs(run)
        run_reference(exe, run, npe, n)
        rst = os.path.join(d, "run_restart")
        shutil.rmtree(rst, ignore_errors=True)
        os.makedirs(rst)
        for f in outputs(run, npe):
            if f != "rast3d.lis":
                shutil.copy(os.path.join(run, f), rst)
        run_reference(exe, rst, npe, n - 1, nstart=1)
        for src, dst in ((run, name), (rst, name + "_restart")):
            out = os.path.join(HERE, dst)
            shutil.rmtree(out, ignore_errors=True)
            os.makedirs(out)
            for f in outputs(src, npe):
                shutil.copy(os.path.join(src, f), out)
            size = sum(os.path.getsize(os.path.join(out, f)) for f in os.listdir(out))
            print(f"{dst}: {len(os.listdir(out))} files, {size / 1024:.0f} KiB", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
