"""Recipe that compiles the Fortran reference where it lies and runs it
(TEST INFRASTRUCTURE ONLY; needs /root/reference, so it runs only in the
build container).

The reference sizes its arrays from an INCLUDEd PARAMETER block, so there is
one build per box.  For each box a directory oracle/_ref/<name>/ receives the
generated 3dmhdparam.f / 3dmhdset.f and *symlinks* to the reference sources;
Fortran INCLUDE resolves next to the (symlinked) source, so the reference
compiles against the generated parameter block without its source being
copied.  The symlinks are removed once the objects exist, so oracle/_ref
holds only generated files, objects, executables and run output.

Toolchain: amdflang (ROCm flang) with the conda MPICH Fortran bindings,
-O2 on baseline x86-64 with FP contraction off.
"""

from __future__ import annotations

import importlib.util
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
REF_SRC = "/root/reference"
REF_OUT = os.path.join(HERE, "_ref")
MPI_INC = "/opt/conda/include"
MPI_LIB = "/opt/conda/lib"
MPIEXEC = "/opt/conda/bin/mpiexec"
FC = "amdflang"
FFLAGS = ["-O2", "-ffixed-form", "-ffp-contract=off", f"-I{MPI_INC}"]


def _load(path, name):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


mkparam = _load(os.path.join(REPO, "rast-3dmhd_amd", "fortran", "mkparam.py"), "mkparam")
refconfigs = _load(os.path.join(HERE, "refconfigs.py"), "refconfigs")


def available() -> bool:
    return os.path.isfile(os.path.join(REF_SRC, "3dmhdsub.f")) and shutil.which(FC) is not None


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)} failed in {cwd}:\n{r.stdout}\n{r.stderr}")
    return r


# Statements of the reference main program that the harness runs as they
# are: the COMMON set-up (constants, SETUP, IPAR/PAR to COMMON) and the
# conductivity profile.  (1-based line range of 3dmhd.f, expected first and
# last statement.)
DRIVER_BLOCKS = {
    "drvset1.inc": (105, 240, "IXC=IXCON", "endif"),
    "drvset2.inc": (269, 279, "IF (.NOT.LREM) THEN", "ENDIF"),
}


def write_driver_blocks(d: str) -> None:
    """Extract DRIVER_BLOCKS from the reference's 3dmhd.f into the build
    directory d (git-ignored): the harness INCLUDEs them, so the reference's
    own set-up statements are never kept in this repository."""
    with open(os.path.join(REF_SRC, "3dmhd.f")) as f:
        lines = f.read().split("\n")
    for name, (a, b, first, last) in DRIVER_BLOCKS.items():
        block = lines[a - 1:b]
        if block[0].strip() != first or block[-1].strip() != last:
            raise RuntimeError(f"3dmhd.f:{a}-{b} is not the expected block for {name}")
        with open(os.path.join(d, name), "w") as f:
            f.write("\n".join(block) + "\n")


def build(name: str, cfg: dict | None = None, main_exe: bool = False) -> str:
    """Compile the reference for one box; returns the build directory."""
    cfg = cfg or refconfigs.CONFIGS[name]
    d = os.path.join(REF_OUT, name)
    os.makedirs(d, exist_ok=True)
    mkparam.write_config(d, cfg["params"], par=cfg["par"], ipar=cfg.get("ipar"),
                         env_steps=main_exe, kappa=cfg.get("kappa"))
    write_driver_blocks(d)
    links = {"3dmhdsub.f": os.path.join(REF_SRC, "3dmhdsub.f"),
             "refharness.f": os.path.join(HERE, "refharness.f")}
    if main_exe:
        links["3dmhd.f"] = os.path.join(REF_SRC, "3dmhd.f")
    try:
        for ln, target in links.items():
            p = os.path.join(d, ln)
            if os.path.lexists(p):
                os.remove(p)
            os.symlink(target, p)
        for src in links:
            _run([FC, *FFLAGS, "-c", src, "-o", src[:-2] + ".o"], d)
        _run([FC, *FFLAGS, "-c", "3dmhdset.f", "-o", "3dmhdset.o"], d)
    finally:
        for ln in links:
            p = os.path.join(d, ln)
            if os.path.islink(p):
                os.remove(p)
    libs = [f"-L{MPI_LIB}", f"-Wl,-rpath,{MPI_LIB}", "-lmpifort", "-lmpi"]
    _run([FC, "refharness.o", "3dmhdsub.o", "3dmhdset.o", *libs, "-o", "harness.exe"], d)
    if main_exe:
        _run([FC, "3dmhd.o", "3dmhdsub.o", "3dmhdset.o", *libs, "-o", "3dmhd.exe"], d)
    return d


FORTRAN_DIR = os.path.join(REPO, "rast-3dmhd_amd", "fortran")
MPI_LIBS = [f"-L{MPI_LIB}", f"-Wl,-rpath,{MPI_LIB}", "-lmpifort", "-lmpi"]


def build_dropin(name: str, cfg: dict | None = None) -> str:
    """The reference's own main program and set-up routines, unchanged, with
    the GPU STEP shim in place of its STEP: compiles the reference for the
    box (build()), makes the reference's STEP in 3dmhdsub.o a weak symbol
    (objcopy, on the object compiled here) so that the shim's strong STEP is
    the one linked, and links 3dmhd.o + 3dmhdsub.o + 3dmhdset.o + the shim +
    librast3dmhd.so into oracle/_ref/<name>/3dmhd_gpu.exe.  Returns its path."""
    cfg = cfg or refconfigs.DRIVER_CONFIGS[name]
    d = build(name, cfg, main_exe=True)
    _run(["objcopy", "--weaken-symbol=step_", "3dmhdsub.o", "3dmhdsub_nostep.o"], d)
    _run([FC, "-O2", "-J", d, "-c", os.path.join(FORTRAN_DIR, "rast3dmhd_c.f90"), "-o", "rast3dmhd_c.o"], d)
    _run([FC, "-O2", "-ffixed-form", "-I", d, f"-I{MPI_INC}", "-J", d, "-c",
          os.path.join(FORTRAN_DIR, "step_shim.f"), "-o", "step_shim.o"], d)
    libdir = os.path.join(REPO, "rast-3dmhd_amd")
    rpath = "$ORIGIN/" + os.path.relpath(libdir, d)
    exe = os.path.join(d, "3dmhd_gpu.exe")
    _run([FC, "3dmhd.o", "3dmhdsub_nostep.o", "3dmhdset.o", "step_shim.o", "rast3dmhd_c.o",
          f"-L{libdir}", "-lrast3dmhd", f"-Wl,-rpath,{rpath}", *MPI_LIBS, "-o", exe], d)
    syms = _run(["nm", exe], d).stdout
    if not any(line.split()[-1:] == ["step_"] and " T " in line for line in syms.splitlines()):
        raise RuntimeError(f"{exe}: STEP is not the shim's strong symbol")
    return exe


def build_dropins() -> list:
    return [build_dropin(n) for n in refconfigs.DRIVER_CONFIGS]


def build_baseline() -> list:
    """Compile the reference main program for the throughput workload
    (refconfigs.baseline_config) once per rank count of
    refconfigs.BASELINE_RANKS (NPEZ=2), the CPU baseline of bench.py, into
    oracle/_ref/bench512_r<ranks>/3dmhd.exe.  Skips a compile when the
    executable is newer than this recipe."""
    exes = []
    for ranks in refconfigs.BASELINE_RANKS:
        name = refconfigs.baseline_name(ranks)
        cfg = refconfigs.baseline_config(npey=ranks // 2, npez=2)
        exe = os.path.join(REF_OUT, name, "3dmhd.exe")
        if not (os.path.exists(exe) and os.path.getmtime(exe) > max(
                os.path.getmtime(__file__), os.path.getmtime(os.path.join(HERE, "refconfigs.py")))):
            build(name, cfg, main_exe=True)
        exes.append(exe)
    return exes


def run_harness(name: str, cfg: dict | None = None) -> str:
    """Run the harness for one box under MPI; returns the output directory."""
    cfg = cfg or refconfigs.CONFIGS[name]
    d = os.path.join(REF_OUT, name)
    out = os.path.join(d, "run")
    if os.path.isdir(out):
        shutil.rmtree(out)
    os.makedirs(out)
    dumps = cfg["dumps"]
    with open(os.path.join(out, "harness.in"), "w") as f:
        f.write(f"{cfg['steps']}\n")
        f.write(f"{len(dumps)} " + " ".join(str(x) for x in dumps) + "\n")
        f.write(f"{cfg['ampu']} {cfg['ampb']} {cfg['seed']}\n")
    p = cfg["params"]
    npe = int(p.get("NPEY", 1)) * int(p.get("NPEZ", 2))
    _run([MPIEXEC, "-launcher", "fork", "-n", str(npe), os.path.join(d, "harness.exe")], out)
    return out


if __name__ == "__main__":
    if not available():
        sys.exit("reference sources or amdflang not available")
    names = sys.argv[1:] or list(refconfigs.CONFIGS)
    for n in names:
        print("building", n, build(n), flush=True)
        print("running", n, run_harness(n), flush=True)
