"""The reference's own, unchanged main program with the GPU STEP.

oracle/build_ref.build_dropin links the reference's 3dmhd.f, 3dmhdsub.f
(its STEP made a weak symbol) and the generated 3dmhdset.f with the STEP
shim (rast-3dmhd_amd/fortran/step_shim.f) and librast3dmhd.so, for the
decks of tests/golden/drv12 and drvh12 (NPEY = NPEZ = 2, the decomposition
the reference's set-up supports).  Run under MPI on 4 ranks, the shim
gathers the slabs onto rank 0, which steps on the GPU; DT and UMACH are
broadcast and the slabs scattered back on dump steps.  The files it writes
must be the reference CPU build's own files byte for byte: every rank's
rast3d.dat0.NNNN, rast3d.par and the rast3d.lis log (fresh run; restart of
the hydrodynamic deck).  The reference's restart of the MHD deck is NaN;
the GPU step reports it and stops.

The executables are built where the reference sources exist; the tests
skip with that reason elsewhere."""

import os
import shutil
import subprocess

import numpy as np
import pytest

import helpers as H
from test_driver_files import GOLD, NAMES, deck, dio, layout, same_files

REF = os.path.join(H.REPO, "oracle", "_ref")
MPIEXEC = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"


def exe_of(name):
    exe = os.path.join(REF, name, "3dmhd_gpu.exe")
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built (needs the reference sources: oracle/build_ref.build_dropins)")
    return exe


def run(name, cwd, ntotal, nstart=0, precision=None, timeout=120, rebase=None):
    env = dict(os.environ, RAST_NTOTAL=str(ntotal), RAST_NSTEP0="1", RAST_NSTART=str(nstart))
    if precision:
        env["RAST3DMHD_PRECISION"] = precision
    if rebase is not None:
        env["RAST3DMHD_REBASE"] = rebase
    _, c = deck(name)
    npe = int(c["params"]["NPEY"]) * int(c["params"]["NPEZ"])
    return subprocess.run([MPIEXEC, "-launcher", "fork", "-n", str(npe), exe_of(name)], cwd=cwd, env=env,
                          capture_output=True, text=True, timeout=timeout)


def test_dropin_driver_stops_every_rank_on_a_shim_error(tmp_path):
    """An invalid setting found on rank 0 stops all ranks together (no
    rank is left waiting in a collective): checked before any device call,
    so this runs without a GPU."""
    r = run("drv12", tmp_path, 3, precision="bogus", timeout=60)
    assert "RAST3DMHD_PRECISION is fp64 or mixed" in r.stdout, r.stdout + r.stderr
    assert not os.path.exists(tmp_path / "rast3d.dat0.0000")
    r = run("drv12", tmp_path, 3, precision="mixed", rebase="often", timeout=60)
    assert "RAST3DMHD_REBASE is a number of steps" in r.stdout, r.stdout + r.stderr
    assert not os.path.exists(tmp_path / "rast3d.dat0.0000")


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_reference_driver_with_gpu_step_writes_the_reference_files(tmp_path, name):
    print("case reference driver + GPU STEP, fresh run", name, flush=True)
    _, c = deck(name)
    r = run(name, tmp_path, int(c["nsteps"]))
    assert r.returncode == 0, r.stdout + r.stderr
    assert same_files(os.path.join(GOLD, name), tmp_path) == []


@pytest.mark.gpu
def test_reference_driver_with_gpu_step_restarts_from_the_reference_dump(tmp_path):
    """The hydrodynamic deck's restart from the reference's dump 1 writes the
    reference's own restart files, file for file, and continues as the
    uninterrupted run."""
    name = "drvh12"
    print("case reference driver + GPU STEP, restart", name, flush=True)
    _, c = deck(name)
    for f in os.listdir(os.path.join(GOLD, name)):
        if not f.endswith(".lis"):
            shutil.copy(os.path.join(GOLD, name, f), tmp_path)
    r = run(name, tmp_path, int(c["nsteps"]) - 1, nstart=1)
    assert r.returncode == 0, r.stdout + r.stderr
    lay = layout(c)
    for k in (1, 2):
        a = dio().read_dump(os.path.join(tmp_path, "rast3d"), k, lay, False)
        b = dio().read_dump(os.path.join(GOLD, name, "rast3d"), k + 1, lay, False)
        for f in a:
            assert np.array_equal(a[f], b[f]), (k, f)
    assert same_files(os.path.join(GOLD, name + "_restart"), tmp_path) == []


def test_reference_mhd_restart_is_not_finite():
    """The reference's own restart of the MHD deck (drv12_restart, written by
    its CPU build) is NaN: a restart reads interiors only, so the physical z
    ghost planes stay zero, 1/rho is infinite there and the boundary planes
    of BX, BY turn non-finite in the first step (BZ = 0 times an infinite
    velocity, 3dmhdsub.f:2228-2240); the reference carries on."""
    _, c = deck("drv12")
    a = dio().read_dump(os.path.join(GOLD, "drv12_restart", "rast3d"), 1, layout(c), True)
    assert all(np.isnan(a[f]).any() for f in ("bx", "by"))


@pytest.mark.gpu
def test_reference_driver_with_gpu_step_stops_the_nonfinite_mhd_restart(tmp_path):
    """The same restart with the GPU step: the non-finite CFL limit is
    reported (MHD_ERR_NONFINITE) and every rank stops, instead of writing
    NaN dumps."""
    print("case reference driver + GPU STEP, MHD restart", flush=True)
    _, c = deck("drv12")
    for f in os.listdir(os.path.join(GOLD, "drv12")):
        if not f.endswith(".lis"):
            shutil.copy(os.path.join(GOLD, "drv12", f), tmp_path)
    r = run("drv12", tmp_path, int(c["nsteps"]) - 1, nstart=1)
    assert "non-finite CFL limit or Mach number" in r.stdout, r.stdout + r.stderr
    a = dio().read_dump(os.path.join(tmp_path, "rast3d"), 1, layout(c), True)
    assert all(np.isfinite(a[f]).all() for f in a)  # dump 1 is the input: nothing was written


@pytest.mark.gpu
@pytest.mark.parametrize("rebase", [None, "1"])
def test_reference_driver_with_mixed_gpu_step(tmp_path, rebase):
    """The mixed form under the reference driver, with and without the
    background re-centred after every step (RAST3DMHD_REBASE)."""
    print("case reference driver + GPU STEP, mixed form, rebase", rebase, flush=True)
    _, c = deck("drv12")
    r = run("drv12", tmp_path, int(c["nsteps"]), precision="mixed", rebase=rebase)
    assert r.returncode == 0, r.stdout + r.stderr
    lay = layout(c)
    for k in (1, 2, 3):
        a = dio().read_dump(os.path.join(tmp_path, "rast3d"), k, lay, True)
        b = dio().read_dump(os.path.join(GOLD, "drv12", "rast3d"), k, lay, True)
        bscale = np.sqrt(sum(np.mean(b[f] ** 2) for f in ("bx", "by", "bz")))
        for f in a:
            scale = bscale if f in ("bx", "by", "bz") else np.sqrt(np.mean(b[f] ** 2))
            e = np.sqrt(np.mean((a[f] - b[f]) ** 2)) / scale
            assert e <= 1e-6, (k, f, e)
