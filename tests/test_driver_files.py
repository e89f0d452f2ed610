"""The host driver and the reference's file formats, on the CPU.

Fixtures tests/golden/drv12 and drvh12 are the files the reference's own
main program wrote (tests/golden/make_driver_golden.py): per-rank
rast3d.dat0.NNNN, rast3d.par and rast3d.lis of a fresh start with a dump
after every step, on 2 x 2 ranks (MHD + tilted rotation + Bx layer +
depth-dependent kappa; hydro with a fixed-flux top and fixed-T bottom), and
of a restart from dump 1.

Here the host driver (rast-3dmhd_amd/driver.py) runs the same decks with
the C oracle standing in for the step (the oracle is bit-identical to the
reference's STEP); the GPU step replaces it in test_gpu_driver.py.  Every
output file must equal the reference's byte for byte: this pins the set-up
(constants, metrics, conductivity profile, STATIC), the initial time-step
diagnostics, the .dat0/.par/.lis writers, and the restart reader.
"""

import importlib
import os
import shutil
import sys

import numpy as np
import pytest

import helpers as H

sys.path.insert(0, os.path.join(H.REPO, "oracle"))
import refconfigs  # noqa: E402

GOLD = os.path.join(H.REPO, "tests", "golden")
NAMES = ["drv12", "drvh12"]


def D():
    return importlib.import_module("rast-3dmhd_amd.driver")


def dio():
    return importlib.import_module("rast-3dmhd_amd.dumpio")


class OracleStepper:
    """The oracle behind the driver's stepper interface (tests only)."""

    def __init__(self, meta):
        self.meta = meta
        self.O = H.oracle_mod()

    def upload(self, st):
        self.o = self.O.Oracle(self.meta, st)

    def step(self):
        self.o.step()
        return self.o.meta["dt"], self.o.meta["umach"]

    def clock(self):
        return self.o.meta["nit"], self.o.meta["timc"], self.o.meta["timt"]

    def set_clock(self, nit, timc, timi):
        self.o.meta.update(nit=nit, timc=timc, timi=timi, timt=timi + timc)

    def download(self):
        return self.o.state()


def deck(name):
    c = refconfigs.DRIVER_CONFIGS[name]
    return D().Deck(c["params"], c["par"], c["ipar"]), c


def layout(c):
    p = c["params"]
    return dio().Layout(p["NPX"], p["NPY"], p["NPZ"], p["NPEY"], p["NPEZ"])


def same_files(ref_dir, out_dir):
    bad = []
    for f in sorted(os.listdir(ref_dir)):
        a = open(os.path.join(ref_dir, f), "rb").read()
        b = open(os.path.join(out_dir, f), "rb").read() if os.path.exists(os.path.join(out_dir, f)) else b""
        if a != b:
            bad.append(f)
    return bad


def run_fresh(name, out, stepper=None, precision="fp64"):
    dk, c = deck(name)
    return D().Run(dk, str(out), precision=precision, ntotal=c["nsteps"], nstep0=1, stepper=stepper).run()


def run_restart(name, out, stepper=None, precision="fp64"):
    """Restart from dump 1 of the reference's files, as the reference's own
    restart fixture was made: the .par and .dat0 files copied in first."""
    dk, c = deck(name)
    for f in os.listdir(os.path.join(GOLD, name)):
        if not f.endswith(".lis"):
            shutil.copy(os.path.join(GOLD, name, f), out)
    return D().Run(dk, str(out), precision=precision, ntotal=c["nsteps"] - 1, nstep0=1, nstart=1,
                   stepper=stepper).run()


@pytest.mark.parametrize("name", NAMES)
def test_fresh_run_writes_the_reference_files(tmp_path, name):
    r = run_fresh(name, tmp_path, OracleStepper)
    assert r["dumps"] == 3 and not r["shutdown"]
    assert same_files(os.path.join(GOLD, name), tmp_path) == []


@pytest.mark.parametrize("name", NAMES)
def test_restart_continues_as_the_uninterrupted_run(tmp_path, name):
    dk, c = deck(name)
    lay, lmag = layout(c), name == "drv12"
    run_restart(name, tmp_path, OracleStepper)
    for k in (1, 2):
        a = dio().read_dump(os.path.join(tmp_path, "rast3d"), k, lay, lmag)
        b = dio().read_dump(os.path.join(GOLD, name, "rast3d"), k + 1, lay, lmag)
        for f in a:
            assert np.array_equal(a[f], b[f]), (k, f)
    if name == "drvh12":
        # hydro: the reference's own restart gives the same files
        assert same_files(os.path.join(GOLD, name + "_restart"), tmp_path) == []


def test_reference_mhd_restart_goes_non_finite():
    """Why the driver fills the restart's z ghost planes: the reference's own
    MHD restart leaves them zero and its state turns NaN in the first step
    (BZ = 0 times an infinite velocity on the boundary planes)."""
    dk, c = deck("drv12")
    st = dio().read_dump(os.path.join(GOLD, "drv12_restart", "rast3d"), 1, layout(c), True)
    assert np.isnan(st["bx"]).any() and np.isnan(st["by"]).any()
    assert "NaN" in dio().read_lis(os.path.join(GOLD, "drv12_restart", "rast3d.lis"))[71]


def test_dump_files_round_trip(tmp_path):
    """Reading the reference's rank files and writing them back gives the
    same bytes; the .par layout record gives back the decomposition."""
    dk, c = deck("drv12")
    lay = layout(c)
    src = os.path.join(GOLD, "drv12", "rast3d")
    assert dio().count_dumps(src, lay, True) == 3
    for k in (1, 2, 3):
        dio().write_dump(os.path.join(tmp_path, "rast3d"), k, dio().read_dump(src, k, lay, True), lay, True)
    for r in range(lay.npe):
        assert open(dio().dat0_path(src, r), "rb").read() == open(dio().dat0_path(os.path.join(tmp_path, "rast3d"), r), "rb").read()
    p1 = dio().read_par(src + ".par", 1)
    assert dio().par_layout(p1) == lay
    assert p1[14] == 1.0 and dio().read_par(src + ".par", 3)[14] == 3.0


def test_dump_of_a_known_state_lands_in_the_right_rank_records(tmp_path):
    """A state whose value encodes its global (i, j, k) and field: every
    record of every rank file holds exactly that rank's interior block."""
    lay = dio().Layout(6, 4, 6, 2, 3)
    st = {}
    for n, f in enumerate(dio().DUMP_FIELDS):
        i, j, k = np.meshgrid(np.arange(8), np.arange(6), np.arange(8), indexing="ij")
        st[f] = np.asfortranarray(1000.0 * n + 100.0 * k + 10.0 * j + i)
    dio().write_dump(os.path.join(tmp_path, "x"), 2, st, lay, True)
    for r in range(lay.npe):
        raw = np.fromfile(dio().dat0_path(os.path.join(tmp_path, "x"), r), dtype=np.float64)
        recs = raw.reshape(-1, 6 * 2 * 2)
        assert recs.shape[0] == 16  # dump 2 of 8 records: records 9..16
        py, pz = r % 2, r // 2
        for n, f in enumerate(dio().DUMP_FIELDS):
            blk = recs[8 + n].reshape((6, 2, 2), order="F")
            i, j, k = np.meshgrid(np.arange(1, 7), 1 + py * 2 + np.arange(2), 1 + pz * 2 + np.arange(2), indexing="ij")
            assert np.array_equal(blk, 1000.0 * n + 100.0 * k + 10.0 * j + i), (r, f)


@pytest.mark.parametrize("x,w,d,text", [
    (150.0, 11, 4, " 1.5000D+02"), (0.7853981852531433, 11, 4, " 7.8540D-01"), (0.0, 11, 4, " 0.0000D+00"),
    (0.021486119, 10, 4, "2.1486D-02"), (-0.021486119, 10, 4, "**********"), (float("nan"), 10, 4, "       NaN"),
    (1.5e-120, 10, 4, "1.5000-120"), (-2.5, 11, 4, "-2.5000D+00")])
def test_fortran_d_edit_descriptor(x, w, d, text):
    assert dio().fmt_1pd(x, w, d) == text


def test_fortran_i_and_rank_strings():
    assert dio().fmt_i(47, 8) == "      47" and dio().fmt_i(12345, 3) == "***"
    assert [dio().pe_string(r) for r in (0, 7, 123, 4095)] == ["0000", "0007", "0123", "4095"]


def test_initial_dt_and_conductivity_match_the_log():
    """The header's time-step diagnostics come from the state the driver
    builds; the conductivity profile matches the reference's (mhdrot16 has
    PZP != 0)."""
    g = H.load("mhdrot16")
    c = refconfigs.CONFIGS["mhdrot16"]
    lit = D().literal
    theta, grav = float(g["theta"]), float(g["grav"])
    rkapst = (lit(c["par"][11]) + 1.0) * theta / grav
    rk, dk = D().kappa_profile(g["zee"], lit(c["par"][9]), lit(c["par"][10]), rkapst)
    assert np.array_equal(rk, g["rkapa"]) and np.array_equal(dk, g["dkapa"])
