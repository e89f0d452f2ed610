"""The host driver with the GPU step writes the reference's files.

The same decks and fixtures as test_driver_files.py (the reference's own
main program on 2 x 2 MPI ranks), now with the HIP library's FP64 step:
every .dat0 rank file, the .par records and the .lis log of a fresh run are
the reference's bytes, and a restart from the reference's dump 1 continues
as its uninterrupted run (for the hydro deck: as its own restart, file for
file).  The mixed form runs the same driver and its dumps stay within 1e-6
of the reference's."""

import os

import numpy as np
import pytest

import helpers as H
from test_driver_files import GOLD, NAMES, deck, dio, layout, run_fresh, run_restart, same_files

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", NAMES)
def test_gpu_fresh_run_writes_the_reference_files(tmp_path, name):
    print("case driver fresh run", name, flush=True)
    r = run_fresh(name, tmp_path)
    assert r["dumps"] == 3 and not r["shutdown"]
    assert same_files(os.path.join(GOLD, name), tmp_path) == []


@pytest.mark.parametrize("name", NAMES)
def test_gpu_restart_from_the_reference_dump(tmp_path, name):
    print("case driver restart", name, flush=True)
    _, c = deck(name)
    lay, lmag = layout(c), name == "drv12"
    run_restart(name, tmp_path)
    for k in (1, 2):
        a = dio().read_dump(os.path.join(tmp_path, "rast3d"), k, lay, lmag)
        b = dio().read_dump(os.path.join(GOLD, name, "rast3d"), k + 1, lay, lmag)
        for f in a:
            assert np.array_equal(a[f], b[f]), (k, f)
    if name == "drvh12":
        assert same_files(os.path.join(GOLD, name + "_restart"), tmp_path) == []


def test_gpu_mixed_form_driver_run(tmp_path):
    print("case driver mixed form", flush=True)
    _, c = deck("drv12")
    lay = layout(c)
    r = run_fresh("drv12", tmp_path, precision="mixed")
    assert r["dumps"] == 3
    for k in (1, 2, 3):
        a = dio().read_dump(os.path.join(tmp_path, "rast3d"), k, lay, True)
        b = dio().read_dump(os.path.join(GOLD, "drv12", "rast3d"), k, lay, True)
        bscale = np.sqrt(sum(np.mean(b[f] ** 2) for f in ("bx", "by", "bz")))
        for f in a:
            scale = bscale if f in ("bx", "by", "bz") else np.sqrt(np.mean(b[f] ** 2))
            e = np.sqrt(np.mean((a[f] - b[f]) ** 2)) / scale
            assert e <= 1e-6, (k, f, e)
