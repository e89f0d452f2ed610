"""Pins the C oracle's optional FLUXES/BCON branches against the reference.

Each fixture (tests/golden/<name>.npz, written by the reference's own build
through oracle/refharness.f) switches on one branch: the prognostic top
temperature (ITCON=3), the half-domain cooler (ITCON=4), the plume boundary
(TP with ITCON=0 and, relaxing in time, ITCON=1), the embedded heat loss
(ITCON=2, with and without its tanh ramp), shear boundaries (LSHR), density
diffusion (ID=1, ID=2), Rempel conduction (LREM) and flux relaxation (RLAX).
The harness evaluates FLUXES once on the initial state, restores the state
and steps; so does the oracle here (with LSHR the temperature right-hand side
is never reset, so that first evaluation is part of the run).  Every
comparison is bit for bit."""

import numpy as np
import pytest

import helpers as H

BRANCHES = ["itc3", "itc4", "plume0", "plume1", "loss2", "loss2s", "shear", "ddiff1", "ddiff2", "rem", "rlax"]
FLUX = ["fu", "fv", "fw", "fr", "ft", "ww1", "ww2", "ww3"]


@pytest.fixture(scope="module")
def O():
    return H.oracle_mod()


def run_harness_sequence(O, name):
    g = H.load(name)
    init = H.initial_state(name)
    assert H.state_sha(init) == str(g["init_sha"])
    o = O.Oracle(H.meta_of(g), init)
    o.fluxes()
    flux = {k: o.arrays[k].copy() for k in FLUX}
    for k in H.FIELDS:
        o.arrays[k][...] = init[k]
    return g, o, flux


@pytest.mark.parametrize("name", BRANCHES)
def test_branch_fluxes_bitwise(O, name):
    g, _, flux = run_harness_sequence(O, name)
    names = FLUX if int(g["lmag"]) else FLUX[:5]
    for k in names:
        a = flux[k][1:-1, 1:-1, 1:-1]
        b = g[f"flux_{k}"][1:-1, 1:-1, 1:-1]
        assert np.array_equal(a, b), f"{name} {k}: max |diff| {np.abs(a - b).max()}"


@pytest.mark.parametrize("name", BRANCHES)
def test_branch_steps_bitwise(O, name):
    g, o, _ = run_harness_sequence(O, name)
    dumps = [int(x) for x in g["dumps"]]
    for n in range(1, max(dumps) + 1):
        o.step()
        if n in dumps:
            assert o.meta["dt"] == float(g[f"step{n}_dt"]), (name, n)
            assert o.meta["umach"] == float(g[f"step{n}_umach"]), (name, n)
            assert o.meta["timc"] == float(g[f"step{n}_timc"])
            if f"step{n}_ru" in g:
                for k in H.FIELDS:
                    a, b = o.arrays[k], g[f"step{n}_{k}"]
                    assert np.array_equal(a, b), f"{name} step {n} {k}: max |diff| {np.nanmax(np.abs(a - b))}"
            assert H.state_sha(o.state()) == str(g[f"step{n}_sha"]), f"{name} step {n}"


def test_branch_fixtures_switch_their_branch():
    """Each fixture has its branch on (and the parameters it needs)."""
    want = {"itc3": ("itc", 3), "itc4": ("itc", 4), "loss2": ("itc", 2), "loss2s": ("itc", 2),
            "shear": ("lshr", 1), "ddiff1": ("id", 1), "ddiff2": ("id", 2), "rem": ("lrem", 1)}
    for name, (k, v) in want.items():
        assert int(H.load(name)[k]) == v, name
    assert float(H.load("plume0")["tp"]) != 0 and float(H.load("plume0")["tc"]) == 0
    assert float(H.load("plume1")["tc"]) != 0 and int(H.load("plume1")["itc"]) == 1
    assert float(H.load("loss2")["tc"]) != 0 and float(H.load("loss2s")["tc"]) == 0
    r = H.load("rlax")
    assert float(r["rlax"]) > 0 and int(r["npez"]) == 3
