"""Pins the C oracle against the reference Fortran build.

The fixtures in tests/golden were written by the reference itself (its own
STATIC / FLUXES / STEP linked into oracle/refharness.f, run under MPI, see
tests/golden/make_golden.py).  Every comparison here is bit-for-bit.
"""

import numpy as np
import pytest

import helpers as H

FLUX = ["fu", "fv", "fw", "fr", "ft", "ww1", "ww2", "ww3"]


@pytest.fixture(scope="module")
def O():
    return H.oracle_mod()


@pytest.mark.parametrize("name", H.SMALL)
def test_fluxes_bitwise(O, name):
    g = H.load(name)
    o = O.Oracle(H.meta_of(g), H.initial_state(name))
    o.fluxes()
    names = FLUX if int(g["lmag"]) else FLUX[:5]
    for k in names:
        a = o.arrays[k][1:-1, 1:-1, 1:-1]
        b = g[f"flux_{k}"][1:-1, 1:-1, 1:-1]
        assert np.array_equal(a, b), f"{name} {k}: max |diff| {np.abs(a - b).max()}"


@pytest.mark.parametrize("name", H.ALL)
def test_steps_bitwise(O, name):
    g = H.load(name)
    init = H.initial_state(name)
    assert H.state_sha(init) == str(g["init_sha"])
    o = O.Oracle(H.meta_of(g), init)
    dumps = [int(x) for x in g["dumps"]]
    for n in range(1, max(dumps) + 1):
        o.step()
        if n in dumps:
            assert o.meta["dt"] == float(g[f"step{n}_dt"])
            assert o.meta["umach"] == float(g[f"step{n}_umach"])
            assert o.meta["timc"] == float(g[f"step{n}_timc"])
            assert o.meta["nit"] == int(g[f"step{n}_nit"])
            assert H.state_sha(o.state()) == str(g[f"step{n}_sha"]), f"{name} step {n}"
            if f"step{n}_ru" in g:
                for k in H.FIELDS:
                    assert np.array_equal(o.arrays[k], g[f"step{n}_{k}"]), f"{name} step {n} {k}"


def test_reference_decomposition_invariance():
    """The reference gives the same bits on 2x2 and 4x3 rank layouts, which
    is what makes a single resident domain equivalent to it."""
    a, b = H.load("mhdrot16"), H.load("mhdrot16_d43")
    assert str(a["init_sha"]) == str(b["init_sha"])
    for n in [int(x) for x in a["dumps"]]:
        assert str(a[f"step{n}_sha"]) == str(b[f"step{n}_sha"])


@pytest.mark.parametrize("name", ["hydro16", "mhdrot16", "mhdflux24"])
def test_initial_state_rebuild(name):
    """STATIC slicing + B layer + RAN2 noise + periodic fill, rebuilt from the
    1-D profiles, equals the reference's initial state bit for bit."""
    g = H.load(name)
    st = H.harness_initial_state(g, *H.harness_params(name))
    for k in H.FIELDS:
        assert np.array_equal(st[k], g[f"init_{k}"]), k


def test_ran2_known_values():
    """RAN2 is an integer LCG plus a 97-slot shuffle; a brute-force
    restatement of Numerical Recipes' published ran2-style generator with
    M=714025, IA=1366, IC=150889 gives the same first draws."""
    R = H.ics().Ran2
    r = R(-62659)
    got = [r() for _ in range(5)]
    # brute force, independent of the class
    m, ia, ic = 714025, 1366, 150889
    idum = (ic + 62659) % m
    tab = []
    for _ in range(97):
        idum = (ia * idum + ic) % m
        tab.append(idum)
    idum = (ia * idum + ic) % m
    iy = idum
    rm = float(np.float32(1.0) / np.float32(714025.0))
    exp = []
    for _ in range(5):
        j = (97 * iy) // m
        iy = tab[j]
        exp.append(iy * rm)
        idum = (ia * idum + ic) % m
        tab[j] = idum
    assert got == exp
    assert all(0.0 <= v < 1.0 for v in got)
