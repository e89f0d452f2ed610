"""The reference's initial state built from scratch on the host.

STATIC's column is integrated by the restated Bulirsch-Stoer path
(rast-3dmhd_amd/static.py) and compared bit for bit with the columns the
reference's own STATIC wrote into the golden fixtures (arctan z grids with
ZZ2 = sqrt 2 and 1, constant and depth-dependent conductivity); the Bx
layer likewise; and the whole initial state (slicing, B layer, RAN2
temperature noise) built from those columns equals the reference's initial
state.  The closed-form polytrope that the integration approximates pins
the method independently of the fixtures."""

import importlib
import math
import os
import sys

import numpy as np
import pytest

import helpers as H

sys.path.insert(0, os.path.join(H.REPO, "oracle"))
import refconfigs  # noqa: E402

NAMES = ["hydro16", "mhdrot16", "mhdflux24", "mhdrot64"]


def S():
    return importlib.import_module("rast-3dmhd_amd.static")


def _lit(v, default="0.0E00"):
    cfg = importlib.import_module("rast-3dmhd_amd.config")
    return cfg.f32(float(str(v if v is not None else default).upper().replace("D", "E")))


def _inputs(name):
    c = refconfigs.CONFIGS[name]
    g = H.load(name)
    par, p = c["par"], c["params"]
    theta, grav = float(g["theta"]), float(g["grav"])
    polys = _lit(par.get(11))
    return g, dict(theta=theta, grav=grav, zmax=_lit(par.get(19), "1.0E00"), ngrid=1,
                   zz2=float(str(p.get("ZZ2", "1.4142135623730951D+00")).replace("D", "E")),
                   pzp=_lit(par.get(9)), sigma=_lit(par.get(10)), rkapst=(polys + 1.0) * theta / grav), par


@pytest.mark.parametrize("name", NAMES)
def test_static_profiles_bitwise_vs_reference(name):
    g, kw, _ = _inputs(name)
    t, r = S().profiles(int(g["npz"]), **kw)
    tp, rp = S().slice_profiles(t, r)
    assert np.array_equal(tp, g["prof_t"]), np.abs(tp - g["prof_t"]).max()
    assert np.array_equal(rp, g["prof_r"]), np.abs(rp - g["prof_r"]).max()


@pytest.mark.parametrize("name", ["mhdrot16", "mhdflux24", "mhdrot64"])
def test_bx_layer_bitwise_vs_reference(name):
    g, _, par = _inputs(name)
    col = S().bx_layer(g["zee"], _lit(par[21]), _lit(par[22]), _lit(par[23]))
    assert np.array_equal(col, g["prof_bx"])


@pytest.mark.parametrize("name", ["hydro16", "mhdrot16", "mhdflux24", "mhdrot64"])
def test_initial_state_from_scratch(name):
    """Integrated columns + slicing + B layer + RAN2 noise (+ the harness's
    velocity/field perturbation) = the reference's initial state."""
    g, kw, par = _inputs(name)
    t, r = S().profiles(int(g["npz"]), **kw)
    tp, rp = S().slice_profiles(t, r)
    bx = S().bx_layer(g["zee"], _lit(par.get(21)), _lit(par.get(22)), _lit(par.get(23))) if int(g["lmag"]) else None
    st = H.harness_initial_state(g, *H.harness_params(name), cols=(tp, rp, bx))
    assert H.state_sha(st) == str(g["init_sha"])


@pytest.mark.parametrize("npz,theta,zz2", [(32, 9.0, math.sqrt(2.0)), (64, 29.0, math.sqrt(2.0)), (20, 2.0, 1.0)])
def test_static_profiles_approach_the_closed_form(npz, theta, zz2):
    """With constant conductivity the ODE's solution is the polytrope
    T = 1 + THETA z, rho = T**(GRAV/THETA - 1); each accepted step keeps a
    relative error below EPS = 1e-12, so the column agrees to ~NPZ * 1e-12.
    A wrong substep sequence, a dropped extrapolation column or a transposed
    metric factor moves it by far more.  The integration advances the
    computational coordinate by the float32 step 1.0E00/FLOAT(NPZ-1), so
    plane k sits at s = k * step accumulated in double, not at the grid's
    k/(NPZ-1): near the bottom the two differ by ~1e-8, and the closed form
    is evaluated where the integration is."""
    grav = 2.0 * theta
    step = float(np.float32(1.0) / np.float32(npz - 1))
    s = [0.0]
    for _ in range(npz - 1):
        s.append(s[-1] + step)
    for ngrid in (1, 0):
        zmap = S().ZMap(ngrid, 1.0, 0.0, zz2)
        z = np.array([zmap(v)[0] for v in s])
        t, r = S().profiles(npz, theta=theta, grav=grav, zmax=1.0, ngrid=ngrid, zz2=zz2)
        tc = 1.0 + theta * z
        rc = tc ** (grav / theta - 1.0)
        assert np.abs(t / tc - 1.0).max() <= npz * 2e-12, (ngrid, np.abs(t / tc - 1.0).max())
        assert np.abs(r / rc - 1.0).max() <= npz * 1e-11, (ngrid, np.abs(r / rc - 1.0).max())
