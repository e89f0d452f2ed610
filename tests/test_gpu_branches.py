"""The optional FLUXES/BCON branches on the GPU (plain FP64 form).

Fixtures: the reference's own runs of one small box per branch
(tests/golden/<name>.npz: ITCON 2/3/4, the plume boundary steady and relaxing
in time, shear boundaries, density diffusion ID=1/2, Rempel conduction,
flux relaxation on 2 x 3 ranks).  As the harness that wrote them does, one
FLUXES evaluation on the initial state (mhd_fluxes), then three steps: the
right-hand side, DT, UMACH and the state (every field, ghosts included) equal
the reference's bit for bit.

Ragged boxes against the oracle: several branches at once on boxes whose
tiles and z chunks are partial, with NPEY/NPEZ decompositions of the
horizontal means and the relaxation integral, bit for bit over two steps."""

import numpy as np
import pytest

import helpers as H
from test_gpu_parity import random_box
from test_oracle_branches import BRANCHES

pytestmark = pytest.mark.gpu

FLUX = ["fu", "fv", "fw", "fr", "ft", "ww1", "ww2", "ww3"]


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("name", BRANCHES)
def test_branch_bitwise_vs_reference(name):
    say("case branch fixture", name)
    g = H.load(name)
    init = H.initial_state(name)
    dumps = [int(x) for x in g["dumps"]]
    with H.pkg().Box(H.meta_of(g)) as box:
        box.upload(init)
        f = box.fluxes()
        for k in (FLUX if int(g["lmag"]) else FLUX[:5]):
            a, b = f[k][1:-1, 1:-1, 1:-1], g[f"flux_{k}"][1:-1, 1:-1, 1:-1]
            assert np.array_equal(a, b), f"{name} {k}: max |diff| {np.abs(a - b).max()}"
        for n in range(1, max(dumps) + 1):
            dt, um = box.step()
            if n in dumps:
                assert dt == float(g[f"step{n}_dt"]) and um == float(g[f"step{n}_umach"]), (name, n, dt, um)
                st = box.download()
                if f"step{n}_ru" in g:
                    for k in H.FIELDS:
                        a, b = st[k], g[f"step{n}_{k}"]
                        assert np.array_equal(a, b), f"{name} step {n} {k}: max |diff| {np.nanmax(np.abs(a - b))}"
                assert H.state_sha(st) == str(g[f"step{n}_sha"]), (name, n)


BRANCH_SETS = {
    # density diffusion (1/ROM) + time-dependent heat loss, MHD, ragged tiles, NPEY=3
    "id1_loss": dict(shape=(70, 36, 20), lmag=True, over=dict(id=1, dh=0.02, dv=0.03, itc=2, tp=0.3, tc=0.05,
                                                             qfh=0.02, hh=0.6, xp=1.5, yp=1.2, zp=0.4, npey=3)),
    # boundary-confined diffusion + Rempel conduction, fixed-flux top, NPEY=4
    "id2_rem": dict(shape=(40, 24, 18), lmag=False, over=dict(id=2, dh=0.01, dv=0.02, lrem=1, itc=1, tu=0.9,
                                                             dztu=1e-3, dztb=-2e-3, zmax=1.0, npey=4)),
    # flux relaxation over 3 ranks + the relaxing plume top, several z chunks
    "rlax_plume": dict(shape=(64, 16, 72), lmag=True, over=dict(rlax=0.5, tstart=-1.0, toff=0.05, itc=0, izc=1,
                                                                tp=0.6, tc=0.03, hh=0.8, xp=2.0, yp=1.0,
                                                                npey=2, npez=3)),
    # shear boundaries + the half-domain cooler
    "shear_itc4": dict(shape=(34, 20, 12), lmag=True, over=dict(lshr=1, itc=4, tp=0.2)),
    # prognostic top temperature
    "itc3": dict(shape=(70, 37, 20), lmag=False, over=dict(itc=3)),
}


@pytest.mark.parametrize("case", list(BRANCH_SETS))
def test_branches_bitwise_vs_oracle(case):
    c = BRANCH_SETS[case]
    say("case branches vs oracle", case, c["shape"])
    meta, st = random_box(*c["shape"], lmag=c["lmag"], lrot=True, seed=sum(c["shape"]))
    meta.update(c["over"])
    O = H.oracle_mod()
    o = O.Oracle(meta, st)
    o.fluxes()  # as the harness: one evaluation before stepping
    for k in H.FIELDS:
        o.arrays[k][...] = st[k]
    with H.pkg().Box(meta) as box:
        box.upload(st)
        f = box.fluxes()
        for k in (FLUX if c["lmag"] else FLUX[:5]):
            assert np.array_equal(f[k][1:-1, 1:-1, 1:-1], o.arrays[k][1:-1, 1:-1, 1:-1]), (case, k)
        for n in range(2):
            dt, um = box.step()
            o.step()
            assert dt == o.meta["dt"] and um == o.meta["umach"], (case, n, dt, o.meta["dt"])
        got = box.download()
    for k in (H.FIELDS if c["lmag"] else H.FIELDS[:5]):
        assert np.array_equal(got[k], o.arrays[k]), (case, k, np.nanmax(np.abs(got[k] - o.arrays[k])))
