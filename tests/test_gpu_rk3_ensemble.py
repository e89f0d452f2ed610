"""Batched ensemble (persistent launches over all boxes, HIP-graph replay,
autotuned tiling) against the single-box path.

Every box of an ensemble must advance bit for bit as the single-box
library advances it alone (FP64: which is itself bit-identical to the
reference; mixed: the same mixed kernels), with its own DT and clock,
whatever the tiling, with or without graph replay, and after an autotune.
"""

import importlib

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

# uneven boxes: partial 64-wide tiles, odd heights, tiny and several chunks
SPECS = [
    dict(npx=40, npy=32, npz=33, ra=2e4, ta=1e3, contrast=5.0, seed=101),
    dict(npx=96, npy=56, npz=40, ra=3e5, ta=1e5, contrast=20.0, seed=102),
    dict(npx=72, npy=88, npz=64, ra=1e6, ta=1e6, contrast=30.0, seed=103),
    dict(npx=32, npy=96, npz=47, ra=1e4, ta=1e2, contrast=3.0, seed=104),
    dict(npx=64, npy=40, npz=32, ra=6e4, ta=4e4, contrast=10.0, seed=105),
]


def say(*a):
    print(*a, flush=True)


def boxes(specs=SPECS):
    wl = importlib.import_module("rast-3dmhd_amd.workload")
    return [wl.ensemble_box(s) for s in specs]


def alone(meta, st, precision, n):
    with H.pkg().Box(meta, precision=precision) as box:
        box.upload(st)
        dt, um = box.steps(n)
        return box.download(), dt, um, box.clock()


def check_equal(ens, refs):
    for b, (st_ref, dt, um, clk) in enumerate(refs):
        got = ens.download(b)
        assert ens.clock(b) == clk, (b, ens.clock(b), clk)
        for k in H.FIELDS:
            assert np.array_equal(got[k], st_ref[k]), (b, k, np.abs(got[k] - st_ref[k]).max())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ensemble_bitwise_equal_single_box(precision):
    say("case ensemble vs single", precision)
    bx = boxes()
    nsteps = 7  # one direct step, pair replays, a single-step replay
    refs = [alone(m, s, precision, nsteps) for m, s in bx]
    pkg = H.pkg()
    with pkg.Ensemble([m for m, _ in bx], precision=precision) as ens:
        for b, (_, s) in enumerate(bx):
            ens.upload(b, s)
        dt, um = ens.steps(nsteps)
        for b, r in enumerate(refs):
            assert dt[b] == r[1] and um[b] == r[2], (b, dt[b], r[1])
        check_equal(ens, refs)


# (tile_x, tile_y, kdepth, grid); the 64-wide shapes through the 3-value call
TILINGS = [(64, 4, 8, 0), (64, 8, 0, 256), (64, 4, 16, 512), (64, 8, 32, 1024), (32, 8, 16, 0), (32, 8, 0, 512)]
# odd sizes: partial tiles in x and y for every tile shape
ODD = dict(npx=45, npy=37, npz=35, ra=5e4, ta=1e3, contrast=8.0, seed=106)


@pytest.mark.parametrize("tiling", TILINGS)
def test_ensemble_tiling_invariance(tiling):
    say("case ensemble tiling", tiling)
    bx = boxes(SPECS[:3] + [ODD])
    refs = [alone(m, s, "mixed", 4) for m, s in bx]
    with H.pkg().Ensemble([m for m, _ in bx], precision="mixed") as ens:
        if tiling[0] == 64:
            ens.set_tiling(*tiling[1:])
            assert ens.tiling() == tiling[1:]
        else:
            ens.set_tile(*tiling)
        assert ens.tile() == tiling
        for b, (_, s) in enumerate(bx):
            ens.upload(b, s)
        ens.steps(2, graph=False)
        ens.steps(2, graph=True)
        check_equal(ens, refs)


def test_ensemble_autotune_keeps_state():
    say("case ensemble autotune")
    bx = boxes()
    refs = [alone(m, s, "fp64", 5) for m, s in bx]
    with H.pkg().Ensemble([m for m, _ in bx], precision="fp64") as ens:
        for b, (_, s) in enumerate(bx):
            ens.upload(b, s)
        ens.steps(2)
        report = ens.autotune(nsteps=2)
        say("autotune", ens.tiling(), min(r["ms_step"] for r in report))
        assert len(report) == 56  # 2 shapes x 7 depths x 4 grids
        best = min(report, key=lambda r: r["ms_step"])
        assert ens.tiling() == (best["tile_y"], best["kdepth"], best["grid"])
        ens.steps(3)
        check_equal(ens, refs)


def test_ensemble_autotune_mixed_tile_shapes():
    """The mixed form also tries 32x8 tiles: 3 shapes x 7 depths x 4 grids."""
    say("case ensemble autotune mixed")
    bx = boxes(SPECS[:2] + [ODD])
    refs = [alone(m, s, "mixed", 3) for m, s in bx]
    with H.pkg().Ensemble([m for m, _ in bx], precision="mixed") as ens:
        for b, (_, s) in enumerate(bx):
            ens.upload(b, s)
        ens.steps(1)
        report = ens.autotune(nsteps=1)
        assert len(report) == 84
        assert sum(r["tile_x"] == 32 for r in report) == 28
        best = min(report, key=lambda r: r["ms_step"])
        say("autotune", ens.tile(), best["ms_step"])
        assert ens.tile() == (best["tile_x"], best["tile_y"], best["kdepth"], best["grid"])
        ens.steps(2)
        check_equal(ens, refs)
        with pytest.raises(H.pkg().MhdError) as e:
            ens.set_tile(32, 4, 8, 0)
        assert e.value.code == H.pkg().MHD_ERR_ARG


def test_ensemble_reupload_one_box():
    """Uploading one box again restarts that box only."""
    say("case ensemble re-upload")
    bx = boxes(SPECS[:2])
    with H.pkg().Ensemble([m for m, _ in bx], precision="fp64") as ens:
        for b, (_, s) in enumerate(bx):
            ens.upload(b, s)
        ens.steps(3)
        ens.upload(1, bx[1][1])
        ens.set_clock(1, 0, 0.0, 0.0)
        ens.steps(2)
        r0 = alone(*bx[0], "fp64", 5)
        r1 = alone(*bx[1], "fp64", 2)
        check_equal(ens, [r0, r1])


def test_ensemble_errors():
    say("case ensemble errors")
    pkg = H.pkg()
    bx = boxes(SPECS[:2])
    with pkg.Ensemble([m for m, _ in bx], precision="fp64") as ens:
        with pytest.raises(pkg.MhdError) as e:
            ens.steps(1)
        assert e.value.code == pkg.MHD_ERR_STATE
        with pytest.raises(pkg.MhdError) as e:
            ens.set_tiling(5, 8, 0)
        assert e.value.code == pkg.MHD_ERR_ARG
        with pytest.raises(pkg.MhdError) as e:
            ens.set_tile(32, 8, 8, 0)  # 32-wide tiles are built for the mixed form only
        assert e.value.code == pkg.MHD_ERR_ARG
        # the C entry point checks the box index itself
        assert ens.lib.mhd_ens_download(ens.h, 7, *([None] * 8)) == pkg.MHD_ERR_ARG
        assert ens.lib.mhd_ens_download(ens.h, -1, *([None] * 8)) == pkg.MHD_ERR_ARG
    m0 = dict(bx[0][0])
    m0["lmag"] = 0
    with pytest.raises(pkg.MhdError) as e:
        pkg.Ensemble([m0, bx[1][0]])
    assert e.value.code == pkg.MHD_ERR_ARG
