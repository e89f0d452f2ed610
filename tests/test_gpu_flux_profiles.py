"""Horizontal means and mean flux profiles (mhd_flux_profiles) against the
oracle's restatement of HORIZONTAL_MEAN / FCONM / FRADM, for both forms."""

import numpy as np
import pytest

import helpers as H
from test_gpu_parity import random_box

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _scales(st):
    """Per interior plane: the mean |term| of the convective-flux sum."""
    i = (slice(1, -1), slice(1, -1), slice(1, -1))
    ro, tt, ru, rv, rw = (st[k][i] for k in ("ro", "tt", "ru", "rv", "rw"))
    tp = tt - tt.mean(axis=(0, 1))
    u2 = (ru * ru + rv * rv + rw * rw) / (ro * ro)
    return np.abs(rw * (2.5 * tp + 0.5 * u2)).mean(axis=(0, 1))


def _compare(got, ref, st):
    sc = _scales(st)
    for k in ("ttm", "rom"):
        e = np.abs(got[k] - ref[k][1:-1]) / np.abs(ref[k][1:-1])
        assert e.max() <= 1e-13, (k, e.max())
    e = np.abs(got["fconm"] - ref["fconm"][1:-1])
    assert np.all(e <= 1e-12 * sc), (e / sc).max()
    e = np.abs(got["fradm"] - ref["fradm"][1:-1])
    assert e.max() <= 1e-11 * np.abs(ref["fradm"]).max(), e.max()


@pytest.mark.parametrize("shape,lmag,lrot", [((70, 37, 20), True, True), ((130, 9, 11), False, False)])
def test_flux_profiles_match_oracle_fp64(shape, lmag, lrot):
    say("case flux profiles fp64", shape)
    O = H.oracle_mod()
    meta, st = random_box(*shape, lmag=lmag, lrot=lrot, seed=sum(shape) + 1)
    o = O.Oracle(meta, st)
    with H.pkg().Box(meta) as box:
        box.upload(st)
        _compare(box.flux_profiles(), o.flux_profiles(), st)
        box.steps(3)
        o.step(3)
        st3 = o.state()
        a, b = box.flux_profiles(), box.flux_profiles()
        for k in a:
            assert np.array_equal(a[k], b[k]), k  # the same order on every call
        _compare(a, o.flux_profiles(), st3)


def test_flux_profiles_mixed_form():
    """The mixed form's profiles (from the stored values, residuals and
    background) against the oracle's on the state the box downloads."""
    say("case flux profiles mixed")
    O = H.oracle_mod()
    meta, st = random_box(70, 37, 20, True, True, seed=91)
    with H.pkg().Box(meta, precision="mixed") as box:
        box.upload(st)
        box.steps(3)
        got = box.flux_profiles()
        down = box.download()
    ref = O.Oracle(meta, down).flux_profiles()
    _compare(got, ref, down)
