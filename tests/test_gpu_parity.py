"""Parity of the HIP step (through the C-ABI) with the reference and the oracle.

Bit-for-bit: the reference's golden fixtures (FLUXES and whole steps), the
oracle on box shapes that do not fill the 64x4 tiles and that split in z,
and at the 512x512x256 size a periodic-shift equivariance check.
"""

import importlib
import sys

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

FLUX = ["fu", "fv", "fw", "fr", "ft", "ww1", "ww2", "ww3"]


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("name", H.SMALL)
def test_fluxes_bitwise_vs_reference(name):
    say("case fluxes", name)
    g = H.load(name)
    with H.pkg().Box(H.meta_of(g)) as box:
        box.upload(H.initial_state(name))
        f = box.fluxes()
    names = FLUX if int(g["lmag"]) else FLUX[:5]
    for k in names:
        a, b = f[k][1:-1, 1:-1, 1:-1], g[f"flux_{k}"][1:-1, 1:-1, 1:-1]
        assert np.array_equal(a, b), f"{name} {k}: max |diff| {np.abs(a - b).max()}"


@pytest.mark.parametrize("name", H.ALL)
def test_steps_bitwise_vs_reference(name):
    say("case steps", name)
    g = H.load(name)
    dumps = [int(x) for x in g["dumps"]]
    with H.pkg().Box(H.meta_of(g)) as box:
        box.upload(H.initial_state(name))
        for n in range(1, max(dumps) + 1):
            dt, um = box.step()
            if n in dumps:
                assert dt == float(g[f"step{n}_dt"]), (n, dt, float(g[f"step{n}_dt"]))
                assert um == float(g[f"step{n}_umach"])
                nit, timc, timt = box.clock()
                assert nit == n and timc == float(g[f"step{n}_timc"]) and timt == float(g[f"step{n}_timt"])
                st = box.download()
                if f"step{n}_ru" in g:
                    for k in H.FIELDS:
                        a, b = st[k], g[f"step{n}_{k}"]
                        assert np.array_equal(a, b), f"{name} step {n} {k}: {np.abs(a - b).max()}"
                assert H.state_sha(st) == str(g[f"step{n}_sha"]), f"{name} step {n}"


def random_box(npx, npy, npz, lmag, lrot, seed, ngrid=1, tilt=45.0):
    cfg = importlib.import_module("rast-3dmhd_amd.config")
    par = {1: "1.5E02", 2: "1.0E00", 3: "3.0E00", 4: "6.0E00", 5: "1.5E00", 6: np.float32(np.radians(tilt)),
           7: "1.2E02", 8: "4.0E01", 17: "4.0E00", 18: "3.0E00", 19: "1.0E00"}
    mt = cfg.metrics(npx, npy, npz, xmax=4.0, ymax=3.0, zmax=1.0, ngrid=ngrid)
    t, r = cfg.polytrope(mt["zee"], 3.0, 6.0)
    meta = cfg.constants(npx, npy, npz, par, lmag=lmag, lrot=lrot, izcon=1, itcon=0, tb=float(t[-1]))
    meta.update(mt)
    rng = np.random.default_rng(seed)
    shape = (npx + 2, npy + 2, npz + 2)
    st = {k: np.zeros(shape, order="F") for k in H.FIELDS}
    st["tt"][:, :, :] = t[None, None, :] * (1.0 + 1e-3 * rng.standard_normal(shape))
    st["ro"][:, :, :] = r[None, None, :] * (1.0 + 1e-3 * rng.standard_normal(shape))
    for k in ("ru", "rv", "rw"):
        st[k][:, :, 1:-1] = st["ro"][:, :, 1:-1] * 2e-2 * rng.standard_normal((npx + 2, npy + 2, npz))
    st["rw"][:, :, 1] = 0.0
    st["rw"][:, :, -2] = 0.0
    if lmag:
        for k in ("bx", "by", "bz"):
            st[k][:, :, 1:-1] = 3e-2 * rng.standard_normal((npx + 2, npy + 2, npz))
        st["bz"][:, :, 1] = 0.0
        st["bz"][:, :, -2] = 0.0
    m = H.ics()
    for k in H.FIELDS:
        m.periodic_fill(st[k])
    return meta, st


@pytest.mark.parametrize("shape,lmag,lrot", [
    ((70, 37, 20), True, True),     # partial tiles in x and y
    ((130, 9, 11), False, False),   # hydro, three x tiles, tiny y
    ((64, 16, 72), True, True),     # several z chunks per column
    ((16, 8, 5), True, False),      # one interior plane between the boundary planes' stencils
])
def test_steps_bitwise_vs_oracle(shape, lmag, lrot):
    say("case oracle", shape, lmag, lrot)
    O = H.oracle_mod()
    meta, st = random_box(*shape, lmag=lmag, lrot=lrot, seed=sum(shape))
    o = O.Oracle(meta, st)
    with H.pkg().Box(meta) as box:
        box.upload(st)
        f = box.fluxes()
        o.fluxes()
        names = FLUX if lmag else FLUX[:5]
        for k in names:
            a, b = f[k][1:-1, 1:-1, 1:-1], o.arrays[k][1:-1, 1:-1, 1:-1]
            assert np.array_equal(a, b), f"fluxes {k}: {np.abs(a - b).max()}"
        o = O.Oracle(meta, st)
        for n in range(3):
            dt, um = box.step()
            o.step()
            assert dt == o.meta["dt"] and um == o.meta["umach"], n
        got = box.download()
    for k in (H.FIELDS if lmag else H.FIELDS[:5]):
        assert np.array_equal(got[k], o.arrays[k]), f"{k}: {np.abs(got[k] - o.arrays[k]).max()}"


def test_steps_batched_equal_single():
    """mhd_steps(n) (no host sync between steps) equals n calls of mhd_step."""
    say("case batched steps")
    meta, st = random_box(40, 24, 16, True, True, seed=3)
    with H.pkg().Box(meta) as a, H.pkg().Box(meta) as b:
        a.upload(st)
        b.upload(st)
        for _ in range(4):
            a.step()
        b.steps(4)
        sa, sb = a.download(), b.download()
        assert a.clock() == b.clock()
    for k in H.FIELDS:
        assert np.array_equal(sa[k], sb[k]), k


def test_full_size_shift_equivariance():
    """At 512x512x256 on a uniform horizontal grid, stepping a periodically
    shifted state gives the shifted result bit for bit (every tile, chunk and
    XCD mapping sees different data)."""
    say("case full-size shift 512x512x256")
    npx, npy, npz = 512, 512, 256
    meta, st = random_box(npx, npy, npz, True, True, seed=7, ngrid=0)
    si, sj = 37, 101

    def shift(a):
        inner = a[1:-1, 1:-1, :]
        out = np.zeros_like(a)
        out[1:-1, 1:-1, :] = np.roll(np.roll(inner, si, axis=0), sj, axis=1)
        H.ics().periodic_fill(out)
        return out

    st2 = {k: shift(v) for k, v in st.items()}
    with H.pkg().Box(meta) as box:
        box.upload(st)
        d1 = box.step()
        r1 = box.download()
        box.upload(st2)
        box.set_clock(0, 0.0, 0.0)
        d2 = box.step()
        r2 = box.download()
    assert d1 == d2
    for k in H.FIELDS:
        assert np.array_equal(shift(r1[k]), r2[k]), k
