"""mhd_step_noalloc: one FP64 step on the caller's device arrays with no
device allocation in the call, against the handle-based step (mhd_upload +
mhd_step + mhd_download) on the same state: bit for bit, DT, UMACH and the
clock included.  Calls in a row share the library's static scratch array:
a 64x64x32 MHD box, then a smaller 16x16x12 one, then a hydrodynamic box
(5 fields), then the first box's result stepped again (against two handle
steps)."""

import ctypes as C

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a):
    """A padded Fortran-order array as a flat device tensor in the same byte order."""
    return torch.from_numpy(np.ravel(np.asarray(a, dtype=np.float64), order="F").copy()).to("cuda")


def noalloc_step(meta, st, clock):
    pkg = H.pkg()
    lib = pkg.load()
    d = pkg.make_desc(meta)
    mt = [torch.from_numpy(a.copy()).to("cuda") for a in pkg.metric_arrays(meta, d)]
    dm = pkg.MhdDevMetrics(*[t.data_ptr() for t in mt])
    nf = 8 if d.lmag else 5
    shape = (d.npx + 2, d.npy + 2, d.npz + 2)
    fs = [_dev(st[k]) for k in pkg.FIELDS[:nf]]
    ptrs = (C.c_void_p * 8)(*([t.data_ptr() for t in fs] + [None] * (8 - nf)))
    nit, timc, timt = C.c_int(clock[0]), C.c_double(clock[1]), C.c_double(0.0)
    dt, um = C.c_double(), C.c_double()
    torch.cuda.synchronize()
    rc = lib.mhd_step_noalloc(C.byref(d), C.byref(dm), ptrs, C.byref(nit), C.byref(timc), C.byref(timt),
                              float(clock[2]), C.byref(dt), C.byref(um),
                              C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, lib.mhd_last_error(None).decode()
    out = {k: t.cpu().numpy().reshape(shape, order="F") for k, t in zip(pkg.FIELDS, fs)}
    return out, dt.value, um.value, (nit.value, timc.value, timt.value)


def handle_steps(meta, st, clock, n=1):
    with H.pkg().Box(meta) as box:
        box.set_clock(clock[0], clock[1], clock[2])
        box.upload(st)
        dt, um = box.steps(n)
        return box.download(), dt, um, box.clock()


def same(a, b, lmag):
    for k in H.FIELDS[:8 if lmag else 5]:
        assert np.array_equal(a[k], b[k]), (k, np.abs(a[k] - b[k]).max())


def test_noalloc_step_matches_the_handle_step():
    cases = [("mhdrot64", (7, 0.25, 1.5)), ("mhdrot16", (0, 0.0, 0.0)), ("hydro16", (3, 0.125, 2.0))]
    first = None
    for name, clock in cases:
        print("case mhd_step_noalloc vs handle step", name, clock, flush=True)
        g = H.load(name)
        meta, st = H.meta_of(g), H.initial_state(name)
        lmag = bool(int(meta["lmag"]))
        got, dt, um, clk = noalloc_step(meta, st, clock)
        ref, dtr, umr, clkr = handle_steps(meta, st, clock)
        assert dt == dtr and um == umr and clk == clkr, (name, dt, dtr, um, umr, clk, clkr)
        same(got, ref, lmag)
        if first is None:
            first = (meta, st, got, clk, clock[2])
    # the first box's result stepped again after the other calls = two handle steps
    meta, st0, got1, clk1, timi = first
    print("case mhd_step_noalloc second step of mhdrot64", flush=True)
    got2, dt2, um2, clk2 = noalloc_step(meta, got1, (clk1[0], clk1[1], timi))
    ref2, dtr2, umr2, clkr2 = handle_steps(meta, st0, cases[0][1], n=2)
    assert dt2 == dtr2 and um2 == umr2 and clk2 == clkr2
    same(got2, ref2, True)
