"""Error behaviour of the C-ABI.

CPU: mhd_create validates the descriptor before any HIP call, so the
reference's own STOP conditions and the branches that are not built come
back as error codes with a message on a machine without a GPU.
GPU: a state that makes the CFL limits or the Mach number non-finite is
reported as MHD_ERR_NONFINITE (the reference carries on silently), in both
forms and in the ensemble."""

import ctypes as C

import numpy as np
import pytest

import helpers as H
from test_gpu_parity import random_box


def _create(meta, **over):
    pkg = H.pkg()
    lib = pkg.load()
    d = pkg.make_desc(meta)
    for k, v in over.items():
        setattr(d, k, v)
    h = C.c_void_p()
    rc = lib.mhd_create(C.byref(d), C.byref(h))
    return rc, lib.mhd_last_error(None).decode(), h


@pytest.mark.parametrize("over,code,text", [
    (dict(izcon=2), "MHD_ERR_SCOPE", "Invalid IZCON"),
    (dict(itcon=5), "MHD_ERR_SCOPE", "ITCON=5"),
    (dict(id=3), "MHD_ERR_SCOPE", "ID=3"),
    (dict(rlax=0.5, itcon=1, izcon=0), "MHD_ERR_SCOPE", "RLAX > 0"),
    (dict(lshr=1, lrem=1), "MHD_ERR_SCOPE", "LSHR with LREM"),
    (dict(itcon=2, precision=1), "MHD_ERR_SCOPE", "MHD_FP64 form only"),
    (dict(tp=0.5, precision=1), "MHD_ERR_SCOPE", "MHD_FP64 form only"),
    (dict(npey=3), "MHD_ERR_ARG", "NPEY=3"),
    (dict(precision=7), "MHD_ERR_ARG", "precision 7"),
    (dict(npz=3), "MHD_ERR_ARG", "too small"),
])
def test_create_rejects_descriptors_before_touching_the_device(over, code, text):
    meta, _ = random_box(16, 8, 6, True, True, seed=1)
    rc, msg, h = _create(meta, **over)
    assert rc == getattr(H.pkg(), code), (rc, msg)
    assert text in msg and h.value is None


def test_null_arguments():
    lib = H.pkg().load()
    assert lib.mhd_create(None, None) == H.pkg().MHD_ERR_ARG
    assert lib.mhd_step(None, None, None) == H.pkg().MHD_ERR_ARG
    assert b"null" in lib.mhd_last_error(None)
    assert lib.mhd_rebase(None) == H.pkg().MHD_ERR_ARG
    assert lib.mhd_set_rebase_interval(None, 10) == H.pkg().MHD_ERR_ARG
    assert b"null" in lib.mhd_last_error(None)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nonfinite_state_is_reported(precision):
    print("case non-finite state", precision, flush=True)
    pkg = H.pkg()
    meta, st = random_box(70, 37, 20, True, True, seed=12)
    st["tt"][30, 20, 10] = np.nan
    with pkg.Box(meta, precision=precision) as box:
        box.upload(st)
        with pytest.raises(pkg.MhdError) as e:
            box.step()
        assert e.value.code == pkg.MHD_ERR_NONFINITE
        assert "non-finite" in str(e.value)
    # a finite state steps normally again on a fresh handle
    meta, st = random_box(70, 37, 20, True, True, seed=12)
    with pkg.Box(meta, precision=precision) as box:
        box.upload(st)
        dt, um = box.step()
        assert np.isfinite(dt) and np.isfinite(um)


@pytest.mark.gpu
def test_nonfinite_box_in_an_ensemble_is_reported():
    print("case non-finite ensemble box", flush=True)
    pkg = H.pkg()
    m0, s0 = random_box(40, 24, 12, True, True, seed=13)
    m1, s1 = random_box(32, 16, 10, True, True, seed=14)
    s1["ro"][5, 5, 5] = np.nan
    with pkg.Ensemble([m0, m1], precision="fp64") as ens:
        ens.upload(0, s0)
        ens.upload(1, s1)
        with pytest.raises(pkg.MhdError) as e:
            ens.steps(1, graph=False)
        assert e.value.code == pkg.MHD_ERR_NONFINITE


@pytest.mark.parametrize("key,delta", [("dxxdx", -1), ("ddy", 1), ("rkapa", -2), ("ddz", -1)])
def test_metric_lengths_checked_before_create(key, delta):
    """A metric array of the wrong length is refused in Python before any
    library call: the C side copies NX, NY or NZ doubles from each array."""
    pkg = H.pkg()
    meta, _ = random_box(16, 8, 6, True, True, seed=1)
    a = np.asarray(meta[key])
    meta[key] = np.resize(a, a.size + delta)
    desc = pkg.make_desc(meta)
    with pytest.raises(ValueError, match=key):
        pkg.metric_arrays(meta, desc)
    with pytest.raises(ValueError, match=key):
        pkg.Box(meta)  # raises before mhd_create, so no device is needed
    with pytest.raises(ValueError, match=key):
        pkg.Ensemble([meta])


def _noalloc_call(meta, **over):
    pkg = H.pkg()
    lib = pkg.load()
    d = pkg.make_desc(meta)
    for k, v in over.items():
        setattr(d, k, v)
    dm = pkg.MhdDevMetrics()
    ptrs = (C.c_void_p * 8)()
    nit, timc, timt = C.c_int(0), C.c_double(0.0), C.c_double(0.0)
    rc = lib.mhd_step_noalloc(C.byref(d), C.byref(dm), ptrs, C.byref(nit), C.byref(timc), C.byref(timt), 0.0,
                              None, None, None)
    return rc, lib.mhd_last_error(None).decode()


def test_noalloc_step_rejects_before_touching_the_device():
    """mhd_step_noalloc checks its arguments, the precision and the size of
    its static scratch before any HIP call."""
    pkg = H.pkg()
    meta, _ = random_box(128, 128, 64, True, True, seed=1)  # 3 registers x 8 fields: 207 MB > 64 MiB
    rc, msg = _noalloc_call(meta)
    assert rc == pkg.MHD_ERR_ARG and "scratch" in msg and "maximum" in msg, (rc, msg)
    meta, _ = random_box(16, 8, 6, True, True, seed=1)
    rc, msg = _noalloc_call(meta, precision=pkg.MHD_MIXED)
    assert rc == pkg.MHD_ERR_SCOPE and "MHD_FP64" in msg, (rc, msg)
    rc, msg = _noalloc_call(meta, itcon=3)
    assert rc == pkg.MHD_ERR_SCOPE and "main path only" in msg, (rc, msg)
    rc, msg = _noalloc_call(meta)  # fits, but the metric pointers are null
    assert rc == pkg.MHD_ERR_ARG and "metric array 0 is null" in msg, (rc, msg)
    lib = pkg.load()
    assert lib.mhd_step_noalloc(None, None, None, None, None, None, 0.0, None, None, None) == pkg.MHD_ERR_ARG
