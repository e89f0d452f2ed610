"""CPU-side checks: the C-ABI library loads and exports every declared entry
point (no GPU calls), and the host set-up reproduces the reference driver's
constants and metric arrays bit for bit."""

import ctypes
import importlib
import os
import re

import numpy as np
import pytest

import helpers as H

HEADER = os.path.join(H.REPO, "include", "rast3dmhd.h")
HEADERS = sorted(os.path.join(H.REPO, "include", f) for f in os.listdir(os.path.join(H.REPO, "include"))
                 if f.endswith(".h"))


def declared_symbols():
    names = set()
    for h in HEADERS:
        txt = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(mhd_[a-z_0-9]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    pkg = H.pkg()
    assert os.path.exists(pkg.LIB_PATH), "build the HIP library first (__graft_entry__.build())"
    lib = ctypes.CDLL(pkg.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 28
    for n in names:
        assert hasattr(lib, n), n


def test_python_binding_declares_every_symbol():
    """load() sets argtypes for every entry point of include/*.h."""
    src = open(os.path.join(H.REPO, "rast-3dmhd_amd", "__init__.py")).read()
    for n in declared_symbols():
        assert f"lib.{n}.argtypes" in src, n


def test_desc_layout_matches_header():
    """mhd_desc in Python has the header's members in the header's order."""
    txt = open(HEADER).read()
    body = txt[txt.index("typedef struct mhd_desc"):txt.index("} mhd_desc;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    members = []
    for typ, names in re.findall(r"\b(int|double)\s+([^;]+);", body):
        members += [(typ, n.strip()) for n in names.split(",")]
    py = [("int" if t is ctypes.c_int else "double", n) for n, t in H.pkg().MhdDesc._fields_]
    assert py == members


def test_missing_library_raises(tmp_path):
    pkg = importlib.import_module("rast-3dmhd_amd")
    saved = pkg._lib
    pkg._lib = None
    try:
        with pytest.raises(RuntimeError):
            pkg.load(str(tmp_path / "nope.so"))
    finally:
        pkg._lib = saved


@pytest.mark.parametrize("name", H.ALL)
def test_constants_and_metrics_match_reference(name):
    import sys
    cfg = importlib.import_module("rast-3dmhd_amd.config")
    sys.path.insert(0, os.path.join(H.REPO, "oracle"))
    try:
        import refconfigs
    finally:
        sys.path.pop(0)
    g = H.load(name)
    c = refconfigs.CONFIGS[name]
    p = c["params"]
    par = dict(c["par"])
    par.setdefault(17, "4.0E00")
    par.setdefault(18, "4.0E00")
    par.setdefault(19, "1.0E00")
    m = cfg.constants(int(g["npx"]), int(g["npy"]), int(g["npz"]), par,
                      lmag=p.get("LMAG") == ".TRUE.", lrot=p.get("LROT") == ".TRUE.",
                      izcon=int(g["izc"]), itcon=int(g["itc"]), tb=float(g["tb"]))
    for k, v in m.items():
        if k in g and np.ndim(g[k]) == 0:
            assert float(v) == float(g[k]), k
    zz = tuple(float(s.replace("D", "e")) for s in (p.get("ZZ1", "0.0D+00"), p.get("ZZ2", "1.4142135623730951D+00")))
    mt = cfg.metrics(int(g["npx"]), int(g["npy"]), int(g["npz"]), xmax=cfg.f32(float(par[17])),
                     ymax=cfg.f32(float(par[18])), zmax=cfg.f32(float(par[19])), zz=zz)
    for k in ("dxxdx", "d2xxdx2", "ddx", "dyydy", "d2yydy2", "ddy", "dzzdz", "d2zzdz2", "ddz"):
        assert np.array_equal(mt[k], g[k]), k


def test_float32_folded_constants():
    """The survey-confirmed amdflang values of the float32-literal constants."""
    cfg = importlib.import_module("rast-3dmhd_amd.config")
    m = cfg.constants(64, 64, 32, {1: "1.0E02", 7: "1.0E02", 8: "1.0E00", 5: "1.0E00"}, lmag=True, lrot=True)
    assert m["gam1"] == 5.33333361148834229e-01
    assert m["c13"] == 3.33333343267440796e-01
    assert m["gamma"] == 1.66666662693023682e+00
    assert m["dzf"] == 1.0 / 31.0 or abs(m["dzf"] - 3.22580635547637939e-02) == 0.0
    assert m["sf"] == 3.14999997615814209e-01


def test_physics_mapping_round_trip():
    cfg = importlib.import_module("rast-3dmhd_amd.config")
    ph = cfg.physics(contrast=10.0, ra=1e5, pr=1.0, pm=1.0, ta=1e4, tilt_deg=0.0)
    theta, grav, re, prp, ry, rm = ph[3], ph[4], ph[1], ph[2], ph[5], ph[7]
    assert theta == pytest.approx(9.0) and grav == pytest.approx(18.0)
    tmid = 1.0 + theta / 2
    nu = 1.0 / (re * tmid)
    cp = 2.5
    kappa = 1.0 / (re * prp * tmid * cp)
    assert nu / kappa == pytest.approx(1.0)
    assert grav * (theta - grav / cp) / (nu * kappa * tmid) == pytest.approx(1e5)
    assert (1.0 / ry / nu) ** 2 == pytest.approx(1e4)
    assert nu / (1.0 / rm) == pytest.approx(1.0)
