"""ctypes front end of the C oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, the smoke entry point and bench.py's CPU-baseline leg use this
module, and only as the checker.  It owns a full set of the reference's /BIG/
arrays (26 padded Fortran-order float64 arrays) and calls the restated
FLUXES / BCON / COMMUNICATE / CFL / STEP of mhd_oracle.c on them.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

INT_KEYS = ["nx", "ny", "nz", "lmag", "lrot", "izc", "itc", "ibc", "nit"]
DBL_KEYS = ["hx", "h2x", "hy", "h2y", "hz", "h2z", "c13", "c23", "c43",
            "gam1", "gam2", "gam3", "zeta1", "zeta2",
            "cv", "ocv", "ore", "re", "repr", "theta", "grav", "sf", "gamma",
            "omx", "omz", "orm", "rm", "obeta", "tb", "tp", "dzf",
            "dt", "timt", "timc", "timi", "umach"]
METRIC_KEYS = ["dxxdx", "d2xxdx2", "ddx", "dyydy", "d2yydy2", "ddy",
               "dzzdz", "d2zzdz2", "ddz", "rkapa", "dkapa"]
ARRAY_KEYS = ["ru", "rv", "rw", "ro", "tt", "uu", "vv", "ww", "fu", "fv", "fw", "fr", "ft",
              "zru", "zrv", "zrw", "zro", "ztt", "ww1", "ww2", "ww3",
              "bx", "by", "bz", "zbx", "zby", "zbz"]
STATE_KEYS = ["ru", "rv", "rw", "ro", "tt", "bx", "by", "bz"]
# optional FLUXES/BCON branches and the saved relaxation state (mhd_oracle.h)
OPT_INT_KEYS = ["id", "lrem", "lshr", "npey", "npez", "icall"]
OPT_DBL_KEYS = ["dh", "dv", "tu", "dztu", "dztb", "rlax", "tstart", "toff",
                "xp", "yp", "zp", "tc", "qfh", "hh", "zmax", "ttm_min", "ttm_max", "tfac"]
OPT_DEFAULTS = {"npey": 1, "npez": 1, "ttm_min": 1e99, "ttm_max": -1e99, "tfac": 1.0}
COORD_KEYS = ["exx", "wyy", "zee"]

_P = C.POINTER(C.c_double)


class _OrcState(C.Structure):
    _fields_ = ([(k, C.c_int) for k in INT_KEYS]
                + [(k, C.c_double) for k in DBL_KEYS]
                + [("cflmin", C.c_double * 4)]
                + [(k, _P) for k in METRIC_KEYS]
                + [(k, _P) for k in ARRAY_KEYS]
                + [(k, C.c_int) for k in OPT_INT_KEYS]
                + [(k, C.c_double) for k in OPT_DBL_KEYS]
                + [(k, _P) for k in COORD_KEYS])


def build(force: bool = False) -> str:
    src = os.path.join(HERE, "mhd_oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        l = C.CDLL(LIB_PATH)
        for fn in ("orc_fluxes", "orc_bcon", "orc_communicate", "orc_cfl", "orc_step"):
            getattr(l, fn).argtypes = [C.POINTER(_OrcState)]
            getattr(l, fn).restype = None
        l.orc_check.argtypes = [C.POINTER(_OrcState)]
        l.orc_check.restype = C.c_int
        l.orc_flux_profiles.argtypes = [C.POINTER(_OrcState)] + [_P] * 4
        l.orc_flux_profiles.restype = None
        _lib = l
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_P)


class Oracle:
    """One resident box: scalars in `meta`, /BIG/ arrays in `arrays`."""

    def __init__(self, meta: dict, fields: dict):
        self.meta = {k: meta[k] for k in INT_KEYS + DBL_KEYS + OPT_INT_KEYS + OPT_DBL_KEYS if k in meta}
        for k, v in OPT_DEFAULTS.items():
            self.meta.setdefault(k, v)
        self.meta.setdefault("nit", 0)
        for k in ("dt", "timt", "timc", "timi", "umach"):
            self.meta.setdefault(k, 0.0)
        nx, ny, nz = int(meta["nx"]), int(meta["ny"]), int(meta["nz"])
        self.shape = (nx, ny, nz)
        self.metrics = {k: np.ascontiguousarray(meta[k], dtype=np.float64) for k in METRIC_KEYS}
        for k, n in zip(METRIC_KEYS, [nx] * 3 + [ny] * 3 + [nz] * 5):
            if self.metrics[k].shape != (n,):
                raise ValueError(f"metric {k} has shape {self.metrics[k].shape}, expected ({n},)")
        self.coords = {k: np.ascontiguousarray(meta[k], dtype=np.float64) if k in meta else np.zeros(n)
                       for k, n in zip(COORD_KEYS, (nx, ny, nz))}
        self.arrays = {k: np.zeros(self.shape, dtype=np.float64, order="F") for k in ARRAY_KEYS}
        for k in STATE_KEYS:
            if k in fields:
                a = np.asarray(fields[k], dtype=np.float64)
                if a.shape != self.shape:
                    raise ValueError(f"field {k} has shape {a.shape}, expected {self.shape}")
                self.arrays[k][...] = a
        self._s = _OrcState()
        self._sync_in()
        rc = lib().orc_check(C.byref(self._s))
        if rc:
            raise ValueError(f"configuration outside the oracle's scope (code {rc})")

    def _sync_in(self):
        s = self._s
        for k in INT_KEYS + OPT_INT_KEYS:
            setattr(s, k, int(self.meta.get(k, 0)))
        for k in DBL_KEYS + OPT_DBL_KEYS:
            setattr(s, k, float(self.meta.get(k, 0.0)))
        for k in COORD_KEYS:
            setattr(s, k, _ptr(self.coords[k]))
        for k in METRIC_KEYS:
            setattr(s, k, _ptr(self.metrics[k]))
        for k in ARRAY_KEYS:
            setattr(s, k, _ptr(self.arrays[k]))

    def _sync_out(self):
        for k in ("nit",):
            self.meta[k] = getattr(self._s, k)
        for k in ("dt", "timt", "timc", "timi", "umach", "ttm_min", "ttm_max", "tfac"):
            self.meta[k] = getattr(self._s, k)
        self.meta["icall"] = self._s.icall
        self.cflmin = list(self._s.cflmin)

    def _call(self, fn):
        self._sync_in()
        getattr(lib(), fn)(C.byref(self._s))
        self._sync_out()

    def fluxes(self):
        self._call("orc_fluxes")

    def bcon(self):
        self._call("orc_bcon")

    def communicate(self):
        self._call("orc_communicate")

    def cfl(self):
        self._call("orc_cfl")

    def step(self, n: int = 1):
        for _ in range(n):
            self._call("orc_step")

    def flux_profiles(self) -> dict:
        """TTM, ROM, FCONM, FRADM of the current state (padded K=1..NZ;
        FCONM and FRADM on K=2..NZ-1, zero elsewhere)."""
        self._sync_in()
        nz = self.shape[2]
        out = {k: np.zeros(nz) for k in ("ttm", "rom", "fconm", "fradm")}
        lib().orc_flux_profiles(C.byref(self._s), *[_ptr(out[k]) for k in ("ttm", "rom", "fconm", "fradm")])
        return out

    def state(self) -> dict:
        return {k: self.arrays[k].copy(order="F") for k in STATE_KEYS}
