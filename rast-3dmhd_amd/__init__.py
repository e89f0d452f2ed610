"""MI355X-native explicit compressible-MHD time step (Rast 3dmhd STEP).

The product is the C-ABI shared library librast3dmhd.so built from csrc/
(declared in include/rast3dmhd.h) and the Fortran STEP shim in fortran/.
This module is the Python binding of that C-ABI used by the tests and the
benchmark: it loads the library, fills the mhd_desc struct from the
driver's COMMON scalars and moves padded Fortran-order arrays in and out.
There is no fallback: if the HIP library is missing, loading raises.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librast3dmhd.so")
FIELDS = ("ru", "rv", "rw", "ro", "tt", "bx", "by", "bz")
RHS = ("fu", "fv", "fw", "fr", "ft", "ww1", "ww2", "ww3")

MHD_OK, MHD_ERR_ARG, MHD_ERR_SCOPE, MHD_ERR_HIP, MHD_ERR_STATE, MHD_ERR_NONFINITE = range(6)
MHD_FP64, MHD_MIXED = 0, 1
PRECISIONS = {"fp64": MHD_FP64, "mixed": MHD_MIXED}

_DESC_INTS = ["npx", "npy", "npz", "lmag", "lrot", "izcon", "itcon", "ibcon", "precision"]
_DESC_DBLS = ["hx", "h2x", "hy", "h2y", "hz", "h2z", "c13", "c23", "c43",
              "gam1", "gam2", "gam3", "zeta1", "zeta2",
              "cv", "ocv", "ore", "re", "repr", "theta", "grav", "sf", "gamma",
              "omx", "omz", "orm", "rm", "obeta", "tb", "tp", "dzf"]
METRICS = ("dxxdx", "d2xxdx2", "ddx", "dyydy", "d2yydy2", "ddy",
           "dzzdz", "d2zzdz2", "ddz", "rkapa", "dkapa")


# optional FLUXES/BCON branches (mhd_desc members after dzf)
_DESC_OPT_INTS = ["id", "lrem", "lshr", "npey", "npez"]
_DESC_OPT_DBLS = ["dh", "dv", "tu", "dztu", "dztb", "rlax", "tstart", "toff", "xp", "yp", "zp", "tc", "qfh", "hh",
                  "zmax"]
COORDS = ("exx", "wyy", "zee")


class MhdDesc(C.Structure):
    _fields_ = ([(k, C.c_int) for k in _DESC_INTS] + [(k, C.c_double) for k in _DESC_DBLS]
                + [(k, C.c_int) for k in _DESC_OPT_INTS] + [(k, C.c_double) for k in _DESC_OPT_DBLS])


class MhdDevMetrics(C.Structure):
    """mhd_dev_metrics: device pointers to the 11 metric arrays."""
    _fields_ = [(k, C.c_void_p) for k in METRICS]


class MhdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None
_P = C.POINTER(C.c_double)


def load(path: str | None = None):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build() "
                           "(the HIP library is the only implementation)")
    lib = C.CDLL(path)
    lib.mhd_create.argtypes = [C.POINTER(MhdDesc), C.POINTER(C.c_void_p)]
    lib.mhd_set_metrics.argtypes = [C.c_void_p] + [_P] * 11
    lib.mhd_set_background.argtypes = [C.c_void_p, _P, _P]
    lib.mhd_rebase.argtypes = [C.c_void_p]
    lib.mhd_set_rebase_interval.argtypes = [C.c_void_p, C.c_int]
    lib.mhd_set_coordinates.argtypes = [C.c_void_p, _P, _P, _P]
    lib.mhd_upload.argtypes = [C.c_void_p] + [_P] * 8
    lib.mhd_download.argtypes = [C.c_void_p] + [_P] * 8
    lib.mhd_step.argtypes = [C.c_void_p, _P, _P]
    lib.mhd_steps.argtypes = [C.c_void_p, C.c_int, _P, _P]
    lib.mhd_get_clock.argtypes = [C.c_void_p, C.POINTER(C.c_int), _P, _P]
    lib.mhd_set_clock.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
    lib.mhd_fluxes.argtypes = [C.c_void_p] + [_P] * 8
    lib.mhd_set_fixed_dt.argtypes = [C.c_void_p, C.c_double]
    lib.mhd_box_filter.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _P]
    lib.mhd_time_steps.argtypes = [C.c_void_p, C.c_int, _P, _P, _P, _P]
    lib.mhd_flux_profiles.argtypes = [C.c_void_p, _P, _P, _P, _P]
    lib.mhd_step_noalloc.argtypes = [C.POINTER(MhdDesc), C.POINTER(MhdDevMetrics), C.POINTER(C.c_void_p),
                                     C.POINTER(C.c_int), _P, _P, C.c_double, _P, _P, C.c_void_p]
    lib.mhd_step_noalloc.restype = C.c_int
    lib.mhd_last_error.argtypes = [C.c_void_p]
    lib.mhd_last_error.restype = C.c_char_p
    lib.mhd_destroy.argtypes = [C.c_void_p]
    lib.mhd_destroy.restype = None
    V, I, D = C.c_void_p, C.c_int, C.c_double
    lib.mhd_ens_create.argtypes = [C.POINTER(MhdDesc), I, C.POINTER(V)]
    lib.mhd_ens_set_metrics.argtypes = [V, I] + [_P] * 11
    lib.mhd_ens_set_background.argtypes = [V, I, _P, _P]
    lib.mhd_ens_upload.argtypes = [V, I] + [_P] * 8
    lib.mhd_ens_download.argtypes = [V, I] + [_P] * 8
    lib.mhd_ens_set_clock.argtypes = [V, I, I, D, D]
    lib.mhd_ens_get_clock.argtypes = [V, I, C.POINTER(I), _P, _P]
    lib.mhd_ens_set_fixed_dt.argtypes = [V, I, D]
    lib.mhd_ens_steps.argtypes = [V, I, I, _P, _P]
    lib.mhd_ens_set_tiling.argtypes = [V, I, I, I]
    lib.mhd_ens_get_tiling.argtypes = [V, C.POINTER(I), C.POINTER(I), C.POINTER(I)]
    lib.mhd_ens_set_tile.argtypes = [V, I, I, I, I]
    lib.mhd_ens_get_tile.argtypes = [V, C.POINTER(I), C.POINTER(I), C.POINTER(I), C.POINTER(I)]
    lib.mhd_ens_autotune.argtypes = [V, I, C.c_char_p, I]
    lib.mhd_ens_time_steps.argtypes = [V, I, _P]
    lib.mhd_ens_zones.argtypes = [V]
    lib.mhd_ens_zones.restype = C.c_longlong
    lib.mhd_ens_last_error.argtypes = [V]
    lib.mhd_ens_last_error.restype = C.c_char_p
    lib.mhd_ens_destroy.argtypes = [V]
    lib.mhd_ens_destroy.restype = None
    F = C.POINTER(C.c_float)
    lib.mhd_net_create.argtypes = [I, I, I, C.POINTER(V)]
    lib.mhd_net_num_params.argtypes = [V]
    lib.mhd_net_num_params.restype = C.c_longlong
    lib.mhd_net_set_params.argtypes = [V, F]
    lib.mhd_net_get_params.argtypes = [V, F]
    lib.mhd_net_forward.argtypes = [V, I, I, I, F, F]
    lib.mhd_net_gradient.argtypes = [V, I, I, I, F, F, F, F]
    lib.mhd_net_add_sample.argtypes = [V, I, I, I, F, F]
    lib.mhd_net_num_samples.argtypes = [V]
    lib.mhd_net_clear_samples.argtypes = [V]
    lib.mhd_net_train.argtypes = [V, I, C.c_float, _P, _P]
    lib.mhd_net_loss.argtypes = [V, I, _P]
    lib.mhd_net_time_forward.argtypes = [V, I, I, I, I, _P]
    lib.mhd_set_closure.argtypes = [V, V, _P, _P, _P, _P, D, D]
    lib.mhd_set_closure.restype = C.c_int
    lib.mhd_closure_active.argtypes = [V]
    lib.mhd_closure_active.restype = C.c_int
    lib.mhd_net_last_error.argtypes = [V]
    lib.mhd_net_last_error.restype = C.c_char_p
    lib.mhd_net_destroy.argtypes = [V]
    lib.mhd_net_destroy.restype = None
    for fn in ("mhd_net_create", "mhd_net_set_params", "mhd_net_get_params", "mhd_net_forward", "mhd_net_gradient",
               "mhd_net_add_sample", "mhd_net_num_samples", "mhd_net_clear_samples", "mhd_net_train",
               "mhd_net_loss", "mhd_net_time_forward"):
        getattr(lib, fn).restype = C.c_int
    for fn in ("mhd_ens_create", "mhd_ens_set_metrics", "mhd_ens_set_background", "mhd_ens_upload",
               "mhd_ens_download", "mhd_ens_set_clock", "mhd_ens_get_clock", "mhd_ens_set_fixed_dt", "mhd_ens_steps", "mhd_ens_set_tiling",
               "mhd_ens_get_tiling", "mhd_ens_set_tile", "mhd_ens_get_tile", "mhd_ens_autotune",
               "mhd_ens_time_steps"):
        getattr(lib, fn).restype = C.c_int
    for fn in ("mhd_create", "mhd_set_metrics", "mhd_set_coordinates", "mhd_set_background", "mhd_upload", "mhd_download", "mhd_step", "mhd_steps",
               "mhd_get_clock", "mhd_set_clock", "mhd_fluxes", "mhd_time_steps", "mhd_set_fixed_dt", "mhd_box_filter",
               "mhd_flux_profiles", "mhd_rebase", "mhd_set_rebase_interval"):
        getattr(lib, fn).restype = C.c_int
    _lib = lib
    return lib


def make_desc(meta: dict, precision: str = "fp64") -> MhdDesc:
    """mhd_desc from a dict of COMMON scalars (names as in the reference,
    lower case; npx/npy/npz are interior sizes)."""
    d = MhdDesc()
    d.precision = PRECISIONS[precision]
    d.npx, d.npy, d.npz = int(meta["npx"]), int(meta["npy"]), int(meta["npz"])
    d.lmag, d.lrot = int(meta.get("lmag", 0)), int(meta.get("lrot", 0))
    d.izcon = int(meta.get("izc", meta.get("izcon", 0)))
    d.itcon = int(meta.get("itc", meta.get("itcon", 0)))
    d.ibcon = int(meta.get("ibc", meta.get("ibcon", 0)))
    for k in _DESC_DBLS + _DESC_OPT_DBLS:
        setattr(d, k, float(meta.get(k, 0.0)))
    for k in _DESC_OPT_INTS:
        setattr(d, k, int(meta.get(k, 0)))
    return d


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def metric_arrays(meta: dict, desc: MhdDesc) -> list:
    """The 11 metric arrays of a box as contiguous FP64 arrays, each checked
    against the padded length the C-ABI copies (NX = npx+2 for the x
    metrics, NY = npy+2 for y, NZ = npz+2 for z, RKAPA and DKAPA); raises
    ValueError before any library call when one is short or long."""
    nx, ny, nz = desc.npx + 2, desc.npy + 2, desc.npz + 2
    want = dict(zip(METRICS, (nx, nx, nx, ny, ny, ny, nz, nz, nz, nz, nz)))
    out = []
    for k in METRICS:
        a = np.ascontiguousarray(meta[k], dtype=np.float64)
        n = want[k]
        if a.ndim != 1 or a.size != n:
            raise ValueError(f"metric {k}: {a.shape}, expected ({n},)")
        out.append(a)
    return out


class Box:
    """One resident box on the current GPU, driven through the C-ABI.

    precision="fp64" is the drop-in form (bit-identical to the reference);
    "mixed" carries FP32 perturbations about an FP64 background (rho0(z),
    T0(z)).  For the mixed form, set_background() fixes the background; if
    it was not called, the first upload() uses the horizontal means of the
    uploaded RO and TT planes."""

    def __init__(self, meta: dict, precision: str = "fp64"):
        self.lib = load()
        self.meta = meta
        self.precision = precision
        self.desc = make_desc(meta, precision)
        self._have_bg = False
        self.shape = (self.desc.npx + 2, self.desc.npy + 2, self.desc.npz + 2)
        self.lmag = bool(self.desc.lmag)
        self._metrics = metric_arrays(meta, self.desc)
        h = C.c_void_p()
        rc = self.lib.mhd_create(C.byref(self.desc), C.byref(h))
        if rc:
            raise MhdError(rc, self.lib.mhd_last_error(None).decode())
        self.h = h
        self._check(self.lib.mhd_set_metrics(self.h, *[_ptr(a) for a in self._metrics]))
        if all(k in meta for k in COORDS):
            # EXX, WYY, ZEE: used by the plume, heat-loss, ID=2 and relaxation branches
            lens = (self.shape[0], self.shape[1], self.shape[2])
            self._coords = [np.ascontiguousarray(meta[k], dtype=np.float64) for k in COORDS]
            for k, a, n in zip(COORDS, self._coords, lens):
                if a.shape != (n,):
                    raise ValueError(f"coordinate {k}: {a.shape}, expected ({n},)")
            self._check(self.lib.mhd_set_coordinates(self.h, *[_ptr(a) for a in self._coords]))

    def _check(self, rc):
        if rc:
            raise MhdError(rc, self.lib.mhd_last_error(self.h).decode())

    def _padded(self, a):
        a = np.asarray(a, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"array shape {a.shape}, expected {self.shape}")
        return np.asfortranarray(a)

    def set_background(self, ro0, tt0):
        """FP64 background profiles, arrays of npz+2 plane values (z ghost
        planes included); mixed form only."""
        ro0 = np.ascontiguousarray(ro0, dtype=np.float64)
        tt0 = np.ascontiguousarray(tt0, dtype=np.float64)
        if ro0.shape != (self.shape[2],) or tt0.shape != (self.shape[2],):
            raise ValueError(f"background profiles must have {self.shape[2]} planes")
        self._check(self.lib.mhd_set_background(self.h, _ptr(ro0), _ptr(tt0)))
        self._have_bg = True

    def rebase(self):
        """Mixed form: re-centre the background on the current horizontal
        means of rho and T (the state is re-expressed about it in place)."""
        self._check(self.lib.mhd_rebase(self.h))

    def set_rebase_interval(self, nsteps: int):
        """Mixed form: rebase() after every nsteps-th step (0: never)."""
        self._check(self.lib.mhd_set_rebase_interval(self.h, int(nsteps)))

    def upload(self, fields: dict):
        if self.precision == "mixed" and not self._have_bg:
            self.set_background(*_horizontal_means(fields))
        arrs = [self._padded(fields[k]) if (k in fields and (self.lmag or k not in ("bx", "by", "bz"))) else None
                for k in FIELDS]
        self._check(self.lib.mhd_upload(self.h, *[_ptr(a) for a in arrs]))

    def download(self) -> dict:
        out = {k: np.zeros(self.shape, order="F") for k in FIELDS}
        ptrs = [_ptr(out[k]) if (self.lmag or k not in ("bx", "by", "bz")) else None for k in FIELDS]
        self._check(self.lib.mhd_download(self.h, *ptrs))
        return out

    def step(self):
        dt, um = C.c_double(), C.c_double()
        self._check(self.lib.mhd_step(self.h, C.byref(dt), C.byref(um)))
        return dt.value, um.value

    def steps(self, n: int):
        dt, um = C.c_double(), C.c_double()
        self._check(self.lib.mhd_steps(self.h, int(n), C.byref(dt), C.byref(um)))
        return dt.value, um.value

    def time_steps(self, n: int) -> dict:
        """Mean milliseconds per step and per kernel launch over n steps,
        from HIP events on the library's stream."""
        v = [C.c_double() for _ in range(4)]
        self._check(self.lib.mhd_time_steps(self.h, int(n), *[C.byref(x) for x in v]))
        return dict(step=v[0].value, stage=v[1].value, cfl=v[2].value, bcon=v[3].value)

    def clock(self):
        nit, timc, timt = C.c_int(), C.c_double(), C.c_double()
        self._check(self.lib.mhd_get_clock(self.h, C.byref(nit), C.byref(timc), C.byref(timt)))
        return nit.value, timc.value, timt.value

    def set_clock(self, nit: int, timc: float, timi: float):
        self._check(self.lib.mhd_set_clock(self.h, int(nit), float(timc), float(timi)))

    def set_fixed_dt(self, dt: float):
        """dt > 0: step with this DT from now on; 0: back to the CFL DT."""
        self._check(self.lib.mhd_set_fixed_dt(self.h, float(dt)))

    def set_closure(self, net, in_profile=None, in_std=None, out_profile=None, out_std=None, val_error: float = 0.0,
                    threshold: float = 0.1):
        """Attach a trained Net (mixed form).  Inputs enter as (value -
        in_profile[c][k]) / in_std[c]; the forcing added to the right-hand
        sides of RU, RV, RW, TT, BX, BY, BZ is output * out_std[c] +
        out_profile[c][k] (profiles [8][npz] and [7][npz]).  net=None, or
        val_error > threshold, leaves the box stepping plainly."""
        nz = self.desc.npz
        arrs = [None] * 4
        if net is not None:
            arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (in_profile, in_std, out_profile, out_std)]
            want = [(8, nz), (8,), (7, nz), (7,)]
            for a, w in zip(arrs, want):
                if a.shape != w:
                    raise ValueError(f"closure normalisation shapes {[a.shape for a in arrs]}, expected {want}")
        self._check(self.lib.mhd_set_closure(self.h, net.h if net is not None else None, *[_ptr(a) for a in arrs],
                                             float(val_error), float(threshold)))
        self._net = net  # keep it alive while attached

    def closure_active(self) -> bool:
        return bool(self.lib.mhd_closure_active(self.h))

    def box_filter(self, factor: int, zmode: int = 1, source: str = "state") -> np.ndarray:
        """Box-filtered state ("state") or right-hand side ("rhs") on the grid
        coarser by `factor`: array [field][K][J][I]."""
        d = self.desc
        nzc = d.npz // factor if zmode == 0 else (d.npz - 1) // factor + 1
        nf = 8 if self.lmag else 5
        out = np.zeros((nf, nzc, d.npy // factor, d.npx // factor))
        self._check(self.lib.mhd_box_filter(self.h, int(factor), int(zmode), 0 if source == "state" else 1, _ptr(out)))
        return out

    def flux_profiles(self) -> dict:
        """Horizontal means TTM, ROM and the mean convective / radiative
        flux profiles FCONM, FRADM of the current state, one value per
        interior plane (reference K = 2..NPZ+1)."""
        nz = self.desc.npz
        out = {k: np.zeros(nz) for k in ("ttm", "rom", "fconm", "fradm")}
        self._check(self.lib.mhd_flux_profiles(self.h, *[_ptr(out[k]) for k in ("ttm", "rom", "fconm", "fradm")]))
        return out

    def fluxes(self) -> dict:
        out = {k: np.zeros(self.shape, order="F") for k in RHS}
        ptrs = [_ptr(out[k]) if (self.lmag or k not in ("ww1", "ww2", "ww3")) else None for k in RHS]
        self._check(self.lib.mhd_fluxes(self.h, *ptrs))
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.mhd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _horizontal_means(fields: dict):
    ro = np.asarray(fields["ro"], dtype=np.float64)
    tt = np.asarray(fields["tt"], dtype=np.float64)
    return ro[1:-1, 1:-1, :].mean(axis=(0, 1)), tt[1:-1, 1:-1, :].mean(axis=(0, 1))


class Ensemble:
    """Independent boxes stepped together on the current GPU
    (include/rast3dmhd_ensemble.h): one persistent launch per RK stage over
    all boxes, replayed from a captured HIP graph.  metas: one dict of
    COMMON scalars + metrics per box (as for Box); all boxes share lmag and
    the precision.  For the mixed form the background of each box is the
    horizontal mean of its first upload unless set_background was called."""

    def __init__(self, metas: list, precision: str = "fp64"):
        self.lib = load()
        self.metas = metas
        self.precision = precision
        self.descs = (MhdDesc * len(metas))(*[make_desc(m, precision) for m in metas])
        self.shapes = [(d.npx + 2, d.npy + 2, d.npz + 2) for d in self.descs]
        self.lmag = bool(self.descs[0].lmag)
        self.n = len(metas)
        metrics = [metric_arrays(m, d) for m, d in zip(metas, self.descs)]
        h = C.c_void_p()
        rc = self.lib.mhd_ens_create(self.descs, self.n, C.byref(h))
        if rc:
            raise MhdError(rc, self.lib.mhd_ens_last_error(None).decode())
        self.h = h
        self._have_bg = [False] * self.n
        for b, arrs in enumerate(metrics):
            self._check(self.lib.mhd_ens_set_metrics(self.h, b, *[_ptr(a) for a in arrs]))

    def _check(self, rc):
        if rc:
            raise MhdError(rc, self.lib.mhd_ens_last_error(self.h).decode())

    def set_background(self, b: int, ro0, tt0):
        ro0 = np.ascontiguousarray(ro0, dtype=np.float64)
        tt0 = np.ascontiguousarray(tt0, dtype=np.float64)
        nz = self.shapes[b][2]
        if ro0.shape != (nz,) or tt0.shape != (nz,):
            raise ValueError(f"background profiles of box {b} must have {nz} planes")
        self._check(self.lib.mhd_ens_set_background(self.h, b, _ptr(ro0), _ptr(tt0)))
        self._have_bg[b] = True

    def upload(self, b: int, fields: dict):
        if self.precision == "mixed" and not self._have_bg[b]:
            self.set_background(b, *_horizontal_means(fields))
        arrs = []
        for k in FIELDS:
            if k in ("bx", "by", "bz") and not self.lmag:
                arrs.append(None)
                continue
            a = np.asfortranarray(np.asarray(fields[k], dtype=np.float64))
            if a.shape != self.shapes[b]:
                raise ValueError(f"box {b} {k}: shape {a.shape}, expected {self.shapes[b]}")
            arrs.append(a)
        self._check(self.lib.mhd_ens_upload(self.h, b, *[_ptr(a) for a in arrs]))

    def download(self, b: int) -> dict:
        out = {k: np.zeros(self.shapes[b], order="F") for k in FIELDS}
        ptrs = [_ptr(out[k]) if (self.lmag or k not in ("bx", "by", "bz")) else None for k in FIELDS]
        self._check(self.lib.mhd_ens_download(self.h, b, *ptrs))
        return out

    def steps(self, n: int, graph: bool = True):
        dt = np.zeros(self.n)
        um = np.zeros(self.n)
        self._check(self.lib.mhd_ens_steps(self.h, int(n), 1 if graph else 0, _ptr(dt), _ptr(um)))
        return dt, um

    def clock(self, b: int):
        nit, timc, timt = C.c_int(), C.c_double(), C.c_double()
        self._check(self.lib.mhd_ens_get_clock(self.h, b, C.byref(nit), C.byref(timc), C.byref(timt)))
        return nit.value, timc.value, timt.value

    def set_clock(self, b: int, nit: int, timc: float, timi: float):
        self._check(self.lib.mhd_ens_set_clock(self.h, b, int(nit), float(timc), float(timi)))

    def set_fixed_dt(self, b: int, dt: float):
        self._check(self.lib.mhd_ens_set_fixed_dt(self.h, b, float(dt)))

    def set_tiling(self, tile_y: int, kdepth: int, grid: int):
        self._check(self.lib.mhd_ens_set_tiling(self.h, int(tile_y), int(kdepth), int(grid)))

    def tiling(self):
        v = [C.c_int() for _ in range(3)]
        self._check(self.lib.mhd_ens_get_tiling(self.h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_tile(self, tile_x: int, tile_y: int, kdepth: int, grid: int):
        """Tiling with the tile width: 64x4, 64x8 or (mixed form) 32x8."""
        self._check(self.lib.mhd_ens_set_tile(self.h, int(tile_x), int(tile_y), int(kdepth), int(grid)))

    def tile(self):
        """(tile_x, tile_y, kdepth, grid) of the current tiling."""
        v = [C.c_int() for _ in range(4)]
        self._check(self.lib.mhd_ens_get_tile(self.h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def autotune(self, nsteps: int = 5):
        """Times the candidate tilings, keeps the fastest (state untouched);
        returns the list of candidates with their ms per step."""
        import json
        buf = C.create_string_buffer(16384)
        self._check(self.lib.mhd_ens_autotune(self.h, int(nsteps), buf, len(buf)))
        return json.loads(buf.value.decode())

    def time_steps(self, n: int) -> float:
        ms = C.c_double()
        self._check(self.lib.mhd_ens_time_steps(self.h, int(n), C.byref(ms)))
        return ms.value

    def zones(self) -> int:
        return int(self.lib.mhd_ens_zones(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.lib.mhd_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _fptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


class Net:
    """The closure network (include/rast3dmhd_closure.h) on the current GPU:
    3x3x3 convolutions cin -> width -> width -> cout, GELU, BF16 MFMA."""

    def __init__(self, cin: int = 8, width: int = 32, cout: int = 7):
        self.lib = load()
        self.cin, self.width, self.cout = cin, width, cout
        h = C.c_void_p()
        rc = self.lib.mhd_net_create(cin, width, cout, C.byref(h))
        if rc:
            raise MhdError(rc, self.lib.mhd_net_last_error(None).decode())
        self.h = h
        self.nparam = int(self.lib.mhd_net_num_params(self.h))

    def _check(self, rc):
        if rc:
            raise MhdError(rc, self.lib.mhd_net_last_error(self.h).decode())

    def set_params(self, vec):
        v = np.ascontiguousarray(vec, dtype=np.float32)
        if v.size != self.nparam:
            raise ValueError(f"{v.size} parameters, expected {self.nparam}")
        self._check(self.lib.mhd_net_set_params(self.h, _fptr(v)))

    def params(self):
        v = np.zeros(self.nparam, dtype=np.float32)
        self._check(self.lib.mhd_net_get_params(self.h, _fptr(v)))
        return v

    @staticmethod
    def _vol(x, c):
        x = np.ascontiguousarray(x, dtype=np.float32)
        if x.ndim != 4 or x.shape[3] != c:
            raise ValueError(f"expected [nz][ny][nx][{c}], got {x.shape}")
        return x

    def forward(self, x):
        x = self._vol(x, self.cin)
        nz, ny, nx, _ = x.shape
        y = np.zeros((nz, ny, nx, self.cout), dtype=np.float32)
        self._check(self.lib.mhd_net_forward(self.h, nx, ny, nz, _fptr(x), _fptr(y)))
        return y

    def gradient(self, x, dy, want_dx=False):
        x = self._vol(x, self.cin)
        dy = self._vol(dy, self.cout)
        nz, ny, nx, _ = x.shape
        g = np.zeros(self.nparam, dtype=np.float32)
        dx = np.zeros_like(x) if want_dx else None
        self._check(self.lib.mhd_net_gradient(self.h, nx, ny, nz, _fptr(x), _fptr(dy), _fptr(g), _fptr(dx)))
        return (g, dx) if want_dx else g

    def add_sample(self, x, t):
        x = self._vol(x, self.cin)
        t = self._vol(t, self.cout)
        nz, ny, nx, _ = x.shape
        self._check(self.lib.mhd_net_add_sample(self.h, nx, ny, nz, _fptr(x), _fptr(t)))

    def clear_samples(self):
        self._check(self.lib.mhd_net_clear_samples(self.h))

    def train(self, iters: int, lr: float):
        a, b = C.c_double(), C.c_double()
        self._check(self.lib.mhd_net_train(self.h, int(iters), float(lr), C.byref(a), C.byref(b)))
        return a.value, b.value

    def loss(self, s: int):
        v = C.c_double()
        self._check(self.lib.mhd_net_loss(self.h, int(s), C.byref(v)))
        return v.value

    def time_forward(self, nx, ny, nz, reps=10):
        v = C.c_double()
        self._check(self.lib.mhd_net_time_forward(self.h, nx, ny, nz, reps, C.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.mhd_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
