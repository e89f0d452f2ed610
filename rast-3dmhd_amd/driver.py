"""Host driver: the reference main program's run sequence around the GPU step.

THREEDMHD (3dmhd.f) sets up COMMON from the compile-time PARAMETER block and
the SETUP deck, builds the initial state (STATIC) or reads a dump back,
computes the initial time-step diagnostics, writes the log header, calls
STEP once per time step and dumps every NSTEP0 steps; a DT below 1e-8 ends
the run with an emergency dump.  This module does the same for one resident
domain on the GPU (the step is the C-ABI's mhd_step), and writes and reads
the reference's own files (dumpio), so that

  * a fresh run writes the .dat0 rank files, .par and .lis that the
    reference's MPI build writes for the same deck, byte for byte;
  * a run restarts from the reference's dump files and continues as the
    reference's uninterrupted run does.

One deliberate difference: the reference's restart leaves the physical
z ghost planes of every field at zero (it reads interiors only,
3dmhd.f:334-350), so 1/rho is infinite there and, with LMAG, the boundary
planes of BX and BY turn non-finite in the first step (BZ = 0 times an
infinite velocity, 3dmhdsub.f:2228-2240).  This driver fills those ghost
planes with the adjacent boundary plane instead; the ghosts only enter
products with the zero boundary values of RW and BZ and fields that BCON
overwrites, so every interior value is the one the uninterrupted run has.
A second: a non-finite DT or Mach number shuts the run down like a DT below
1e-8 (the reference's test DT.LT.1.0E-08 is false for NaN and it carries
on).

A deck is the pair the reference compiles in: PARAMETER overrides
(3dmhdparam.f names) and SETUP's PAR/IPAR slots (3dmhdset.f), as literal
text or numbers, with the defaults of fortran/mkparam.py.

    python rast-3dmhd_amd/driver.py --deck deck.json --out rundir [--precision mixed]
"""

from __future__ import annotations

import argparse
import importlib
import importlib.util
import json
import math
import os
import sys
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))


def _mkparam():
    spec = importlib.util.spec_from_file_location("mkparam", os.path.join(HERE, "fortran", "mkparam.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _mods():
    return (importlib.import_module("rast-3dmhd_amd.config"), importlib.import_module("rast-3dmhd_amd.static"),
            importlib.import_module("rast-3dmhd_amd.dumpio"))


DT_MIN = float(np.float32(1.0e-08))  # the DT.LT.1.0E-08 abort test, a default-REAL literal


def literal(v) -> float:
    """Value of a Fortran literal as the compiler folds it: D literals in
    double, default-REAL (E) literals and their quotients in float32,
    widened; numbers are taken as float32 values like E literals."""
    if isinstance(v, bool):
        raise TypeError("logical where a number is expected")
    if not isinstance(v, str):
        return float(np.float32(v))
    s = v.strip().upper()
    if "D" in s:
        return float(s.replace("D", "E"))
    if "/" in s:
        a, b = s.split("/")
        return float(np.float32(float(a)) / np.float32(float(b)))
    return float(np.float32(float(s)))


def logical(v) -> bool:
    return v if isinstance(v, bool) else str(v).strip().upper() == ".TRUE."


@dataclass
class Deck:
    """PARAMETER overrides and SETUP deck of one run (names and slots of
    3dmhdparam.f and 3dmhdset.f)."""
    params: dict
    par: dict = field(default_factory=dict)
    ipar: dict = field(default_factory=dict)
    fname: str = "rast3d"

    @classmethod
    def from_json(cls, path: str) -> "Deck":
        d = json.load(open(path))
        return cls(d["params"], {int(k): v for k, v in d.get("par", {}).items()},
                   {int(k): int(v) for k, v in d.get("ipar", {}).items()}, d.get("fname", "rast3d"))


class Setup:
    """COMMON as the main program fills it (3dmhd.f:105-279, 378-380):
    constants, IPAR/PAR after the driver's own assignments, grid metrics and
    the conductivity profile, as the step's metadata dictionary."""

    def __init__(self, deck: Deck, ntotal=None, nstep0=None, nstart=None):
        mk = _mkparam()
        cfg, _, dio = _mods()
        P = dict(mk.PARAM_DEFAULTS)
        for k, v in deck.params.items():
            if k.upper() not in P:
                raise KeyError(f"unknown PARAMETER {k}")
            P[k.upper()] = v
        self.P = P
        self.lmag, self.lrot = logical(P["LMAG"]), logical(P["LROT"])
        for sw in ("LREM", "LSHR"):
            if logical(P[sw]):
                raise ValueError(f"{sw} is not supported by this build's step")
        npx, npy, npz = int(P["NPX"]), int(P["NPY"]), int(P["NPZ"])
        self.layout = dio.Layout(npx, npy, npz, int(P["NPEY"]), int(P["NPEZ"]))
        self.ngrid = int(P["NGRID"])
        if self.ngrid not in (0, 1):
            raise ValueError(f"NGRID={self.ngrid}: the host set-up covers NGRID 0 and 1")
        # SETUP: zeroed deck, then the defaults and the run's values
        ipar = [0] * 33
        for k, v in {**mk.IPAR_DEFAULTS, **deck.ipar}.items():
            ipar[k] = int(v)
        for k, v in ((3, ntotal), (4, nstep0), (5, nstart)):
            if v is not None:
                ipar[k] = int(v)
        par = [0.0] * 65
        for k, v in {**mk.PAR_DEFAULTS, **deck.par}.items():
            par[k] = literal(v)
        self.ncase, self.ncasep, self.ntotal, self.nstep0, self.nstart = ipar[1:6]
        izc, itc, ibc = int(P["IZCON"]), int(P["ITCON"]), int(P["IBCON"])
        sf = literal(P["SFF"])
        # the driver's own IPAR/PAR assignments (3dmhd.f:142-218)
        for k, v in ((6, P["IXCON"]), (7, P["IYCON"]), (8, izc), (9, itc), (10, ibc), (11, npx), (12, npy),
                     (13, npz), (14, self.layout.npe), (17, self.ngrid), (18, P["NCOL"]), (19, P["NROW"]),
                     (20, P["ID"]), (21, P["NPEY"])):
            ipar[k] = int(v)
        for k, name in ((24, "XX1"), (25, "XX2"), (26, "YY1"), (27, "YY2"), (28, "ZZ1"), (29, "ZZ2")):
            par[k] = literal(P[name])
        par[30] = sf
        for k, name in zip(range(39, 51), ("XA", "XB", "XC", "XD", "YA", "YB", "YC", "YD", "ZA", "ZB", "ZC", "ZD")):
            par[k] = literal(P[name])
        par[56], par[57] = literal(P["DH"]), literal(P["DV"])
        if int(P["ID"]) != 0:
            raise ValueError("ID != 0 (density diffusion) is not supported by this build's step")
        if int(P["IXCON"]) != 0 or int(P["IYCON"]) != 0:
            raise ValueError("FLUXES: Non-periodic horizontal boundaries")
        self.ipar, self.par = ipar[1:], par[1:]
        names = ("RE", "PR", "THETA", "GRAV", "RY", "ANG", "RM", "BETA", "PZP", "SIGMA", "POLYS", "TP", "TC",
                 "XP", "YP", "ZP", "XMAX", "YMAX", "ZMAX", "AMPT", "AMPB", "BFH", "BZP")
        self.named = {n: par[i + 1] for i, n in enumerate(names)}
        self.named.update(QFH=par[53], GAMMA=par[54], HH=par[55], TSTART=par[58], TOFF=par[59], RLAX=par[60])
        if self.named["RLAX"] > 0.0:
            raise ValueError("RLAX > 0 (flux relaxation) is not supported by this build's step")
        n = self.named
        meta = cfg.constants(npx, npy, npz, {k: par[k] for k in range(1, 65)}, lmag=self.lmag, lrot=self.lrot,
                             sff=sf, izcon=izc, itcon=itc, ibcon=ibc)
        mt = cfg.metrics(npx, npy, npz, xmax=n["XMAX"], ymax=n["YMAX"], zmax=n["ZMAX"], ngrid=self.ngrid,
                         xx=(par[24], par[25]), yy=(par[26], par[27]), zz=(par[28], par[29]))
        meta.update(mt)
        self.rkapst = 1.0 if n["GRAV"] == 0.0 else (n["POLYS"] + 1.0) * n["THETA"] / n["GRAV"]
        meta["rkapa"], meta["dkapa"] = kappa_profile(mt["zee"], n["PZP"], n["SIGMA"], self.rkapst)
        self.meta = meta
        self.sf = sf

    def fresh_state(self):
        """STATIC's state with the first COMMUNICATE (3dmhd.f:290, 365)."""
        _, S, _ = _mods()
        n, m = self.named, self.meta
        st, tp, _ = S.initial_state(
            m["npx"], m["npy"], m["npz"], m["zee"], theta=n["THETA"], grav=n["GRAV"], zmax=n["ZMAX"],
            ngrid=self.ngrid, zz1=self.par[27], zz2=self.par[28], pzp=n["PZP"], sigma=n["SIGMA"],
            rkapst=self.rkapst, ampt=n["AMPT"], lmag=self.lmag, ampb=n["AMPB"], bfh=n["BFH"], bzp=n["BZP"],
            obeta=m["obeta"])
        return st, float(tp[-1])


def kappa_profile(zee, pzp: float, sigma: float, rkapst: float):
    """RKAPA, DKAPA of the main program (3dmhd.f:269-279): 1 and 0, or the
    tanh transition at depth PZP of width SIGMA and its z derivative."""
    zee = np.asarray(zee, dtype=np.float64)
    if pzp == 0.0:
        return np.ones(zee.size), np.zeros(zee.size)
    rk = np.empty(zee.size)
    dk = np.empty(zee.size)
    for k, z in enumerate(zee):
        a = (z - pzp) / sigma
        rk[k] = 1.0 + (rkapst - 1.0) / 2.0 * (1.0 + math.tanh(a))
        c = math.cosh(a)
        dk[k] = (rkapst - 1.0) / 2.0 / sigma / (c * c)
    return rk, dk


def initial_dt(meta: dict, st: dict):
    """The pointwise time-step limits the main program logs before the first
    step (3dmhd.f:378-496): DT1 (fast mode), DT2 (conduction), DT3
    (viscosity), DT4 (resistivity, 0 without LMAG), DT = SF*min and
    NSOUND = NINT(1/DT)."""
    sl = (slice(1, -1), slice(1, -1), slice(1, -1))
    ru, rv, rw, ro, tt = (np.asarray(st[k], dtype=np.float64)[sl] for k in ("ru", "rv", "rw", "ro", "tt"))
    gamma, obeta = float(meta["gamma"]), float(meta["obeta"])
    cv = 1.0 / (gamma - 1.0)
    uu = (ru * ru + rv * rv + rw * rw) / (ro * ro)
    vv = gamma * tt
    lmag = bool(int(meta["lmag"]))
    if lmag:
        bx, by, bz = (np.asarray(st[k], dtype=np.float64)[sl] for k in ("bx", "by", "bz"))
        vv = vv + (bx * bx + by * by + bz * bz) * obeta / ro
    uu = uu + vv + 2.0 * np.sqrt(uu * vv)
    ddx = np.asarray(meta["ddx"])[1:-1][:, None, None]
    ddy = np.asarray(meta["ddy"])[1:-1][None, :, None]
    ddz = np.asarray(meta["ddz"])[1:-1][None, None, :]
    dd = np.minimum(np.minimum(ddx, ddy), ddz)
    rk = np.asarray(meta["rkapa"])[1:-1][None, None, :]
    repr_, re, rm, sf = float(meta["repr"]), float(meta["re"]), float(meta["rm"]), float(meta["sf"])
    dt1 = float((dd / np.sqrt(uu)).min())
    dt2 = float((0.5 * dd * dd * repr_ * cv * ro / rk).min())
    dt3 = float((0.375 * dd * dd * re * ro).min())
    if lmag:
        dt4 = float((0.5 * dd * dd * rm).min())
        dt = sf * min(dt1, dt2, dt3, dt4)
    else:
        dt4 = 0.0
        dt = sf * min(dt1, dt2, dt3)
    nsound = int(math.floor(1.0 / dt + 0.5)) if dt > 0 else 0  # NINT
    return (dt1, dt2, dt3, dt4), dt, nsound


def fill_restart_ghosts(st: dict) -> dict:
    """Physical z ghost planes of a restarted state: the adjacent boundary
    plane of every field (see the module docstring)."""
    for a in st.values():
        a[:, :, 0] = a[:, :, 1]
        a[:, :, -1] = a[:, :, -2]
    return st


class Run:
    """One run of the deck in directory `outdir`: fresh (NSTART = 0) or a
    restart from dump NSTART of the files FINP there.  `stepper` makes the
    object that steps the state (default: the GPU library's Box in the given
    precision); it needs upload, step, download, clock and set_clock."""

    def __init__(self, deck: Deck, outdir: str, precision: str = "fp64", ntotal=None, nstep0=None, nstart=None,
                 stepper=None, log=None):
        self.setup = Setup(deck, ntotal, nstep0, nstart)
        self.deck, self.outdir = deck, outdir
        self.precision = precision
        self.stepper_factory = stepper
        self.log = log or (lambda *a: None)
        os.makedirs(outdir, exist_ok=True)

    def _prefix(self, name):
        return os.path.join(self.outdir, name)

    def run(self) -> dict:
        _, _, dio = _mods()
        s, m = self.setup, self.setup.meta
        lay, lmag = s.layout, s.lmag
        fin, fout = self._prefix(self.deck.fname), self._prefix(self.deck.fname)
        timi = 0.0
        restart_size = None
        if s.nstart == 0:
            st, tb = s.fresh_state()
        else:
            p1 = dio.read_par(fin + ".par", 1)
            old = dio.par_layout(p1)
            if (old.npx, old.nry, old.nrz) != (lay.npx, lay.nry, lay.nrz):
                raise ValueError("MAIN:  Grid size mismatch, no interpolation")
            pn = dio.read_par(fin + ".par", s.nstart)
            timi = pn[63] + pn[64]
            st = fill_restart_ghosts(dio.read_dump(fin, s.nstart, lay, lmag))
            tb = float(st["tt"][1, 1, lay.npz])  # TB = TT(2,2,NZ-ILAP/2) of the bottom ranks
            restart_size = (old.npx, old.nry * lay.npey, old.nrz * lay.npez)
        m["tb"] = tb
        dts, dt0, nsound = initial_dt(m, st)
        lis = dio.LisFile(fout + ".lis")
        nm = s.named
        hdr = dio.lis_header(ncase=s.ncase, ncasep=s.ncasep, nstart=s.nstart, npx=lay.npx, npy=lay.npy,
                             npz=lay.npz, lrot=s.lrot, lmag=lmag, izc=int(m["izc"]), itc=int(m["itc"]), par=nm,
                             sf=s.sf, ampt=nm["AMPT"], ampb=nm["AMPB"], finp=self.deck.fname,
                             fout=self.deck.fname, timi=timi, restart_size=restart_size, ntotal=s.ntotal,
                             nstep0=s.nstep0, dts=dts, dt=dt0, nsound=nsound)
        for i, line in enumerate(hdr):
            lis.write(i + 1, line)
        irec = 59
        stepper = self.stepper_factory(m) if self.stepper_factory else \
            importlib.import_module("rast-3dmhd_amd").Box(m, precision=self.precision)
        summary = {"dumps": 0, "steps": 0, "shutdown": False}
        try:
            stepper.upload(st)
            stepper.set_clock(0, 0.0, timi)
            ndump0 = 0
            nit = 0
            for nk in range(1, s.ntotal + 1):
                try:
                    dt, umach = stepper.step()
                except Exception as e:  # a non-finite DT or Mach number: shut down as for a tiny DT
                    if getattr(e, "code", None) != 5:
                        raise
                    dt, umach = float("nan"), float("nan")
                nit, timc, timt = stepper.clock()
                if nk == s.ntotal:
                    t64, t65 = timt, 0.0
                else:
                    t64, t65 = timi, timc
                rec = dio.par1_record(s.ipar, s.par, nit, dt, t64, t65)
                summary["steps"] = nit
                if not (dt >= DT_MIN):
                    ndump0 += 1
                    dio.write_dump(fout, ndump0, stepper.download(), lay, lmag)
                    dio.write_par(fout + ".par", ndump0, rec)
                    base = 58 + (7 * nit) // s.nstep0
                    for i, line in enumerate(dio.lis_shutdown(dt, nit - 1)):
                        lis.write(base + 1 + i, line)
                    self.log(f"shut down at step {nit}: DT = {dt}")
                    summary.update(shutdown=True, dumps=ndump0, dt=dt)
                    return summary
                if nit % s.nstep0 == 0:
                    ndump0 += 1
                    dio.write_dump(fout, ndump0, stepper.download(), lay, lmag)
                    dio.write_par(fout + ".par", ndump0, rec)
                    irec = 59 + (7 * nit) // s.nstep0
                    for i, line in enumerate(dio.lis_dump_block(nit, timt, timc, umach, dt)):
                        lis.write(irec - 6 + i, line)
                    self.log(f"dump {ndump0}: step {nit}, DT {dt:.6e}, UMACH {umach:.4e}")
            for i, line in enumerate(dio.lis_trailer()):
                lis.write(irec + 1 + i, line)
            summary.update(dumps=ndump0, dt=dt, umach=umach, timt=timt)
            return summary
        finally:
            close = getattr(stepper, "close", None)
            if close:
                close()


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--deck", required=True, help="JSON with params, par, ipar (and fname)")
    ap.add_argument("--out", required=True, help="run directory (restart files are read from here)")
    ap.add_argument("--precision", default="fp64", choices=("fp64", "mixed"))
    ap.add_argument("--ntotal", type=int)
    ap.add_argument("--nstep0", type=int)
    ap.add_argument("--nstart", type=int)
    a = ap.parse_args()
    r = Run(Deck.from_json(a.deck), a.out, a.precision, a.ntotal, a.nstep0, a.nstart,
            log=lambda *x: print(*x, flush=True)).run()
    print(json.dumps(r))


if __name__ == "__main__":
    main()
