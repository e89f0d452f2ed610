"""The reference's dump, restart and log files, written and read on the host.

The reference main program (3dmhd.f) writes three kinds of file, and reads
the first two back on a restart (3dmhd.f:289-363):

* FOUT.dat0.NNNN, one per MPI rank NNNN (PEEXTN, 3dmhdsub.f:2960-2978):
  unformatted direct access, RECL = IWORD*NPX*NRY*NRZ bytes, record
  8*n+1..8*n+8 (5*n+1..5*n+5 without LMAG) of dump n+1 holding the rank's
  interior block of RU, RV, RW, TT, RO[, BX, BY, BZ] in that order, x
  fastest (3dmhd.f:926-959).  Records are contiguous with no markers.
* FOUT.par: rank 0, RECL = IWORD*96, record n+1 = PAR1(96): IPAR(1:32) as
  reals, then PAR(1:64), with PAR1(15) = NIT, PAR1(63) = DT and PAR1(64),
  PAR1(65) = TIMI, TIMC (TIMT, 0 on the last step) (3dmhd.f:914-922,
  961-965).
* FOUT.lis: formatted direct access, 80-character records, each a line of
  text padded with blanks to 79 characters and closed by a newline
  (3dmhd.f:526-887 header, 967-1007 one block per dump, 1019-1050 trailer,
  1124-1150 shutdown).

Rank r owns y-slab MYPEY = r mod NPEY and z-slab MYPEZ = r div NPEY
(3dmhd.f:87-88) of NRY = NPY/NPEY rows and NRZ = NPZ/NPEZ planes.  This
module maps the single resident domain of this build (padded
(NPX+2, NPY+2, NPZ+2) arrays, Fortran order) to and from those rank files,
so that runs of the GPU step write what the reference writes and restart
from what the reference wrote.
"""

from __future__ import annotations

import math
import os
from dataclasses import dataclass

import numpy as np

DUMP_FIELDS = ("ru", "rv", "rw", "tt", "ro", "bx", "by", "bz")  # record order of a dump
IWORD = 8
LIS_RECL = 80


def pe_string(rank: int) -> str:
    """PEEXTN: the rank written with I4 and its blanks replaced by zeros."""
    s = f"{rank:4d}"
    return "****" if len(s) > 4 else s.replace(" ", "0")


@dataclass(frozen=True)
class Layout:
    """The reference's 2-D decomposition of an NPX x NPY x NPZ box."""
    npx: int
    npy: int
    npz: int
    npey: int
    npez: int

    def __post_init__(self):
        if self.npy % self.npey or self.npz % self.npez:
            raise ValueError(f"NPY={self.npy} / NPEY={self.npey} or NPZ={self.npz} / NPEZ={self.npez} not whole")

    @property
    def nry(self) -> int:
        return self.npy // self.npey

    @property
    def nrz(self) -> int:
        return self.npz // self.npez

    @property
    def npe(self) -> int:
        return self.npey * self.npez

    def slab(self, rank: int):
        """Global padded index ranges (0-based, half-open) of the rank's
        interior block: x all interior columns, its NRY rows, its NRZ planes."""
        py, pz = rank % self.npey, rank // self.npey
        return (slice(1, self.npx + 1), slice(1 + py * self.nry, 1 + (py + 1) * self.nry),
                slice(1 + pz * self.nrz, 1 + (pz + 1) * self.nrz))

    @property
    def recl(self) -> int:
        return IWORD * self.npx * self.nry * self.nrz


def _fields(lmag: bool):
    return DUMP_FIELDS if lmag else DUMP_FIELDS[:5]


def dat0_path(prefix: str, rank: int) -> str:
    return f"{prefix}.dat0.{pe_string(rank)}"


def write_dump(prefix: str, dump: int, state: dict, lay: Layout, lmag: bool) -> None:
    """Dump number `dump` (1-based; records of dump n start at nf*(n-1)+1)
    of a padded single-domain state into every rank's .dat0 file, creating
    the files if needed and leaving other records untouched, as the
    reference's direct-access WRITEs do."""
    names = _fields(lmag)
    nf = len(names)
    for r in range(lay.npe):
        sl = lay.slab(r)
        path = dat0_path(prefix, r)
        with open(path, "r+b" if os.path.exists(path) else "w+b") as f:
            for i, k in enumerate(names):
                block = np.asarray(state[k][sl], dtype=np.float64)
                f.seek((nf * (dump - 1) + i) * lay.recl)
                f.write(block.tobytes(order="F"))


def read_dump(prefix: str, dump: int, lay: Layout, lmag: bool) -> dict:
    """Dump `dump` of the rank files, stitched into padded single-domain
    arrays: interior from the records, x/y ghosts by periodic copy
    (COMMUNICATE), z ghost planes zero (the reference's restart does not set
    them; 3dmhd.f:334-350)."""
    names = _fields(lmag)
    nf = len(names)
    shape = (lay.npx + 2, lay.npy + 2, lay.npz + 2)
    out = {k: np.zeros(shape, order="F") for k in DUMP_FIELDS}
    blk = (lay.npx, lay.nry, lay.nrz)
    for r in range(lay.npe):
        sl = lay.slab(r)
        with open(dat0_path(prefix, r), "rb") as f:
            for i, k in enumerate(names):
                f.seek((nf * (dump - 1) + i) * lay.recl)
                raw = f.read(lay.recl)
                if len(raw) != lay.recl:
                    raise ValueError(f"{dat0_path(prefix, r)}: dump {dump} record {k} is short")
                out[k][sl] = np.frombuffer(raw, dtype=np.float64).reshape(blk, order="F")
    for k in DUMP_FIELDS:
        periodic_fill(out[k])
    return out


def periodic_fill(a: np.ndarray) -> np.ndarray:
    """COMMUNICATE's periodic ghost copies of one resident domain: y first,
    then x, over every plane (3dmhdsub.f:3131-3138)."""
    a[:, 0, :] = a[:, -2, :]
    a[:, -1, :] = a[:, 1, :]
    a[0, :, :] = a[-2, :, :]
    a[-1, :, :] = a[1, :, :]
    return a


def count_dumps(prefix: str, lay: Layout, lmag: bool) -> int:
    """Complete dumps held by rank 0's .dat0 file."""
    size = os.path.getsize(dat0_path(prefix, 0))
    return size // (len(_fields(lmag)) * lay.recl)


# ---- .par -----------------------------------------------------------------

def par1_record(ipar, par, nit: int, dt: float, t64: float, t65: float) -> np.ndarray:
    """PAR1(96): IPAR(1:32) as reals, PAR(1:64), then the step's NIT, DT and
    times in slots 15, 63, 64, 65 (1-based)."""
    rec = np.zeros(96)
    rec[:32] = np.asarray(ipar, dtype=np.float64)[:32]
    rec[32:] = np.asarray(par, dtype=np.float64)[:64]
    rec[14] = float(nit)
    rec[62] = dt
    rec[63] = t64
    rec[64] = t65
    return rec


def write_par(path: str, recno: int, rec: np.ndarray) -> None:
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    if rec.shape != (96,):
        raise ValueError("a .par record holds 96 values")
    with open(path, "r+b" if os.path.exists(path) else "w+b") as f:
        f.seek((recno - 1) * IWORD * 96)
        f.write(rec.tobytes())


def read_par(path: str, recno: int) -> np.ndarray:
    with open(path, "rb") as f:
        f.seek((recno - 1) * IWORD * 96)
        raw = f.read(IWORD * 96)
    if len(raw) != IWORD * 96:
        raise ValueError(f"{path}: record {recno} missing")
    return np.frombuffer(raw, dtype=np.float64).copy()


def par_layout(par1: np.ndarray) -> Layout:
    """Box and decomposition recorded in a .par record: NPX, NPY, NPZ
    (slots 11-13), NPE (14) and NPEY (21), as the restart reads them
    (3dmhd.f:309-311)."""
    npx, npy, npz = (int(round(par1[i])) for i in (10, 11, 12))
    npe, npey = int(round(par1[13])), int(round(par1[20]))
    return Layout(npx, npy, npz, npey, npe // npey)


# ---- .lis -----------------------------------------------------------------

def fmt_i(v: int, w: int) -> str:
    """Fortran Iw."""
    s = f"{int(v):d}"
    return "*" * w if len(s) > w else s.rjust(w)


def fmt_1pd(x: float, w: int, d: int) -> str:
    """Fortran 1PDw.d: one digit before the point, d after, exponent letter
    D with a sign and two digits (the letter is dropped for exponents of
    three digits); asterisks when the field is too narrow."""
    x = float(x)
    if math.isnan(x):
        s = "NaN"
    elif math.isinf(x):
        s = ("Infinity" if w >= 8 + (x < 0) else "Inf")
        s = ("-" if x < 0 else "") + s
    else:
        m, e = f"{x:.{d}E}".split("E")
        e = int(e)
        s = f"{m}D{e:+03d}" if abs(e) <= 99 else f"{m}{e:+04d}"
    return "*" * w if len(s) > w else s.rjust(w)


def fmt_a(text: str, w: int) -> str:
    """Fortran Aw of a CHARACTER variable: its first w characters (the
    variable is blank-padded to its declared length)."""
    return text[:w].ljust(w)


class LisFile:
    """The 80-character-record log, written record by record at given record
    numbers as the reference's direct-access WRITEs do."""

    def __init__(self, path: str):
        self.path = path

    def write(self, recno: int, text: str) -> None:
        if len(text) > LIS_RECL - 1:
            raise ValueError(f"log line longer than 79 characters: {text!r}")
        rec = (text + " " * (LIS_RECL - 1 - len(text)) + "\n").encode("latin-1")
        with open(self.path, "r+b" if os.path.exists(self.path) else "w+b") as f:
            f.seek((recno - 1) * LIS_RECL)
            f.write(rec)


def read_lis(path: str) -> list:
    """Records of a log as text lines (blank padding and newline removed)."""
    raw = open(path, "rb").read()
    return [raw[i:i + LIS_RECL - 1].decode("latin-1").rstrip(" ") for i in range(0, len(raw), LIS_RECL)]


RULE = "-" * 78
BLANK = ""


def lis_header(*, ncase: int, ncasep: int, nstart: int, npx: int, npy: int, npz: int, lrot: bool, lmag: bool,
               izc: int, itc: int, par: dict, sf: float, ampt: float, ampb: float, finp: str, fout: str,
               timi: float, restart_size, ntotal: int, nstep0: int, dts, dt: float, nsound: int) -> list:
    """The header lines of a run's log, records 1.. in order (3dmhd.f:526-887,
    formats 4000-4052).  par: RE, PR, THETA, GRAV, RY, ANG, RM, BETA, PZP,
    SIGMA, POLYS, TP, XP, YP, ZP, XMAX, YMAX, ZMAX by name; dts: DT1..DT4;
    restart_size: (NX1, NY1*NPEY, NZ1*NPEZ) of the restart record."""
    L = [RULE, BLANK, BLANK, "     THREE-D MAGNETOHYDRODYNAMICS", "     Case #" + fmt_i(ncase, 3)]
    if nstart == 0:
        L.append("     Static start")
    elif ncase == ncasep:
        L.append("     Continuation")
    else:
        L.append("     Previous case #" + fmt_i(ncasep, 3))
    L += [BLANK, BLANK, RULE, BLANK,
          "     Array size: NX=" + fmt_i(npx, 4) + " NY=" + fmt_i(npy, 4) + " NZ=" + fmt_i(npz, 4), BLANK,
          "     Rotating domain." if lrot else "     Nonrotating domain.",
          "     Magnetic fields included." if lmag else "     No magnetic fields.",
          "     Horizontally periodic.",
          "     Stress free and impenetrable lower boundary."]
    if izc == 0:
        L.append("     Constant temperature lower boundary.")
    elif izc == 1:
        L.append("     Constant thermal flux lower boundary.")
    L.append({0: "     Constant temperature upper boundary.",
              1: "     Constant thermal flux upper boundary.",
              2: "     Embedded heat loss.",
              3: "     Embedded thermal.",
              4: "     Differential upper boundary temperature."}.get(itc, "     No temperature perturbation."))
    L += [BLANK, "     Parameters for this run:"]
    for label, key in (("Re   ", "RE"), ("Pr   ", "PR"), ("Theta", "THETA"), ("Grav ", "GRAV"), ("Ro   ", "RY"),
                       ("Angle", "ANG"), ("Rm   ", "RM"), ("Beta ", "BETA"), ("Zpen ", "PZP"), ("Sigma", "SIGMA"),
                       ("ms   ", "POLYS"), ("Tp   ", "TP"), ("Xp   ", "XP"), ("Yp   ", "YP"), ("Zp   ", "ZP"),
                       ("Xmax ", "XMAX"), ("Ymax ", "YMAX"), ("Zmax ", "ZMAX")):
        L.append("          " + label + " = " + fmt_1pd(par[key], 11, 4))
    L += [BLANK, "     Time stepping safety factor:", "          SFF = " + fmt_1pd(sf, 10, 4), BLANK]
    if nstart == 0:
        L += ["     Calculation started from a static state.",
              "     Initial temperature perturbations: " + fmt_1pd(ampt, 10, 4),
              ("     Initial magnetic layer amplitude: " + fmt_1pd(ampb, 10, 4)) if lmag
              else "     No magnetic fields."]
    else:
        nx1, ny1, nz1 = restart_size
        L += ["     Started from dump " + fmt_i(nstart, 2) + " of file " + fmt_a(finp, 20),
              "     Previous simulation time: " + fmt_1pd(timi, 10, 4),
              "     Previous size: NX=" + fmt_i(nx1, 4) + " NY=" + fmt_i(ny1, 4) + " NZ=" + fmt_i(nz1, 4)]
    L += [BLANK, "     The results are stored in file " + fmt_a(fout, 15),
          "     Total number of iterations required = " + fmt_i(ntotal, 8),
          "     Variables dumped every " + fmt_i(nstep0, 6) + " iterations.", BLANK,
          "     Advected fast-mode time: " + fmt_1pd(dts[0], 10, 4),
          "     Thermal diffusion time:  " + fmt_1pd(dts[1], 10, 4),
          "     Viscous diffusion time:  " + fmt_1pd(dts[2], 10, 4),
          "     Magnetic diffusion time: " + fmt_1pd(dts[3], 10, 4),
          "     Initial time-step value: " + fmt_1pd(dt, 10, 4),
          "     Iterations per crossing time: " + fmt_i(nsound, 8), BLANK, RULE, RULE]
    return L


def lis_dump_block(nit: int, timt: float, timc: float, umach: float, dt: float) -> list:
    """The seven lines a dump appends (3dmhd.f:967-1007), for records
    IREC-6..IREC with IREC = 59 + 7*NIT/NSTEP0."""
    return [BLANK, "     Iteration " + fmt_i(nit, 7) + " completed", "     " + "-" * 35,
            "     Total simulation time:   " + fmt_1pd(timt, 10, 4),
            "     Present simulation time: " + fmt_1pd(timc, 10, 4),
            "     Maximum Mach number:     " + fmt_1pd(umach, 10, 4),
            "     Present time step:       " + fmt_1pd(dt, 10, 4)]


def lis_trailer() -> list:
    """Records IREC+1..IREC+4 after the last step (3dmhd.f:1027-1048)."""
    return [RULE, "               SIMULATION COMPLETE", RULE, RULE]


def lis_shutdown(dt: float, nit: int) -> list:
    """Records IREC+1..IREC+4 of the DT < 1e-8 shutdown, IREC =
    58 + 7*NIT/NSTEP0 (3dmhd.f:1126-1146, format 4100)."""
    return [RULE, "     Time step " + fmt_1pd(dt, 11, 4) + "  Shut down at step " + fmt_i(nit, 8), RULE, RULE]
