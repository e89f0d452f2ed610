"""Host-side set-up of one resident box: the COMMON scalars and the metric
arrays that the step reads, computed the way the reference driver does.

* constants(): the parameter mapping of the main program (3dmhd.f:105-240,
  378-379).  Default-REAL literals are folded in float32 and widened, exactly
  as the Fortran compiler does (C13 = 1.0E00/3.0E00, GAM1 = 8.0E00/15.0E00,
  GAMMA = 5.00E00/3.00E00, HX = 0.5E00*FLOAT(NPX-1), ...).
* metrics(): the mapped-grid Jacobian of MKGRID / X,Y,ZJACOBI
  (3dmhdsub.f:17-353) for one resident domain: uniform (NGRID=0), arctan
  stretch (NGRID=1), plus a tanh stretch in z (NGRID=3, no counterpart in
  the reference) whose ratio of largest to smallest spacing is a parameter.
* physics(): a mapping from (density contrast, Ra, Pr, Pm, Ta, tilt) of a
  polytropic layer onto the reference's (THETA, GRAV, RE, PR, RM, RY, ANG,
  BETA) inputs, defined at mid-depth.
"""

from __future__ import annotations

import math

import numpy as np


def f32(x) -> float:
    """Value of a default-REAL literal or expression, widened to double."""
    return float(np.float32(x))


def f32div(a, b) -> float:
    return float(np.float32(a) / np.float32(b))


def constants(npx: int, npy: int, npz: int, par: dict, *, lmag: bool, lrot: bool,
              sff: float = 0.315, izcon: int = 1, itcon: int = 0, ibcon: int = 0,
              tb: float = 1.0) -> dict:
    """COMMON scalars of the step.  `par` maps PAR slot numbers (3dmhdset.f
    meanings) to float values that are already float32-representable or to
    literal strings; missing slots are 0.  TB is the bottom temperature of
    the static profile (STATIC, sub:520)."""
    def P(n):
        v = par.get(n, 0.0)
        if isinstance(v, str):
            s = v.upper().replace("D", "E")
            if "/" in s:
                a, b = s.split("/")
                return f32div(float(a), float(b))
            return f32(float(s))
        return f32(v)

    c13 = f32div(1.0, 3.0)
    m = dict(npx=npx, npy=npy, npz=npz, nx=npx + 2, ny=npy + 2, nz=npz + 2,
             lmag=int(lmag), lrot=int(lrot), izc=izcon, itc=itcon, ibc=ibcon)
    m["c13"] = c13
    m["c23"] = 2.0 * c13
    m["c43"] = 4.0 * c13
    m["gam1"] = f32div(8.0, 15.0)
    m["gam2"] = f32div(5.0, 12.0)
    m["gam3"] = f32div(3.0, 4.0)
    m["zeta1"] = -f32div(17.0, 60.0)
    m["zeta2"] = -f32div(5.0, 12.0)
    for ax, n in (("x", npx), ("y", npy), ("z", npz)):
        m[f"h{ax}"] = float(np.float32(0.5) * np.float32(n - 1))
        m[f"h2{ax}"] = float(np.float32(n - 1) * np.float32(n - 1))
    m["sf"] = f32(sff)
    re, pr, theta, grav = P(1), P(2), P(3), P(4)
    ry, ang, rm, beta = P(5), P(6), P(7), P(8)
    gamma = P(54) if 54 in par else f32div(5.0, 3.0)
    m.update(re=re, theta=theta, grav=grav, rm=rm, gamma=gamma, tp=P(12), ampt=P(20))
    if lrot:
        m["omx"] = math.sin(ang) / ry
        m["omz"] = math.cos(ang) / ry
    else:
        m["omx"] = m["omz"] = 0.0
    m["repr"] = re * pr
    m["ore"] = 1.0 / re
    if lmag:
        m["orm"] = 1.0 / rm
        m["obeta"] = 2.0 / beta
    else:
        m["orm"] = m["obeta"] = 0.0
    m["cv"] = 1.0 / (gamma - 1.0)
    m["ocv"] = 1.0 / m["cv"]
    m["tb"] = float(tb)
    m["dzf"] = f32div(1.0, npz - 1)
    return m


def _mkgrid(ngrid: int, scode: np.ndarray, smax: float, s1: float, s2: float):
    """MKGRID (3dmhdsub.f:17-175) for NGRID 0 and 1, vectorised."""
    scode = np.asarray(scode, dtype=np.float64)
    if ngrid == 0:
        return scode * smax, np.full_like(scode, 1.0 / smax), np.zeros_like(scode)
    if ngrid != 1:
        raise ValueError("ngrid must be 0 or 1 here")
    a1 = (s2 - s1) / smax
    a3 = math.atan(s1)
    a2 = math.atan(s2) - a3
    sphys = np.where(scode == 0.0, 0.0,
                     np.where(scode == 1.0, smax, (np.tan(a2 * scode + a3) - s1) / a1))
    x = s1 + a1 * sphys
    dssds = a1 / a2 / (1.0 + x ** 2)
    d2 = -(2.0 * a1 ** 2 * x / a2 / (1.0 + x ** 2) ** 2)
    return sphys, dssds, d2


def _tanh_grid(scode: np.ndarray, smax: float, ratio: float):
    """z(s) = smax * [s + c*(tanh(a(s-1/2)) ... )] is replaced here by the
    closed form of a spacing that varies as 1/cosh^2: dz/ds = smax*(1 +
    (ratio-1)*(1 - sech^2(a*(s-1))))/norm, finest at the top (s=0), coarsest
    at the bottom, with max/min spacing equal to `ratio`.  Returns
    (z, ds/dz, d2s/dz2)."""
    a = 3.0
    s = np.asarray(scode, dtype=np.float64)

    def g(t):  # spacing shape, 1 at s=0 ... ratio at s=1
        return 1.0 + (ratio - 1.0) * (np.tanh(a * t) / np.tanh(a))

    # integral of g from 0 to t
    def G(t):
        return t + (ratio - 1.0) * np.log(np.cosh(a * t)) / (a * np.tanh(a))

    norm = G(1.0)
    z = smax * G(s) / norm
    dzds = smax * g(s) / norm
    d2zds2 = smax * (ratio - 1.0) * a * (1.0 / np.cosh(a * s)) ** 2 / np.tanh(a) / norm
    dsdz = 1.0 / dzds
    d2sdz2 = -d2zds2 * dsdz ** 3
    return z, dsdz, d2sdz2


def metrics(npx: int, npy: int, npz: int, *, xmax: float, ymax: float, zmax: float,
            ngrid: int = 1, xx=(-1.0e-9, 1.0e-9), yy=(-1.0e-9, 1.0e-9), zz=(0.0, math.sqrt(2.0)),
            zratio: float = 3.0) -> dict:
    """Padded metric arrays of /AJACOBI/ for one resident domain
    (XJACOBI sub:177-209, YJACOBI sub:211-281, ZJACOBI sub:283-353).
    ngrid=3 uses the arctan map in x,y and a tanh stretch in z."""
    out = {}
    for ax, n, smax, (s1, s2) in (("x", npx, xmax, xx), ("y", npy, ymax, yy), ("z", npz, zmax, zz)):
        orr = f32div(1.0, n - 1)
        sc = np.arange(n, dtype=np.float64)
        sc = np.float32(sc).astype(np.float64) * orr
        sc[0], sc[-1] = 0.0, 1.0
        if ngrid == 3 and ax == "z":
            phys, d1, d2 = _tanh_grid(sc, smax, zratio)
        else:
            phys, d1, d2 = _mkgrid(1 if ngrid == 3 else ngrid, sc, smax, s1, s2)
        dd = orr * (1.0 / d1)
        pad = lambda a, lo, hi: np.concatenate([[lo], a, [hi]])
        if ax == "x":
            # XJACOBI leaves the x ghost entries unset (zero)
            vals = [pad(phys, 0.0, 0.0), pad(d1, 0.0, 0.0), pad(d2, 0.0, 0.0), pad(dd, 0.0, 0.0)]
        else:
            # Y/ZJACOBI evaluate the edge ghosts at the domain ends
            vals = [pad(phys, phys[0], phys[-1]), pad(d1, d1[0], d1[-1]),
                    pad(d2, d2[0], d2[-1]), pad(dd, dd[0], dd[-1])]
        names = {"x": ("exx", "dxxdx", "d2xxdx2", "ddx"), "y": ("wyy", "dyydy", "d2yydy2", "ddy"),
                 "z": ("zee", "dzzdz", "d2zzdz2", "ddz")}[ax]
        for nm, v in zip(names, vals):
            out[nm] = np.ascontiguousarray(v)
    out["rkapa"] = np.ones(npz + 2)
    out["dkapa"] = np.zeros(npz + 2)
    return out


def polytrope(zee: np.ndarray, theta: float, grav: float):
    """Closed-form polytrope T = 1 + THETA z, rho = T**(GRAV/THETA - 1) on a
    padded z column, ghost planes repeating the boundary values as STATIC
    sets them (sub:525-546)."""
    zee = np.asarray(zee, dtype=np.float64)
    t = 1.0 + theta * zee
    r = t ** (grav / theta - 1.0)
    t[0], r[0] = t[1], r[1]
    t[-1], r[-1] = t[-2], r[-2]
    return t, r


def physics(contrast: float, ra: float, pr: float, pm: float, ta: float, tilt_deg: float,
            m: float = 1.0, gamma: float = 5.0 / 3.0, beta: float = 2.0) -> dict:
    """Reference inputs for a polytropic layer of depth 1 with polytropic
    index m and top temperature 1:  THETA = contrast**(1/m) - 1,
    GRAV = (m+1) THETA.  Dimensionless numbers are taken at mid-depth
    (T_mid = 1 + THETA/2, rho_mid = T_mid**m), with kinematic viscosity
    nu = 1/(RE rho_mid), thermal diffusivity kappa = 1/(RE PR rho_mid c_p),
    c_p = gamma/(gamma-1) (so Pr = nu/kappa = c_p PR), magnetic diffusivity
    eta = 1/RM and 2 Omega = 1/RY:
        Ra = GRAV (THETA - THETA_ad) / (nu kappa T_mid),  THETA_ad = GRAV/c_p
        Ta = (2 Omega / nu)**2,  Pm = nu/eta.
    Returns PAR slot values (floats)."""
    theta = contrast ** (1.0 / m) - 1.0
    grav = (m + 1.0) * theta
    cp = gamma / (gamma - 1.0)
    tmid = 1.0 + 0.5 * theta
    rmid = tmid ** m
    dtheta = theta - grav / cp
    if dtheta <= 0:
        raise ValueError("layer is not superadiabatic")
    nu = math.sqrt(grav * dtheta * pr / (ra * tmid))
    kappa = nu / pr
    re = 1.0 / (nu * rmid)
    prp = 1.0 / (re * rmid * cp * kappa)
    rm = pm / nu
    two_omega = math.sqrt(ta) * nu
    ry = 1.0 / two_omega if ta > 0 else 0.0
    return {1: re, 2: prp, 3: theta, 4: grav, 5: ry, 6: math.radians(tilt_deg), 7: rm, 8: beta,
            17: 4.0, 18: 4.0, 19: 1.0}
