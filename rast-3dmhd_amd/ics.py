"""Host-side initial conditions for a single resident box.

The reference builds its initial state once, on the host, in STATIC
(3dmhdsub.f:355-657): a polytrope T(z), rho(z) integrated by Bulirsch-Stoer,
sliced into the padded arrays, an optional Gaussian Bx layer with the
matching density deficit, and a multiplicative random T perturbation drawn
from RAN2 (3dmhdsub.f:3377-3410) in an order that does not depend on the
processor layout.  This module restates the slicing, the B layer and the
RAN2 perturbation for one resident domain, taking the 1-D profiles as
inputs, plus the periodic ghost fill of COMMUNICATE (3dmhdsub.f:3131-3138).

For boxes where the exact RAN2 sequence is not needed (throughput runs),
`polytrope_profiles` gives the closed-form m-polytrope that STATIC
integrates numerically when PZP=0.
"""

from __future__ import annotations

import numpy as np

FIELDS = ("ru", "rv", "rw", "ro", "tt", "bx", "by", "bz")

# RAN2 constants (3dmhdsub.f:3385).  RM is a default-REAL PARAMETER, so it is
# the float32 value of 1/714025 widened to double.
_M, _IA, _IC = 714025, 1366, 150889
_RM = float(np.float32(1.0) / np.float32(714025.0))


class Ran2:
    """Numerical-Recipes style LCG with a 97-slot shuffle, as RAN2."""

    def __init__(self, idum: int = -62659):
        self.idum = idum
        self.iiy = 0
        self.iir = [0] * 97

    def __call__(self) -> float:
        if self.idum < 0:
            self.idum = (_IC - self.idum) % _M
            for j in range(97):
                self.idum = (_IA * self.idum + _IC) % _M
                self.iir[j] = self.idum
            self.idum = (_IA * self.idum + _IC) % _M
            self.iiy = self.idum
        j = 1 + (97 * self.iiy) // _M
        if j > 97 or j < 1:
            raise RuntimeError("RAN2: index out of range")
        self.iiy = self.iir[j - 1]
        out = self.iiy * _RM
        self.idum = (_IA * self.idum + _IC) % _M
        self.iir[j - 1] = self.idum
        return out


def periodic_fill(a: np.ndarray) -> np.ndarray:
    """COMMUNICATE's in-domain periodic copies: y first, then x, over all z."""
    a[:, 0, :] = a[:, -2, :]
    a[:, -1, :] = a[:, 1, :]
    a[0, :, :] = a[-2, :, :]
    a[-1, :, :] = a[1, :, :]
    return a


def static_state(npx: int, npy: int, npz: int, col_t, col_r, col_bx=None,
                 ampt: float = 0.0, obeta: float = 0.0, lmag: bool = False,
                 seed: int = -62659, fast_noise: bool = False) -> dict:
    """Padded (npx+2, npy+2, npz+2) Fortran-order state as STATIC leaves it
    for one resident domain, before the first ghost fill.

    col_t, col_r: padded columns (length npz+2) of T and rho, ghost planes
    included (top ghost = T(1), bottom ghost = T(NPZ) as in sub:525-546).
    col_bx: padded column of the Bx layer (sub:572-574), or None.
    fast_noise: draw the T perturbation from numpy instead of RAN2 (for
    throughput boxes where the sequence need not match the reference).
    """
    nx, ny, nz = npx + 2, npy + 2, npz + 2
    col_t = np.asarray(col_t, dtype=np.float64)
    col_r = np.asarray(col_r, dtype=np.float64)
    if col_t.shape != (nz,) or col_r.shape != (nz,):
        raise ValueError("profiles must have length npz+2")
    st = {k: np.zeros((nx, ny, nz), order="F") for k in FIELDS}
    st["tt"][1:-1, 1:-1, :] = col_t[None, None, :]
    st["ro"][1:-1, 1:-1, :] = col_r[None, None, :]
    if lmag and col_bx is not None:
        col_bx = np.asarray(col_bx, dtype=np.float64)
        st["bx"][:, :, 1:-1] = col_bx[None, None, 1:-1]
        bx = st["bx"][1:-1, 1:-1, 1:-1]
        tt = st["tt"][1:-1, 1:-1, 1:-1]
        st["ro"][1:-1, 1:-1, 1:-1] = st["ro"][1:-1, 1:-1, 1:-1] - bx * bx / tt * obeta
    if ampt != 0.0:
        if fast_noise:
            rng = np.random.default_rng(abs(seed))
            rnd = 1.0 + ampt * (rng.random((npz, npy, npx)) - 0.5)
        else:
            r = Ran2(seed)
            n = npx * npy * npz
            draws = np.fromiter((r() for _ in range(n)), dtype=np.float64, count=n)
            rnd = 1.0 + ampt * (draws.reshape(npz, npy, npx) - 0.5)
        st["tt"][1:-1, 1:-1, 1:-1] = st["tt"][1:-1, 1:-1, 1:-1] * rnd.transpose(2, 1, 0)
    return st


def vector_potential_field(meta: dict, kmax: int, amp: float, seed: int):
    """Seed field B = curl A from a random vector potential with integer
    wavenumbers |k| <= kmax (horizontal Fourier modes of the periodic box,
    vertical modes sin(pi m z) for Ax, Ay and cos(pi m z) for Az, m >= 1).
    The curl uses the step's own centred, metric-weighted differences, so the
    centred divergence of B vanishes in the interior to round-off.  BZ is set
    to zero on the boundary planes (the IBCON=0 condition).  Scaled so that
    max|B| = amp.  Returns padded (bx, by, bz) with periodic ghosts."""
    npx, npy, npz = int(meta["npx"]), int(meta["npy"]), int(meta["npz"])
    rng = np.random.default_rng(seed)
    zee = np.asarray(meta["zee"], dtype=np.float64)
    zl = zee[-2] - zee[1]
    zc = np.empty(npz + 2)
    zc[1:-1] = (zee[1:-1] - zee[1]) / zl
    zc[0] = 2 * zc[1] - zc[2]
    zc[-1] = 2 * zc[-2] - zc[-3]
    kx = np.fft.fftfreq(npx, 1.0 / npx)[:, None]
    ky = np.fft.rfftfreq(npy, 1.0 / npy)[None, :]
    A = [np.zeros((npx, npy, npz + 2)) for _ in range(3)]
    for m in range(1, kmax + 1):
        sel = (kx ** 2 + ky ** 2 + m ** 2) <= kmax ** 2
        if not sel.any():
            continue
        for c in range(3):
            coef = (rng.standard_normal(sel.shape) + 1j * rng.standard_normal(sel.shape)) * sel
            plane = np.fft.irfft2(coef, s=(npx, npy)) * (npx * npy)
            prof = np.cos(np.pi * m * zc) if c == 2 else np.sin(np.pi * m * zc)
            A[c] += plane[:, :, None] * prof[None, None, :]
    hx, hy, hz = float(meta["hx"]), float(meta["hy"]), float(meta["hz"])
    dx = np.asarray(meta["dxxdx"])[1:-1][:, None, None]
    dy = np.asarray(meta["dyydy"])[1:-1][None, :, None]
    dz = np.asarray(meta["dzzdz"])[1:-1][None, None, :]

    def ddx(f):
        return (np.roll(f, -1, 0) - np.roll(f, 1, 0))[:, :, 1:-1] * hx * dx

    def ddy(f):
        return (np.roll(f, -1, 1) - np.roll(f, 1, 1))[:, :, 1:-1] * hy * dy

    def ddz(f):
        return (f[:, :, 2:] - f[:, :, :-2]) * hz * dz

    bx = ddy(A[2]) - ddz(A[1])
    by = ddz(A[0]) - ddx(A[2])
    bz = ddx(A[1]) - ddy(A[0])
    bz[:, :, 0] = 0.0
    bz[:, :, -1] = 0.0
    scale = amp / max(np.abs(bx).max(), np.abs(by).max(), np.abs(bz).max(), 1e-300)
    out = []
    for b in (bx, by, bz):
        p = np.zeros((npx + 2, npy + 2, npz + 2), order="F")
        p[1:-1, 1:-1, 1:-1] = b * scale
        out.append(periodic_fill(p))
    return tuple(out)


def polytrope_profiles(zee: np.ndarray, theta: float, grav: float):
    """Closed-form polytrope T = 1 + theta z, rho = T**(grav/theta - 1) on a
    padded z column (ghost planes repeat the boundary values).  STATIC
    integrates the same hydrostatic ODE numerically (sub:485-508) when the
    conductivity is uniform (PZP=0), so the two agree to its tolerance."""
    zee = np.asarray(zee, dtype=np.float64)
    t = 1.0 + theta * zee
    r = t ** (grav / theta - 1.0)
    t[0], r[0] = t[1], r[1]
    t[-1], r[-1] = t[-2], r[-2]
    return t, r
