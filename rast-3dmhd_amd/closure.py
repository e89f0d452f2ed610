"""Learned subgrid closure: training data from box-filtered plain (FP64)
fine runs, training of the BF16-MFMA network on the GPU, and closure-
corrected coarse runs (mixed form) judged against the filtered fine run.

Grids.  The fine box has factor*nxc x factor*nyc x factor*(nzc-1)+1 points
and the coarse box nxc x nyc x nzc, with horizontal sizes chosen so that the
coarse spacing is factor fine spacings and both boxes have the same
periods, and the same tanh-stretched vertical map (NGRID=3): coarse plane K
is then exactly fine plane factor*K.  The filter (mhd_box_filter, zmode 1)
averages factor x factor horizontal blocks and the factor vertical spacings
around each coarse plane.

Target.  For a filtered fine state q_f = F(q): the subgrid term of each
equation is S = F(RHS_fine(q)) - RHS_coarse(q_f) for RU, RV, RW, TT, BX, BY,
BZ; the network maps (rho, u, v, w, s, BX, BY, BZ) of q_f to S, both
normalised per channel.

Judge.  The mean convective flux F_conv(z) = -<rho w (2.5 T' + |u|^2/2)>
(horizontal average, T' the deviation from the horizontal mean; the
reference's FCONM, 3dmhdsub.f:1650-1661) and the horizontal kinetic-energy
spectrum of the coarse runs against the filtered fine run after the same
simulated time.  The closure is attached only if its error on held-out
samples is at most `threshold`; otherwise the coarse run is the plain one.
"""

from __future__ import annotations

import importlib
import time

import numpy as np

FIELDS = ("ru", "rv", "rw", "ro", "tt", "bx", "by", "bz")
RHS_OF_TARGET = (0, 1, 2, 4, 5, 6, 7)   # FU, FV, FW, FT, WW1..WW3 in the RHS order
TARGETS = ("ru", "rv", "rw", "tt", "bx", "by", "bz")
INPUTS = ("rho", "u", "v", "w", "s", "bx", "by", "bz")


def _mods():
    return (importlib.import_module("rast-3dmhd_amd"), importlib.import_module("rast-3dmhd_amd.config"),
            importlib.import_module("rast-3dmhd_amd.ics"))


def pair(nxc, nyc, nzc, factor=4, contrast=10.0, ra=1e5, pr=1.0, pm=1.0, ta=1e4, ampt=1e-2, ampb=1e-5, kmax=4,
         seed=77):
    """(fine meta, fine initial state, coarse meta)."""
    pkg, cfg, ics = _mods()
    ph = cfg.physics(contrast=contrast, ra=ra, pr=pr, pm=pm, ta=ta, tilt_deg=0.0)
    nxf, nyf, nzf = factor * nxc, factor * nyc, factor * (nzc - 1) + 1
    xf, yf = float(ph[17]), float(ph[18])
    xc = xf * factor * (nxc - 1) / (nxf - 1)
    yc = yf * factor * (nyc - 1) / (nyf - 1)

    def meta(nx, ny, nz, xmax, ymax):
        par = {k: cfg.f32(v) for k, v in ph.items()}
        par[17], par[18] = cfg.f32(xmax), cfg.f32(ymax)
        mt = cfg.metrics(nx, ny, nz, xmax=par[17], ymax=par[18], zmax=par[19], ngrid=3)
        t, r = cfg.polytrope(mt["zee"], par[3], par[4])
        m = cfg.constants(nx, ny, nz, par, lmag=True, lrot=True, izcon=1, itcon=0, tb=float(t[-1]))
        m.update(mt)
        return m, t, r

    mf, t, r = meta(nxf, nyf, nzf, xf, yf)
    mc, _, _ = meta(nxc, nyc, nzc, xc, yc)
    st = ics.static_state(nxf, nyf, nzf, t, r, ampt=ampt, lmag=True, seed=seed, fast_noise=True)
    st["bx"], st["by"], st["bz"] = ics.vector_potential_field(mf, kmax=kmax, amp=ampb, seed=seed)
    for k in ("ru", "rv", "rw", "ro", "tt"):
        ics.periodic_fill(st[k])
    return mf, st, mc


def padded(coarse):
    """[8][K][J][I] interior arrays -> dict of padded Fortran-order fields
    (periodic x/y ghosts; z ghosts copy the boundary planes)."""
    _, _, ics = _mods()
    out = {}
    for f, k in enumerate(FIELDS):
        a = coarse[f].transpose(2, 1, 0)  # [I][J][K]
        p = np.zeros((a.shape[0] + 2, a.shape[1] + 2, a.shape[2] + 2), order="F")
        p[1:-1, 1:-1, 1:-1] = a
        p[1:-1, 1:-1, 0] = a[:, :, 0]
        p[1:-1, 1:-1, -1] = a[:, :, -1]
        ics.periodic_fill(p)
        out[k] = p
    return out


def interior(fields):
    """padded dict -> [8][K][J][I] interior arrays."""
    return np.stack([fields[k][1:-1, 1:-1, 1:-1].transpose(2, 1, 0) for k in FIELDS])


def inputs_of(q, gamma):
    """Network inputs [K][J][I][8] of a coarse state [8][K][J][I]."""
    ro, tt = q[3], q[4]
    s = np.log(tt) / (gamma - 1.0) - np.log(ro)
    return np.stack([ro, q[0] / ro, q[1] / ro, q[2] / ro, s, q[5], q[6], q[7]], axis=-1)


def convective_flux(q):
    """-<rho w (2.5 T' + |u|^2/2)> per plane of a state [8][K][J][I]."""
    ro, tt = q[3], q[4]
    u, v, w = q[0] / ro, q[1] / ro, q[2] / ro
    tp = tt - tt.mean(axis=(1, 2), keepdims=True)
    return -(ro * w * (2.5 * tp + 0.5 * (u * u + v * v + w * w))).mean(axis=(1, 2))


def ke_spectrum(q):
    """Horizontal spectrum of sqrt(rho) u, summed over shells of integer
    |k| and averaged over the planes."""
    ro = q[3]
    nz, ny, nx = ro.shape
    kx = np.fft.fftfreq(nx) * nx
    ky = np.fft.fftfreq(ny) * ny
    kk = np.sqrt(kx[None, :] ** 2 + ky[:, None] ** 2)
    shells = np.rint(kk).astype(int)
    nk = min(nx, ny) // 2
    spec = np.zeros(nk)
    for c in range(3):
        a = np.sqrt(ro) * q[c] / ro
        e = (np.abs(np.fft.fft2(a, axes=(1, 2))) ** 2).mean(axis=0) / (nx * ny) ** 2
        spec += np.bincount(shells.ravel(), weights=0.5 * e.ravel(), minlength=nk)[:nk]
    return spec


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


class Experiment:
    """One fine/coarse pair on the current GPU."""

    def __init__(self, nxc, nyc, nzc, factor=4, log=print, **kw):
        self.pkg = _mods()[0]
        self.factor = factor
        self.log = log
        self.mf, self.st0, self.mc = pair(nxc, nyc, nzc, factor=factor, **kw)
        self.gamma = float(self.mf["gamma"])
        self.fine = self.pkg.Box(self.mf)
        self.fine.upload(self.st0)

    # -- data -------------------------------------------------------------
    def develop(self, steps, chunk=500):
        done = 0
        while done < steps:
            n = min(chunk, steps - done)
            dt, um = self.fine.steps(n)
            done += n
            self.log(f"fine: {done}/{steps} steps, dt {dt:.3e}, mach {um:.3e}")
        return dt

    def filtered(self):
        return self.fine.box_filter(self.factor, zmode=1, source="state")

    def sample(self):
        """(inputs [K][J][I][8], targets [K][J][I][7], filtered state)."""
        qf = self.filtered()
        rf = self.fine.box_filter(self.factor, zmode=1, source="rhs")
        with self.pkg.Box(self.mc) as c:
            c.upload(padded(qf))
            fc = c.fluxes()
        rc = np.stack([fc[k][1:-1, 1:-1, 1:-1].transpose(2, 1, 0) for k in
                       ("fu", "fv", "fw", "fr", "ft", "ww1", "ww2", "ww3")])
        s = np.stack([rf[i] - rc[i] for i in RHS_OF_TARGET], axis=-1)
        return inputs_of(qf, self.gamma), s, qf

    def collect(self, nsamples, every):
        X, S, Q = [], [], []
        for i in range(nsamples):
            self.develop(every, chunk=every)
            x, s, q = self.sample()
            X.append(x)
            S.append(s)
            Q.append(q)
            self.log(f"sample {i + 1}/{nsamples}: |S| {np.sqrt((s ** 2).mean(axis=(0, 1, 2)))}")
        return X, S, Q

    # -- training ---------------------------------------------------------
    @staticmethod
    def normalisation(X, S):
        """Per-plane means of the inputs and targets over the samples (the
        horizontally uniform part, e.g. the different discrete hydrostatic
        and diffusive balance of the two grids, is carried by the profile)
        and the standard deviation of what remains, per channel."""
        xs, ss = np.stack(X), np.stack(S)               # [n][K][J][I][c]
        pin = xs.mean(axis=(0, 2, 3))                    # [K][8]
        pout = ss.mean(axis=(0, 2, 3))                   # [K][7]
        sin = (xs - pin[None, :, None, None, :]).std(axis=(0, 1, 2, 3))
        sout = (ss - pout[None, :, None, None, :]).std(axis=(0, 1, 2, 3))
        sin[sin == 0] = 1.0
        sout[sout == 0] = 1.0
        return {"in_profile": pin.T.copy(), "in_std": sin, "out_profile": pout.T.copy(), "out_std": sout}

    @staticmethod
    def encode(x, nrm):
        return (x - nrm["in_profile"].T[:, None, None, :]) / nrm["in_std"]

    @staticmethod
    def decode(y, nrm):
        return y * nrm["out_std"] + nrm["out_profile"].T[:, None, None, :]

    def train(self, X, S, iters=4000, lr=2e-3, val=None, seed=0):
        """Trains on the samples not in `val` (default: every fourth and the
        last) with Adam, lr for the first three quarters of the iterations
        and lr/4 after; returns the network, the normalisation, the mean
        relative L2 error on the held-out samples and training figures."""
        n = len(X)
        val = sorted(set(val if val is not None else list(range(3, n, 4)) + [n - 1]))
        tr = [i for i in range(n) if i not in val]
        nrm = self.normalisation([X[i] for i in tr], [S[i] for i in tr])
        net = self.pkg.Net(8, 32, 7)
        net.set_params(init_params(net.nparam, seed))
        for i in tr:
            net.add_sample(self.encode(X[i], nrm),
                           (S[i] - nrm["out_profile"].T[:, None, None, :]) / nrm["out_std"])
        t0 = time.perf_counter()
        first, _ = net.train(iters - iters // 4, lr)
        _, last = net.train(iters // 4, lr / 4)
        ttrain = time.perf_counter() - t0
        errs, errs_profile = [], []
        for i in val:
            pred = self.decode(net.forward(self.encode(X[i], nrm)), nrm)
            errs.append(rel(pred, S[i]))
            errs_profile.append(rel(np.broadcast_to(nrm["out_profile"].T[:, None, None, :], S[i].shape), S[i]))
        self.log(f"train: loss {first:.4g} -> {last:.4g} in {ttrain:.1f} s; held-out {val}: {errs} "
                 f"(profile alone: {errs_profile})")
        return net, nrm, float(np.mean(errs)), {"loss_first": first, "loss_last": last, "train_s": ttrain,
                                                "train_samples": len(tr), "held_out": val,
                                                "held_out_error_profile_only": float(np.mean(errs_profile))}

    # -- judging ----------------------------------------------------------
    def coarse_run(self, q0, steps, dt, net=None, norm=None, val_error=0.0, threshold=0.1):
        """steps coarse mixed-form steps at fixed dt from the coarse state q0,
        with the closure attached if given; (final state, attached?, ms per
        step measured on a second box)."""
        def box():
            c = self.pkg.Box(self.mc, precision="mixed")
            c.upload(padded(q0))
            c.set_fixed_dt(dt)
            if net is not None:
                c.set_closure(net, norm["in_profile"], norm["in_std"], norm["out_profile"], norm["out_std"],
                              val_error, threshold)
            return c
        with box() as c:
            active = c.closure_active() if net is not None else False
            c.steps(steps)
            q = interior(c.download())
        with box() as c:
            ms = c.time_steps(10)["step"]
        return q, active, ms


def init_params(n, seed):
    """He-style initial weights for the 8 -> 32 -> 32 -> 7 network, in the
    library's parameter order."""
    rng = np.random.default_rng(seed)
    out = []
    for ci, co in ((8, 32), (32, 32), (32, 7)):
        out.append(rng.standard_normal(co * 27 * ci) * np.sqrt(2.0 / (27 * ci)))
        out.append(np.zeros(co))
    v = np.concatenate(out).astype(np.float32)
    assert v.size == n
    return v


def run(nxc=32, nyc=32, nzc=32, factor=4, develop=20000, nsamples=12, every=200, horizon=400, iters=3000,
        threshold=0.1, log=print):
    """The whole experiment; returns a JSON-able report.  horizon (fine
    steps of the a-posteriori comparison) must be a multiple of factor."""
    if horizon % factor:
        raise ValueError("horizon must be a multiple of factor")
    ex = Experiment(nxc, nyc, nzc, factor=factor, log=log)
    t0 = time.perf_counter()
    ex.develop(develop)
    dev_wall = time.perf_counter() - t0
    X, S, Q = ex.collect(nsamples, every)
    net, norm, val_err, tr = ex.train(X, S, iters=iters)
    # a posteriori: from the last (held-out) filtered snapshot
    q0 = Q[-1]
    dtf = ex.fine.step()[0]
    ex.fine.set_fixed_dt(dtf)
    ms_fine = ex.fine.time_steps(5)["step"]
    ex.fine.steps(horizon - 6)
    qf_end = ex.filtered()
    dtc, nc = dtf * factor, horizon // factor
    qp, _, ms_plain = ex.coarse_run(q0, nc, dtc)
    qc, active, ms_clos = ex.coarse_run(q0, nc, dtc, net, norm, val_err, threshold)
    fref, sref = convective_flux(qf_end), ke_spectrum(qf_end)
    rep = {
        "fine": [factor * nxc, factor * nyc, factor * (nzc - 1) + 1], "coarse": [nxc, nyc, nzc],
        "develop_steps": develop, "develop_wall_s": dev_wall, "samples": nsamples, "every": every,
        "horizon_fine_steps": horizon, "coarse_steps": nc, "dt_fine": dtf, "dt_coarse": dtc,
        "held_out_error": val_err, "threshold": threshold, "closure_active": active, "training": tr,
        "flux_error_plain": rel(convective_flux(qp), fref), "flux_error_closure": rel(convective_flux(qc), fref),
        "ke_spectrum_error_plain": rel(ke_spectrum(qp), sref), "ke_spectrum_error_closure": rel(ke_spectrum(qc), sref),
        "ms_step_fine_fp64": ms_fine, "ms_step_coarse_plain": ms_plain, "ms_step_coarse_closure": ms_clos,
        "gpu_time_ratio_fine_over_closure": (ms_fine * horizon) / (ms_clos * nc),
    }
    ex.fine.close()
    net.close()
    return rep


if __name__ == "__main__":
    import argparse
    import json
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    ap = argparse.ArgumentParser()
    ap.add_argument("--coarse", type=int, nargs=3, default=[32, 32, 32])
    ap.add_argument("--develop", type=int, default=20000)
    ap.add_argument("--samples", type=int, default=12)
    ap.add_argument("--every", type=int, default=200)
    ap.add_argument("--horizon", type=int, default=400)
    ap.add_argument("--iters", type=int, default=3000)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    r = run(*a.coarse, develop=a.develop, nsamples=a.samples, every=a.every, horizon=a.horizon, iters=a.iters,
            log=lambda *x: print(*x, file=sys.stderr, flush=True))
    s = json.dumps(r)
    if a.out:
        open(a.out, "w").write(s + "\n")
    print(s)
