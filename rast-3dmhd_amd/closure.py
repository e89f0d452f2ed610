"""Learned subgrid closure: training data from box-filtered plain (FP64)
fine runs, training of the BF16-MFMA network on the GPU, and closure-
corrected coarse runs (mixed form) judged against the filtered fine run.

Grids.  The fine box has factor*nxc x factor*nyc x factor*(nzc-1)+1 points
and the coarse box nxc x nyc x nzc, with the same horizontal periods and the
same tanh-stretched vertical map (NGRID=3): coarse plane K is then exactly
fine plane factor*K.  The filter (mhd_box_filter, zmode 1) averages factor x
factor horizontal blocks and the factor vertical spacings around each coarse
plane.  A fine run may be developed on a box coarser horizontally by
dev_factor (same periods and vertical grid) and prolonged spectrally.

Seed.  A relative temperature perturbation of the low horizontal Fourier
modes (|kx|, |ky| <= kmodes) times sin(pi zeta) of the mapped vertical
coordinate, so the same unstable modes are seeded at every resolution and
the layer turns over after a comparable simulated time.

Target.  For a filtered fine state q_f = F(q): the subgrid term of each
equation is S = F(RHS_fine(q)) - RHS_coarse(q_f) for RU, RV, RW, TT, BX, BY,
BZ; the network maps (rho, u, v, w, s, BX, BY, BZ) of q_f to S.  Inputs and
targets enter as deviations from their per-plane means over the training
samples, divided by one standard deviation per channel.  On the two boundary
planes the targets are the per-plane means alone (the boundary conditions
overwrite RU, RV, RW, TT and BZ there after every stage, and the coarse
right-hand side on them depends on the frozen ghost planes).

Judge.  The mean convective flux F_conv(z) = -<rho w (2.5 T' + |u|^2/2)>
(horizontal average, T' the deviation from the horizontal mean; the
reference's FCONM, 3dmhdsub.f:1650-1661), averaged over snapshots of the
comparison window, and the horizontal kinetic-energy spectrum, of the
coarse runs against the filtered fine run at the same simulated times.

Fallback.  The closure is attached only if (a) its error on held-out
samples is at most `threshold` (the library's own rule) and (b) on the
held-out sampling intervals, coarse runs with the closure reproduce the
filtered fine flux profile at the next sample better than plain coarse runs.
Otherwise the coarse run is the plain one.
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


def low_modes(npx, npy, npz, kmodes, amp, seed):
    """Relative perturbation [K][J][I]: random amplitudes and phases on the
    horizontal modes 0 < max(|kx|, |ky|) <= kmodes, times sin(pi K/(npz-1)),
    scaled to an RMS of amp."""
    rng = np.random.default_rng(seed)
    x = np.arange(npx) / npx
    y = np.arange(npy) / npy
    h = np.zeros((npy, npx))
    for kx in range(-kmodes, kmodes + 1):
        for ky in range(0, kmodes + 1):
            if ky == 0 and kx <= 0:
                continue
            a, ph = rng.standard_normal(), rng.uniform(0.0, 2.0 * np.pi)
            h += a * np.cos(2.0 * np.pi * (kx * x[None, :] + ky * y[:, None]) + ph)
    prof = np.sin(np.pi * np.arange(npz) / (npz - 1))
    p = prof[:, None, None] * h[None]
    return p * (amp / np.sqrt((p * p).mean()))


class Grids:
    """Metas of the fine, coarse and development boxes of one experiment."""

    def __init__(self, nxc, nyc, nzc, factor=4, dev_factor=1, contrast=10.0, ra=1e5, pr=1.0, pm=1.0, ta=1e4):
        _, self.cfg, _ = _mods()
        self.ph = self.cfg.physics(contrast=contrast, ra=ra, pr=pr, pm=pm, ta=ta, tilt_deg=0.0)
        self.factor = factor
        self.fine_shape = (factor * nxc, factor * nyc, factor * (nzc - 1) + 1)
        self.coarse_shape = (nxc, nyc, nzc)
        nxf, nyf, nzf = self.fine_shape
        if nxf % dev_factor or nyf % dev_factor:
            raise ValueError("dev_factor must divide the fine horizontal sizes")
        self.dev_shape = (nxf // dev_factor, nyf // dev_factor, nzf)
        # periods n * dx with dx = xmax / (n - 1); the fine box keeps the default xmax
        self.period = (float(self.ph[17]) * nxf / (nxf - 1), float(self.ph[18]) * nyf / (nyf - 1))

    def meta(self, shape):
        nx, ny, nz = shape
        cfg = self.cfg
        par = {k: cfg.f32(v) for k, v in self.ph.items()}
        par[17] = cfg.f32(self.period[0] * (nx - 1) / nx)
        par[18] = cfg.f32(self.period[1] * (ny - 1) / ny)
        mt = cfg.metrics(nx, ny, nz, xmax=par[17], ymax=par[18], zmax=par[19], ngrid=3)
        t, r = cfg.polytrope(mt["zee"], par[3], par[4])
        m = cfg.constants(nx, ny, nz, par, lmag=True, lrot=True, izcon=1, itcon=0, tb=float(t[-1]))
        m.update(mt)
        return m, t, r

    def initial(self, shape, ampt=5e-2, kmodes=3, ampb=1e-5, kmax=4, seed=77):
        """Padded initial state of a box of this shape: the polytrope with the
        low-mode temperature seed and a seed magnetic field."""
        _, _, ics = _mods()
        m, t, r = self.meta(shape)
        nx, ny, nz = shape
        st = ics.static_state(nx, ny, nz, t, r, ampt=0.0, lmag=True)
        st["tt"][1:-1, 1:-1, 1:-1] *= 1.0 + low_modes(nx, ny, nz, kmodes, ampt, seed).transpose(2, 1, 0)
        st["bx"], st["by"], st["bz"] = ics.vector_potential_field(m, kmax=kmax, amp=ampb, seed=seed)
        for k in ("ru", "rv", "rw", "ro", "tt"):
            ics.periodic_fill(st[k])
        return m, st


def prolong(fields, nx, ny):
    """Spectral horizontal prolongation of padded fields (all planes, ghost
    planes included) to nx x ny interior points; the Nyquist rows of the
    source are dropped.  Periodic ghosts are refilled."""
    _, _, ics = _mods()
    out = {}
    for k, a in fields.items():
        src = np.asarray(a)[1:-1, 1:-1, :]
        mx, my, nz = src.shape
        spec = np.fft.fft2(src, axes=(0, 1))
        big = np.zeros((nx, ny, nz), dtype=complex)
        hx, hy = mx // 2, my // 2
        for sx, dx in ((slice(0, hx), slice(0, hx)), (slice(mx - hx + 1, mx), slice(nx - hx + 1, nx))):
            for sy, dy in ((slice(0, hy), slice(0, hy)), (slice(my - hy + 1, my), slice(ny - hy + 1, ny))):
                big[dx, dy] = spec[sx, sy]
        p = np.zeros((nx + 2, ny + 2, nz), order="F")
        p[1:-1, 1:-1, :] = np.fft.ifft2(big, axes=(0, 1)).real * (nx * ny / (mx * my))
        out[k] = ics.periodic_fill(p)
    return out


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


def judge(ref, got):
    """Errors of coarse snapshots `got` against filtered fine snapshots
    `ref` (same simulated times): the time-mean flux profile, the flux
    profile at the last snapshot, per-snapshot flux errors and the
    time-mean kinetic-energy spectrum."""
    fr = [convective_flux(q) for q in ref]
    fg = [convective_flux(q) for q in got]
    return {"flux_mean": rel(np.mean(fg, axis=0), np.mean(fr, axis=0)), "flux_end": rel(fg[-1], fr[-1]),
            "flux_each": [rel(a, b) for a, b in zip(fg, fr)],
            "ke_mean": rel(np.mean([ke_spectrum(q) for q in got], axis=0),
                           np.mean([ke_spectrum(q) for q in ref], axis=0))}


class Experiment:
    """One fine/coarse pair on the current GPU."""

    def __init__(self, nxc, nyc, nzc, factor=4, dev_factor=1, log=print, seed=77, ampt=5e-2, kmodes=3, **kw):
        self.pkg = _mods()[0]
        self.g = Grids(nxc, nyc, nzc, factor=factor, dev_factor=dev_factor, **kw)
        self.factor, self.dev_factor = factor, dev_factor
        self.log = log
        self.mf, _, _ = self.g.meta(self.g.fine_shape)
        self.mc, _, _ = self.g.meta(self.g.coarse_shape)
        self.gamma = float(self.mf["gamma"])
        md, st = self.g.initial(self.g.dev_shape, ampt=ampt, kmodes=kmodes, seed=seed)
        self.fine = self.pkg.Box(md)
        self.fine.upload(st)
        self.on_fine_grid = dev_factor == 1
        self.times = []

    # -- data -------------------------------------------------------------
    def develop(self, steps, chunk=1000, until_mach=None):
        """steps fine steps (or, with until_mach, at most steps, stopping at
        the first chunk whose Mach number reaches until_mach)."""
        done, um = 0, 0.0
        while done < steps:
            n = min(chunk, steps - done)
            dt, um = self.fine.steps(n)
            done += n
            self.log(f"{'fine' if self.on_fine_grid else 'dev'}: {done}/{steps} steps, dt {dt:.3e}, mach {um:.3e}")
            if until_mach is not None and um >= until_mach:
                break
        return done, um

    def to_fine_grid(self):
        """Prolongs the development box onto the fine grid and continues there."""
        if self.on_fine_grid:
            return
        nit, timc, timt = self.fine.clock()
        st = prolong(self.fine.download(), *self.g.fine_shape[:2])
        self.fine.close()
        self.fine = self.pkg.Box(self.mf)
        self.fine.upload(st)
        self.fine.set_clock(nit, timc, timt - timc)
        self.on_fine_grid = True
        self.log(f"prolonged to {self.g.fine_shape} at t = {timt:.4e}")

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
        """nsamples samples every `every` fine steps; also records the
        simulated time of each."""
        X, S, Q, T = [], [], [], []
        for i in range(nsamples):
            self.develop(every, chunk=every)
            x, s, q = self.sample()
            X.append(x)
            S.append(s)
            Q.append(q)
            T.append(self.fine.clock()[2])
            self.log(f"sample {i + 1}/{nsamples}: t {T[-1]:.4e}, |S| {np.sqrt((s ** 2).mean(axis=(0, 1, 2)))}")
        return X, S, Q, T

    # -- training ---------------------------------------------------------
    @staticmethod
    def normalisation(X, S):
        """Per-plane means of the inputs and targets over the samples (the
        horizontally uniform part, e.g. the different discrete hydrostatic
        and diffusive balance of the two grids, is carried by the profile)
        and the standard deviation of what remains, per channel, over the
        planes off the boundaries."""
        xs, ss = np.stack(X), np.stack(S)               # [n][K][J][I][c]
        pin = xs.mean(axis=(0, 2, 3))                    # [K][8]
        pout = ss.mean(axis=(0, 2, 3))                   # [K][7]
        sin = (xs - pin[None, :, None, None, :])[:, 1:-1].std(axis=(0, 1, 2, 3))
        sout = (ss - pout[None, :, None, None, :])[:, 1:-1].std(axis=(0, 1, 2, 3))
        sin[sin == 0] = 1.0
        sout[sout == 0] = 1.0
        return {"in_profile": pin.T.copy(), "in_std": sin, "out_profile": pout.T.copy(), "out_std": sout}

    @staticmethod
    def encode(x, nrm):
        return (x - nrm["in_profile"].T[:, None, None, :]) / nrm["in_std"]

    @staticmethod
    def decode(y, nrm):
        return y * nrm["out_std"] + nrm["out_profile"].T[:, None, None, :]

    @staticmethod
    def target(s, nrm):
        """Normalised training target; zero on the boundary planes."""
        t = (s - nrm["out_profile"].T[:, None, None, :]) / nrm["out_std"]
        t[0] = 0.0
        t[-1] = 0.0
        return t

    @staticmethod
    def held_out(n):
        """Every fourth sample and the last."""
        return sorted(set(list(range(3, n, 4)) + [n - 1]))

    def train(self, X, S, iters=4000, lr=2e-3, val=None, seed=0):
        """Trains on the samples not in `val` (default: held_out) with Adam,
        lr for the first three quarters of the iterations and lr/4 after;
        returns the network, the normalisation, the mean relative L2 error
        on the held-out samples (planes off the boundaries) and training
        figures."""
        n = len(X)
        val = sorted(set(val if val is not None else self.held_out(n)))
        tr = [i for i in range(n) if i not in val]
        nrm = self.normalisation([X[i] for i in tr], [S[i] for i in tr])
        net = self.pkg.Net(8, 32, 7)
        net.set_params(init_params(net.nparam, seed))
        for i in tr:
            net.add_sample(self.encode(X[i], nrm), self.target(S[i], nrm))
        t0 = time.perf_counter()
        first, _ = net.train(iters - iters // 4, lr)
        _, last = net.train(iters // 4, lr / 4)
        ttrain = time.perf_counter() - t0
        errs, errs_profile, per_channel = [], [], []
        prof = nrm["out_profile"].T[:, None, None, :]
        for i in val:
            pred = self.decode(net.forward(self.encode(X[i], nrm)), nrm)[1:-1]
            ref = S[i][1:-1]
            errs.append(rel(pred, ref))
            errs_profile.append(rel(np.broadcast_to(prof[1:-1], ref.shape), ref))
            per_channel.append([rel(pred[..., c], ref[..., c]) for c in range(7)])
        self.log(f"train: loss {first:.4g} -> {last:.4g} in {ttrain:.1f} s; held-out {val}: {errs} "
                 f"(profile alone: {errs_profile})")
        return net, nrm, float(np.mean(errs)), {
            "loss_first": first, "loss_last": last, "train_s": ttrain, "train_samples": len(tr), "held_out": val,
            "held_out_error_profile_only": float(np.mean(errs_profile)),
            "held_out_error_per_channel": dict(zip(TARGETS, np.mean(per_channel, axis=0).tolist()))}

    # -- judging ----------------------------------------------------------
    def _coarse_box(self, q0, net, norm, val_error, threshold):
        c = self.pkg.Box(self.mc, precision="mixed")
        c.upload(padded(q0))
        if net is not None:
            c.set_closure(net, norm["in_profile"], norm["in_std"], norm["out_profile"], norm["out_std"],
                          val_error, threshold)
        return c

    def coarse_run(self, q0, intervals, net=None, norm=None, val_error=0.0, threshold=0.1):
        """Coarse mixed-form run from the coarse state q0 over intervals
        [(steps, dt), ...] at fixed dt, with the closure attached if given;
        (filtered-grid snapshot after each interval, attached?)."""
        with self._coarse_box(q0, net, norm, val_error, threshold) as c:
            active = c.closure_active() if net is not None else False
            snaps = []
            for m, dt in intervals:
                c.set_fixed_dt(dt)
                c.steps(m)
                snaps.append(interior(c.download()))
        return snaps, active

    def coarse_ms(self, q0, dt, net=None, norm=None):
        """Mean milliseconds per coarse step (HIP events over 10 steps)."""
        with self._coarse_box(q0, net, norm, 0.0, 1.0) as c:
            c.set_fixed_dt(dt)
            return c.time_steps(10)["step"]

    def validate(self, net, norm, Q, T, every):
        """A-posteriori check on the held-out sampling intervals: coarse runs
        from each held-out sample (but the last) to the next sample, plain
        and with the closure; flux-profile errors against the next filtered
        sample."""
        m = every // self.factor
        plain, clos = [], []
        for i in self.held_out(len(Q))[:-1]:
            iv = [(m, (T[i + 1] - T[i]) / m)]
            ref = [Q[i + 1]]
            plain.append(judge(ref, self.coarse_run(Q[i], iv)[0])["flux_end"])
            clos.append(judge(ref, self.coarse_run(Q[i], iv, net, norm)[0])["flux_end"])
        return {"flux_plain": plain, "flux_closure": clos,
                "closure_better": bool(np.mean(clos) < np.mean(plain)) if plain else True}


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


def run(nxc=32, nyc=32, nzc=32, factor=4, dev_factor=1, until_mach=0.05, develop_max=100000, develop=2000,
        nsamples=24, every=200, horizon=1600, nsnap=8, iters=6000, threshold=0.1, log=print, **kw):
    """The whole experiment; returns a JSON-able report.  every and
    horizon/nsnap (fine steps) must be multiples of factor."""
    seg = horizon // nsnap
    if every % factor or seg % factor or seg * nsnap != horizon:
        raise ValueError("every and horizon/nsnap must be multiples of factor")
    ex = Experiment(nxc, nyc, nzc, factor=factor, dev_factor=dev_factor, log=log, **kw)
    t0 = time.perf_counter()
    n_dev, mach_dev = ex.develop(develop_max, until_mach=until_mach)
    ex.to_fine_grid()
    ex.develop(develop)
    dev_wall = time.perf_counter() - t0
    X, S, Q, T = ex.collect(nsamples, every)
    net, norm, val_err, tr = ex.train(X, S, iters=iters)
    post = ex.validate(net, norm, Q, T, every)
    attach = val_err <= threshold and post["closure_better"]
    log(f"a-posteriori validation: {post}; attach {attach}")
    # comparison window after the sampling window, from the last (held-out) sample
    q0, refs, times = Q[-1], [], [T[-1]]
    for _ in range(nsnap):
        ex.develop(seg, chunk=seg)
        refs.append(ex.filtered())
        times.append(ex.fine.clock()[2])
    m = seg // factor
    iv = [(m, (times[i + 1] - times[i]) / m) for i in range(nsnap)]
    plain, _ = ex.coarse_run(q0, iv)
    clos, active = ex.coarse_run(q0, iv, net if attach else None, norm, val_err, threshold)
    with ex.pkg.Net(8, 32, 7) as zero:
        zero.set_params(np.zeros(zero.nparam, dtype=np.float32))
        prof_only, _ = ex.coarse_run(q0, iv, zero, norm)
    jp, jc, jz = judge(refs, plain), judge(refs, clos), judge(refs, prof_only)
    # GPU time per simulated interval: the fine run (FP64, as run, and the
    # mixed form for comparison) against the coarse run with the network
    dt_c = iv[-1][1]
    ms_plain = ex.coarse_ms(refs[-1], dt_c)
    ms_clos = ex.coarse_ms(refs[-1], dt_c, net, norm)
    ms_fine = ex.fine.time_steps(5)["step"]
    st = ex.fine.download()
    ex.fine.close()
    with ex.pkg.Box(ex.mf, precision="mixed") as fm:
        fm.upload(st)
        ms_fine_mixed = fm.time_steps(5)["step"]
    del st
    ms_net = net.time_forward(nxc, nyc, nzc, reps=10)
    net.close()
    rep = {
        "fine": list(ex.g.fine_shape), "coarse": [nxc, nyc, nzc], "dev_factor": dev_factor,
        "physics": {k: kw[k] for k in ("contrast", "ra", "pr", "pm", "ta") if k in kw},
        "develop_steps": [n_dev, develop], "develop_mach": mach_dev, "develop_wall_s": dev_wall,
        "samples": nsamples, "every": every, "sample_times": T, "horizon_fine_steps": horizon, "snapshots": nsnap,
        "window_times": times, "coarse_steps": m * nsnap,
        "held_out_error": val_err, "threshold": threshold, "aposteriori_validation": post,
        "closure_active": active, "training": tr,
        "flux_error_plain": jp["flux_mean"], "flux_error_closure": jc["flux_mean"],
        "flux_error_profile_only": jz["flux_mean"],
        "flux_error_end_plain": jp["flux_end"], "flux_error_end_closure": jc["flux_end"],
        "flux_error_each_plain": jp["flux_each"], "flux_error_each_closure": jc["flux_each"],
        "ke_spectrum_error_plain": jp["ke_mean"], "ke_spectrum_error_closure": jc["ke_mean"],
        "ke_spectrum_error_profile_only": jz["ke_mean"],
        "ms_step_fine_fp64": ms_fine, "ms_step_fine_mixed": ms_fine_mixed,
        "ms_step_coarse_plain": ms_plain, "ms_step_coarse_closure": ms_clos, "ms_net_forward": ms_net,
        "gpu_time_ratio_fine_over_closure": (ms_fine * horizon) / (ms_clos * m * nsnap),
        "gpu_time_ratio_fine_mixed_over_closure": (ms_fine_mixed * horizon) / (ms_clos * m * nsnap),
    }
    return rep


if __name__ == "__main__":
    import argparse
    import json
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    ap = argparse.ArgumentParser()
    ap.add_argument("--coarse", type=int, nargs=3, default=[32, 32, 32])
    ap.add_argument("--dev-factor", type=int, default=1)
    ap.add_argument("--until-mach", type=float, default=0.05)
    ap.add_argument("--develop-max", type=int, default=100000)
    ap.add_argument("--develop", type=int, default=2000)
    ap.add_argument("--samples", type=int, default=24)
    ap.add_argument("--every", type=int, default=200)
    ap.add_argument("--horizon", type=int, default=1600)
    ap.add_argument("--snapshots", type=int, default=8)
    ap.add_argument("--iters", type=int, default=6000)
    ap.add_argument("--ampt", type=float, default=5e-2)
    ap.add_argument("--contrast", type=float, default=10.0, help="density contrast of the layer")
    ap.add_argument("--ra", type=float, default=1e5, help="Rayleigh number")
    ap.add_argument("--pm", type=float, default=1.0, help="magnetic Prandtl number")
    ap.add_argument("--ta", type=float, default=1e4, help="Taylor number (vertical rotation axis)")
    ap.add_argument("--out", default=None)
    ap.add_argument("--time-net", type=int, nargs=3, default=None, metavar=("NX", "NY", "NZ"),
                    help="only time the network's forward pass (He-initialised weights) on an NX x NY x NZ box")
    a = ap.parse_args()
    if a.time_net:
        pkg = _mods()[0]
        with pkg.Net(8, 32, 7) as net:
            net.set_params(init_params(net.nparam, 1))
            print(json.dumps({"net_forward_ms": net.time_forward(*a.time_net, reps=20), "box": a.time_net}))
        sys.exit(0)
    r = run(*a.coarse, dev_factor=a.dev_factor, until_mach=a.until_mach, develop_max=a.develop_max,
            develop=a.develop, nsamples=a.samples, every=a.every, horizon=a.horizon, nsnap=a.snapshots,
            iters=a.iters, ampt=a.ampt, contrast=a.contrast, ra=a.ra, pm=a.pm, ta=a.ta,
            log=lambda *x: print(*x, file=sys.stderr, flush=True))
    s = json.dumps(r)
    if a.out:
        open(a.out, "w").write(s + "\n")
    print(s)
