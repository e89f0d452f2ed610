"""The closure network's BF16 MFMA kernels against the CPU oracle
(oracle/net_oracle.py, pinned by finite differences): forward pass,
parameter and input gradients, and training on the device."""

import os
import sys

import numpy as np
import pytest

import helpers as H

sys.path.insert(0, os.path.join(H.REPO, "oracle"))
import net_oracle as N  # noqa: E402

pytestmark = pytest.mark.gpu

CIN, WIDTH, COUT = 8, 32, 7


def say(*a):
    print(*a, flush=True)


def rel(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


@pytest.mark.parametrize("shape", [(6, 8, 32), (5, 9, 70), (3, 4, 16)])
def test_forward_matches_oracle(shape):
    nz, ny, nx = shape
    say("case net forward", shape)
    rng = np.random.default_rng(sum(shape))
    p = N.init_params(CIN, WIDTH, COUT, seed=1)
    x = N.bf16(rng.standard_normal((nz, ny, nx, CIN)))
    ref = N.Net(p, bf16_emulation=True).forward(x)
    with H.pkg().Net(CIN, WIDTH, COUT) as net:
        net.set_params(N.flatten(p))
        y = net.forward(x)
    say("rel", rel(y, ref), "max", np.abs(y - ref).max() / np.abs(ref).max())
    assert rel(y, ref) <= 2e-3
    assert np.abs(y - ref).max() <= 2e-2 * np.abs(ref).max()


def test_gradients_match_oracle():
    say("case net gradient")
    nz, ny, nx = 4, 8, 32
    rng = np.random.default_rng(7)
    p = N.init_params(CIN, WIDTH, COUT, seed=2)
    x = N.bf16(rng.standard_normal((nz, ny, nx, CIN)))
    dy = N.bf16(rng.standard_normal((nz, ny, nx, COUT)))
    g_ref, dx_ref = N.Net(p, bf16_emulation=True).backward(x, dy)
    with H.pkg().Net(CIN, WIDTH, COUT) as net:
        net.set_params(N.flatten(p))
        g, dx = net.gradient(x, dy, want_dx=True)
    got = N.unflatten(g, CIN, WIDTH, COUT)
    for k in ("W3", "b3", "W2", "b2", "W1", "b1"):
        e = rel(got[k], g_ref[k])
        say(k, e)
        assert e <= 1e-2, (k, e)
    assert rel(dx, dx_ref) <= 1e-2


def test_training_reduces_loss():
    """Fit the outputs of a teacher network with other weights."""
    say("case net training")
    nz, ny, nx = 8, 16, 32
    rng = np.random.default_rng(11)
    teacher = N.init_params(CIN, WIDTH, COUT, seed=3)
    with H.pkg().Net(CIN, WIDTH, COUT) as net:
        net.set_params(N.flatten(teacher))
        xs = [N.bf16(rng.standard_normal((nz, ny, nx, CIN))) for _ in range(2)]
        ts = [net.forward(x) for x in xs]
        net.set_params(N.flatten(N.init_params(CIN, WIDTH, COUT, seed=4)))
        for x, t in zip(xs, ts):
            net.add_sample(x, t)
        l0 = net.loss(0)
        first, last = net.train(600, 2e-3)
        l1 = net.loss(0)
        say("loss", l0, first, last, l1)
    assert first == pytest.approx(l0, rel=1e-6)
    assert l1 < 0.25 * l0


def test_loss_and_adam_step_match_the_oracle():
    """The training loss (k_loss) and one Adam step (k_adam) against their
    CPU restatement: the loss of a sample from the network's own forward
    pass, and the parameters after one training step from the gradient the
    backward pass gives for the loss gradient."""
    say("case net loss and Adam")
    nz, ny, nx = 6, 8, 16
    rng = np.random.default_rng(21)
    x = N.bf16(rng.standard_normal((nz, ny, nx, CIN)))
    t = (0.3 * rng.standard_normal((nz, ny, nx, COUT))).astype(np.float32)
    p0 = N.flatten(N.init_params(CIN, WIDTH, COUT, seed=5)).astype(np.float32)
    with H.pkg().Net(CIN, WIDTH, COUT) as net:
        net.set_params(p0)
        y = net.forward(x)
        net.add_sample(x, t)
        loss = net.loss(0)
        want, dy = N.mse_loss(y, t, COUT)
        assert loss == pytest.approx(want, rel=1e-12)
        g = net.gradient(x, dy)
        first, _ = net.train(1, 1e-3)
        p1 = net.params()
    assert first == pytest.approx(want, rel=1e-12)
    z = np.zeros_like(p0)
    p1w, _, _ = N.adam_step(p0, g, z, z, 1, 1e-3)
    # the same float32 formula: equal to a few float32 ulps of the parameters
    tol = 4 * np.finfo(np.float32).eps * np.abs(p0).max()
    assert np.abs(p1 - p1w).max() <= tol, np.abs(p1 - p1w).max()
    assert np.abs((p0 - p1) - (p0 - p1w)).max() <= tol


@pytest.mark.parametrize("npz,zmode", [(17, 1), (16, 0)])
def test_box_filter_matches_oracle(npz, zmode):
    say("case box filter", npz, zmode)
    from test_gpu_parity import random_box
    meta, st = random_box(64, 32, npz, True, True, seed=npz)
    with H.pkg().Box(meta) as box:
        box.upload(st)
        got = box.box_filter(4, zmode=zmode, source="state")
        got_rhs = box.box_filter(4, zmode=zmode, source="rhs")
        rhs = box.fluxes()
    filt = N.box_filter_vertex if zmode == 1 else N.box_filter
    for f, k in enumerate(H.FIELDS):
        ref = filt(st[k][1:-1, 1:-1, 1:-1].transpose(2, 1, 0), 4)
        assert np.allclose(got[f], ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max()), k
    for f, k in enumerate(["fu", "fv", "fw", "fr", "ft", "ww1", "ww2", "ww3"]):
        ref = filt(rhs[k][1:-1, 1:-1, 1:-1].transpose(2, 1, 0), 4)
        assert np.allclose(got_rhs[f], ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max()), k


def _closure_case():
    import importlib
    wl = importlib.import_module("rast-3dmhd_amd.workload")
    return wl.stratified_box(64, 48, 24, seed=3)


def _run_mixed(meta, st, nsteps, net=None, val_error=0.0, threshold=0.1, out_std=None):
    with H.pkg().Box(meta, precision="mixed") as box:
        box.upload(st)
        if net is not None:
            nz = int(meta["npz"])
            box.set_closure(net, np.zeros((8, nz)), np.ones(8), np.zeros((7, nz)),
                            np.ones(7) if out_std is None else out_std, val_error, threshold)
        dt, _ = box.steps(nsteps)
        return box.download(), dt, (box.closure_active() if net is not None else False)


def test_closure_fallbacks_step_plainly():
    """Rejected network (error above threshold), all-zero network, and a
    network with non-finite output: every one steps bit-identically to the
    plain mixed step."""
    say("case closure fallbacks")
    meta, st = _closure_case()
    ref, dt_ref, _ = _run_mixed(meta, st, 3)
    with H.pkg().Net(CIN, WIDTH, COUT) as net:
        net.set_params(N.flatten(N.init_params(CIN, WIDTH, COUT, seed=5)))
        got, dt, active = _run_mixed(meta, st, 3, net, val_error=0.5, threshold=0.1)
        assert not active and dt == dt_ref
        for k in H.FIELDS:
            assert np.array_equal(got[k], ref[k]), ("rejected", k)
        net.set_params(np.zeros(net.nparam, dtype=np.float32))
        got, dt, active = _run_mixed(meta, st, 3, net)
        assert active and dt == dt_ref
        for k in H.FIELDS:
            assert np.array_equal(got[k], ref[k]), ("zero", k)
        bad = np.zeros(net.nparam, dtype=np.float32)
        bad[-COUT:] = np.nan  # output biases
        net.set_params(bad)
        got, dt, active = _run_mixed(meta, st, 3, net)
        assert active and dt == dt_ref
        for k in H.FIELDS:
            assert np.array_equal(got[k], ref[k]), ("nan", k)


def test_closure_constant_forcing_adds_dt_g():
    """Only output biases set: a constant forcing g per field; one step adds
    dt * g (the RK3 weights sum to one) up to the coupling within the step."""
    say("case closure constant forcing")
    meta, st = _closure_case()
    ref, dt_ref, _ = _run_mixed(meta, st, 1)
    g = np.array([1e-3, -2e-3, 5e-4, 1e-3, 1e-6, -1e-6, 0.0])
    with H.pkg().Net(CIN, WIDTH, COUT) as net:
        p = np.zeros(net.nparam, dtype=np.float32)
        p[-COUT:] = np.sign(g)  # output = sign(g); de-normalised by |g|
        net.set_params(p)
        got, dt, active = _run_mixed(meta, st, 1, net, out_std=np.where(g == 0, 1.0, np.abs(g)))
    assert active and dt == dt_ref
    # away from the walls: next to them a uniform RW forcing meets RW = 0 at
    # the boundary plane and the acoustic response within the step is
    # O(CFL^2), a few per cent of dt * g
    inner = (slice(1, -1), slice(1, -1), slice(3, -3))
    for f, k in enumerate(["ru", "rv", "rw", "tt", "bx", "by"]):
        d = (got[k] - ref[k])[inner]
        want = dt * g[f]
        assert np.abs(d - want).max() <= 1e-2 * abs(want), (k, np.abs(d - want).max(), want)


def test_closure_forcing_matches_host_evaluation():
    """Network inputs built on the device (per-plane input profiles, input
    scales), network output de-normalised with per-plane output profiles:
    with a fixed, tiny dt the step changes each forced field by dt * g, g
    evaluated on the host from the same state with the same network."""
    import importlib
    say("case closure forcing vs host evaluation")
    cl = importlib.import_module("rast-3dmhd_amd.closure")
    meta, st = _closure_case()
    # the case starts at rest with horizontally uniform density: add motion
    # and density fluctuations so that every input channel varies
    rng = np.random.default_rng(21)
    inner3 = (slice(1, -1), slice(1, -1), slice(1, -1))
    st["ro"][inner3] *= 1.0 + 1e-3 * rng.standard_normal(st["ro"][inner3].shape)
    for k in ("ru", "rv", "rw"):
        st[k][inner3] = st["ro"][inner3] * 1e-2 * rng.standard_normal(st[k][inner3].shape)
    st["rw"][:, :, 1] = st["rw"][:, :, -2] = 0.0
    for k in ("ru", "rv", "rw", "ro"):
        H.ics().periodic_fill(st[k])
    _, dt0, _ = _run_mixed(meta, st, 1)
    dt = dt0 * 1e-3
    nz = int(meta["npz"])
    x = cl.inputs_of(cl.interior(st), float(meta["gamma"]))
    in_profile = x.mean(axis=(1, 2)).T.copy()
    in_std = (x - in_profile.T[:, None, None, :]).std(axis=(0, 1, 2))
    rms = np.array([np.sqrt(np.mean(st[k][1:-1, 1:-1, 1:-1] ** 2)) for k in cl.TARGETS])
    out_std = 1e-3 * rms / dt
    kk = np.arange(nz) / (nz - 1)
    out_profile = np.stack([0.5 * out_std[c] * np.cos(np.pi * kk + c) for c in range(7)])
    with H.pkg().Net(CIN, WIDTH, COUT) as net:
        net.set_params(N.flatten(N.init_params(CIN, WIDTH, COUT, seed=9)))
        nrm = {"in_profile": in_profile, "in_std": in_std, "out_profile": out_profile, "out_std": out_std}
        g = cl.Experiment.decode(net.forward(cl.Experiment.encode(x, nrm)), nrm)  # [K][J][I][7]
        outs = []
        for attach in (False, True):
            with H.pkg().Box(meta, precision="mixed") as box:
                box.upload(st)
                box.set_fixed_dt(dt)
                if attach:
                    box.set_closure(net, in_profile, in_std, out_profile, out_std, 0.0, 0.1)
                    assert box.closure_active()
                box.steps(1)
                outs.append(box.download())
    inner = (slice(1, -1), slice(1, -1), slice(3, -3))
    for c, k in enumerate(cl.TARGETS):
        d = (outs[1][k] - outs[0][k])[inner]
        want = dt * g[..., c].transpose(2, 1, 0)[:, :, 2:-2]
        err = np.abs(d - want).max() / np.abs(want).max()
        say(k, err)
        assert err <= 5e-3, (k, err)


def test_closure_pipeline_beats_plain_coarse():
    """Whole pipeline at coarse 32x32x32 against filtered fine 128x128x125:
    the network trained on the GPU passes the held-out and a-posteriori
    checks, is attached, and the closure-corrected coarse run reproduces the
    filtered fine run's mean convective-flux profile and kinetic-energy
    spectrum better than the plain coarse run."""
    import importlib
    say("case closure pipeline 32x32x32")
    cl = importlib.import_module("rast-3dmhd_amd.closure")
    r = cl.run(32, 32, 32, until_mach=0.05, develop=2000, nsamples=24, every=200, horizon=1600, nsnap=8,
               iters=6000, log=say)
    say({k: r[k] for k in ("held_out_error", "closure_active", "flux_error_plain", "flux_error_closure",
                           "ke_spectrum_error_plain", "ke_spectrum_error_closure")})
    assert r["held_out_error"] <= 0.1 and r["aposteriori_validation"]["closure_better"]
    assert r["closure_active"]
    assert r["flux_error_closure"] < 0.5 * r["flux_error_plain"]
    assert r["flux_error_closure"] <= 0.05
    assert r["ke_spectrum_error_closure"] < r["ke_spectrum_error_plain"]
