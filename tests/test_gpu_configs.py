"""The benchmark workloads themselves, each at its own size, through the C-ABI.

  * 64x64x32 MHD + rotation, 200 steps: the reference's own run, bit for bit
    at steps 1, 3, 50, 100 and 200 (test_gpu_parity.py, fixture mhdrot64).
  * 256x256x128 (contrast 10, Ra 1e5, Pr 1, Pm 1, Ta 1e4, tanh z grid, seed
    field |k| <= 4 at 1e-5): FP64 against the oracle bit for bit for two
    steps; mixed against FP64 after 1000 steps at one fixed dt, within 1e-6
    relative L2 per field.
  * 512x512x256 (contrast 30, Ra 1e6, Pm 2, Ta 1e6 at 45 degrees, seed field
    |k| <= 8 at 1e-3), the headline box: mixed against FP64 after 500 steps
    within 1e-6 per field, and periodic-shift equivariance of the mixed step.
  * 64 uneven boxes of the ensemble table (sizes 32..96 x 32..96 x 32..64,
    Ra, Ta, contrast per box): autotuned tiles, graph replay, every box bit
    for bit the single-box mixed run.
  * closure at coarse 128x128x64: the network's forward pass on the whole
    volume against the oracle on sub-blocks, and one closure-corrected step
    against the forcing evaluated on the host.
"""

import importlib
import os
import sys

import numpy as np
import pytest

import helpers as H

sys.path.insert(0, os.path.join(H.REPO, "oracle"))
import net_oracle as N  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = 1e-6  # mixed vs FP64, relative L2 per stored field


def say(*a):
    print(*a, flush=True)


def wl():
    return importlib.import_module("rast-3dmhd_amd.workload")


def run(meta, st, precision, n, fixed_dt=0.0, chunk=250):
    with H.pkg().Box(meta, precision=precision) as box:
        box.upload(st)
        if fixed_dt:
            box.set_fixed_dt(fixed_dt)
        done = 0
        while done < n:
            k = min(chunk, n - done)
            dt, um = box.steps(k)
            done += k
            say(f"  {precision}: {done}/{n} steps")
        return box.download(), dt, um, box.clock()


def first_dt(meta, st):
    with H.pkg().Box(meta) as box:
        box.upload(st)
        return box.step()[0]


def check_errors(meta, mix, ref):
    err = wl().field_errors(meta, mix, ref)
    say("rel_l2", err["rel_l2"], "velocity", err["velocity"], "entropy", err["entropy"],
        "b", err.get("b_field"), "divb", err.get("divb_max"))
    for k, e in err["rel_l2"].items():
        assert e <= TOL, (k, e)
    assert err["velocity"] <= TOL and err["b_field"] <= TOL
    d = err["divb_max"]
    assert abs(d["test"] - d["ref"]) <= 1e-3 * d["ref"] + 1e-12 * d["b_max"], d
    return err


def test_256_fp64_two_steps_bitwise_vs_oracle():
    say("case 256x256x128 fp64 vs oracle, 2 steps")
    meta, st = wl().stratified_box(256, 256, 128)
    O = H.oracle_mod()
    o = O.Oracle(meta, st)
    with H.pkg().Box(meta) as box:
        box.upload(st)
        for n in range(2):
            dt, um = box.step()
            o.step()
            assert dt == o.meta["dt"] and um == o.meta["umach"], n
        got = box.download()
        assert box.clock() == (2, o.meta["timc"], o.meta["timt"])
    for k in H.FIELDS:
        assert np.array_equal(got[k], o.arrays[k]), (k, np.abs(got[k] - o.arrays[k]).max())


def test_256_mixed_vs_fp64_1000_steps():
    say("case 256x256x128 mixed vs fp64, 1000 steps")
    meta, st = wl().stratified_box(256, 256, 128)
    dt0 = first_dt(meta, st)
    ref, _, um64, c64 = run(meta, st, "fp64", 1000, fixed_dt=dt0)
    mix, _, umm, cm = run(meta, st, "mixed", 1000, fixed_dt=dt0)
    assert c64 == cm and c64[0] == 1000
    assert abs(umm - um64) <= 1e-6 * um64
    check_errors(meta, mix, ref)


def test_512_mixed_vs_fp64_500_steps():
    say("case 512x512x256 mixed vs fp64, 500 steps")
    meta, st = wl().convection_box(512, 512, 256)
    dt0 = first_dt(meta, st)
    ref, _, um64, c64 = run(meta, st, "fp64", 500, fixed_dt=dt0)
    mix, _, umm, cm = run(meta, st, "mixed", 500, fixed_dt=dt0)
    assert c64 == cm and c64[0] == 500
    assert abs(umm - um64) <= 1e-6 * um64
    check_errors(meta, mix, ref)


def test_512_mixed_shift_equivariance():
    """Every tile, z chunk and XCD of the full-size launch sees different
    data after a periodic shift; the result is the shifted one bit for bit."""
    say("case 512x512x256 mixed shift")
    meta, st = wl().convection_box(512, 512, 256, seed=77)
    si, sj = 201, 45

    def shift(a):
        out = np.zeros_like(a)
        out[1:-1, 1:-1, :] = np.roll(np.roll(a[1:-1, 1:-1, :], si, axis=0), sj, axis=1)
        H.ics().periodic_fill(out)
        return out

    r1, d1, _, _ = run(meta, st, "mixed", 3)
    r2, d2, _, _ = run(meta, {k: shift(v) for k, v in st.items()}, "mixed", 3)
    assert d1 == d2
    for k in H.FIELDS:
        assert np.array_equal(shift(r1[k]), r2[k]), k


def test_64_box_ensemble_bitwise_equal_single_boxes():
    say("case 64-box ensemble (autotuned, graph replay) vs single boxes")
    specs = wl().ensemble_specs(64)
    bx = [wl().ensemble_box(s) for s in specs]
    nsteps = 5
    pkg = H.pkg()
    with pkg.Ensemble([m for m, _ in bx], precision="mixed") as ens:
        for b, (_, s) in enumerate(bx):
            ens.upload(b, s)
        cands = ens.autotune(2)
        say("autotuned", ens.tile(), "of", len(cands), "tilings")
        dt, um = ens.steps(nsteps, graph=True)
        for b, (m, s) in enumerate(bx):
            with pkg.Box(m, precision="mixed") as box:
                box.upload(s)
                dtb, umb = box.steps(nsteps)
                ref, clk = box.download(), box.clock()
            assert dt[b] == dtb and um[b] == umb, (b, dt[b], dtb)
            assert ens.clock(b) == clk, b
            got = ens.download(b)
            for k in H.FIELDS:
                assert np.array_equal(got[k], ref[k]), (b, specs[b], k)


def _windows(nz, ny, nx, blk=8, halo=3):
    """Output sub-blocks (z0, y0, x0) of size blk: touching the top and the
    bottom of the volume and in its middle."""
    return [(0, 0, 0), (nz - blk, ny // 2, nx - blk // 2), (nz // 2 - blk // 2, ny - 3, 5)]


def _oracle_block(params, x, z0, y0, x0, blk=8, halo=3):
    """Oracle output on the blk^3 sub-block at (z0, y0, x0) of the whole-volume
    network: the input window carries the 3-voxel receptive field of the
    three 3x3x3 layers (periodic in x, y; clipped at the volume's z ends,
    where the network zero-pads), and only the window's centre is kept."""
    nz, ny, nx, _ = x.shape
    za, zb = max(0, z0 - halo), min(nz, z0 + blk + halo)
    ys = np.arange(y0 - halo, y0 + blk + halo) % ny
    xs = np.arange(x0 - halo, x0 + blk + halo) % nx
    win = x[za:zb][:, ys][:, :, xs]
    y = N.Net(params, bf16_emulation=True).forward(win)
    return y[z0 - za:z0 - za + blk, halo:halo + blk, halo:halo + blk]


def test_closure_forward_128_against_oracle_on_sub_blocks():
    say("case closure forward 128x128x64 vs oracle sub-blocks")
    nz, ny, nx = 64, 128, 128
    rng = np.random.default_rng(128)
    p = N.init_params(8, 32, 7, seed=3)
    x = N.bf16(rng.standard_normal((nz, ny, nx, 8)))
    with H.pkg().Net(8, 32, 7) as net:
        net.set_params(N.flatten(p))
        y = net.forward(x)
    assert np.all(np.isfinite(y))
    for z0, y0, x0 in _windows(nz, ny, nx):
        ref = _oracle_block(p, x, z0, y0, x0)
        ys = np.arange(y0, y0 + 8) % ny
        xs = np.arange(x0, x0 + 8) % nx
        got = y[z0:z0 + 8][:, ys][:, :, xs]
        e = float(np.sqrt(np.mean((got - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))
        say("block", (z0, y0, x0), "rel", e)
        assert e <= 2e-3
        assert np.abs(got - ref).max() <= 2e-2 * np.abs(ref).max()


def test_closure_corrected_step_128_matches_host_evaluation():
    """One closure-corrected step of the coarse 128x128x64 mixed box (the
    256x256x128 workload's physics) at a tiny fixed dt changes every forced
    field by dt * g, with g the network's forcing evaluated on the host from
    the same state."""
    say("case closure-corrected step 128x128x64 vs host evaluation")
    cl = importlib.import_module("rast-3dmhd_amd.closure")
    meta, st = wl().stratified_box(128, 128, 64, seed=31)
    rng = np.random.default_rng(64)
    inner3 = (slice(1, -1), slice(1, -1), slice(1, -1))
    st["ro"][inner3] *= 1.0 + 1e-3 * rng.standard_normal(st["ro"][inner3].shape)
    for k in ("ru", "rv", "rw"):
        st[k][inner3] = st["ro"][inner3] * 1e-2 * rng.standard_normal(st[k][inner3].shape)
    st["rw"][:, :, 1] = st["rw"][:, :, -2] = 0.0
    for k in ("ru", "rv", "rw", "ro"):
        H.ics().periodic_fill(st[k])
    dt = first_dt(meta, st) * 1e-3
    nz = int(meta["npz"])
    x = cl.inputs_of(cl.interior(st), float(meta["gamma"]))
    in_profile = x.mean(axis=(1, 2)).T.copy()
    in_std = (x - in_profile.T[:, None, None, :]).std(axis=(0, 1, 2))
    rms = np.array([np.sqrt(np.mean(st[k][inner3] ** 2)) for k in cl.TARGETS])
    out_std = 1e-3 * rms / dt
    kk = np.arange(nz) / (nz - 1)
    out_profile = np.stack([0.5 * out_std[c] * np.cos(np.pi * kk + c) for c in range(7)])
    with H.pkg().Net(8, 32, 7) as net:
        net.set_params(N.flatten(N.init_params(8, 32, 7, seed=11)))
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
