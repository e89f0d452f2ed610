"""The Fortran drop-in: a Fortran driver calls SUBROUTINE STEP (the shim in
rast-3dmhd_amd/fortran/step_shim.f), which drives the HIP library through the
ISO_C_BINDING C-ABI.  After N steps the driver's COMMON arrays and the
/CTIM/, /TRACE/, /ITER/ scalars must equal the reference's own run bit for
bit (golden fixtures)."""

import os
import subprocess

import numpy as np
import pytest

import helpers as H

FBUILD = os.path.join(H.REPO, "rast-3dmhd_amd", "fortran", "build")
SCALARS = ["hx", "h2x", "hy", "h2y", "hz", "h2z", "c13", "c23", "c43",
           "gam1", "gam2", "gam3", "zeta1", "zeta2", "cv", "ocv", "ore", "re", "repr",
           "theta", "grav", "sf", "gamma", "omx", "omz", "orm", "rm", "obeta", "tb", "tp"]
METRICS = ["dxxdx", "d2xxdx2", "ddx", "dyydy", "d2yydy2", "ddy", "dzzdz", "d2zzdz2", "ddz",
           "rkapa", "dkapa"]


def write_replay_input(path, g, state, nsteps):
    lmag = int(g["lmag"])
    hdr = np.array([nsteps, lmag, int(g["lrot"]), int(g["izc"]), int(g["itc"]), int(g["ibc"]),
                    int(g["npx"]), int(g["npy"]), int(g["npz"])], dtype=np.int32)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.array([float(g[k]) for k in SCALARS]).tobytes())
        for k in METRICS:
            f.write(np.ascontiguousarray(g[k], dtype=np.float64).tobytes())
        for k in (H.FIELDS if lmag else H.FIELDS[:5]):
            f.write(np.asarray(state[k], dtype=np.float64).tobytes(order="F"))


def read_replay_output(path, g):
    lmag = int(g["lmag"])
    shape = (int(g["nx"]), int(g["ny"]), int(g["nz"]))
    n = shape[0] * shape[1] * shape[2]
    a = np.fromfile(path, dtype=np.float64)
    names = H.FIELDS if lmag else H.FIELDS[:5]
    st = {k: a[i * n:(i + 1) * n].reshape(shape, order="F") for i, k in enumerate(names)}
    for k in H.FIELDS:
        st.setdefault(k, np.zeros(shape, order="F"))
    dt, umach, timc, timt, nit = a[len(names) * n:]
    return st, dt, umach, timc, timt, int(nit)


def test_replay_inputs_layout(tmp_path):
    """CPU check of the replay file writer against the reader's layout."""
    g = H.load("hydro16")
    st = H.initial_state("hydro16")
    p = tmp_path / "replay.in"
    write_replay_input(p, g, st, 3)
    n = int(g["nx"]) * int(g["ny"]) * int(g["nz"])
    nm = 3 * int(g["nx"]) + 3 * int(g["ny"]) + 5 * int(g["nz"])
    assert os.path.getsize(p) == 9 * 4 + 8 * (30 + nm + 5 * n)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["hydro16", "mhdrot16"])
def test_fortran_dropin_matches_reference(tmp_path, name):
    print("case fortran drop-in", name, flush=True)
    exe = os.path.join(FBUILD, name, "rastreplay")
    assert os.path.exists(exe), "build the Fortran drop-in first (__graft_entry__.build())"
    g = H.load(name)
    n = int(max(g["dumps"]))
    write_replay_input(tmp_path / "replay.in", g, H.initial_state(name), n)
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    st, dt, umach, timc, timt, nit = read_replay_output(tmp_path / "replay.out", g)
    assert nit == n
    assert dt == float(g[f"step{n}_dt"]) and umach == float(g[f"step{n}_umach"])
    assert timc == float(g[f"step{n}_timc"]) and timt == float(g[f"step{n}_timt"])
    assert H.state_sha(st) == str(g[f"step{n}_sha"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["hydro16", "mhdrot16"])
def test_fortran_dropin_mixed_form(tmp_path, name):
    """RAST3DMHD_PRECISION=mixed: the same driver and shim step the
    mixed-precision form; one step agrees with the reference's FP64 step to
    1e-6 relative L2 per field (the B components against the field's RMS
    magnitude), and DT, computed from the same state in FP64, to 1e-12."""
    print("case fortran drop-in mixed", name, flush=True)
    exe = os.path.join(FBUILD, name, "rastreplay")
    assert os.path.exists(exe), "build the Fortran drop-in first (__graft_entry__.build())"
    g = H.load(name)
    write_replay_input(tmp_path / "replay.in", g, H.initial_state(name), 1)
    env = dict(os.environ, RAST3DMHD_PRECISION="mixed")
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120, env=env)
    # Fortran STOP exits with status 0: the output file tells a finished run
    assert r.returncode == 0 and (tmp_path / "replay.out").exists(), r.stdout + r.stderr
    st, dt, umach, timc, timt, nit = read_replay_output(tmp_path / "replay.out", g)
    assert nit == 1
    assert abs(dt - float(g["step1_dt"])) <= 1e-12 * float(g["step1_dt"])
    assert abs(umach - float(g["step1_umach"])) <= 1e-6 * float(g["step1_umach"])
    inner = (slice(1, -1), slice(1, -1), slice(1, -1))
    lmag = int(g["lmag"])
    bscale = np.sqrt(sum(np.mean(np.asarray(g[f"step1_{k}"])[inner] ** 2) for k in ("bx", "by", "bz"))) if lmag else 0
    for k in (H.FIELDS if lmag else H.FIELDS[:5]):
        ref = np.asarray(g[f"step1_{k}"])[inner]
        d = np.sqrt(np.mean((st[k][inner] - ref) ** 2))
        scale = bscale if k in ("bx", "by", "bz") else np.sqrt(np.mean(ref ** 2))
        print(k, d / scale, flush=True)
        assert d <= 1e-6 * scale, (k, d / scale)


def test_shim_rejects_unknown_precision(tmp_path):
    """The shim's precision switch is read on the first STEP call, before any
    GPU work: an unknown value stops the driver with a message."""
    exe = os.path.join(FBUILD, "hydro16", "rastreplay")
    if not os.path.exists(exe):
        pytest.skip("Fortran drop-in not built")
    g = H.load("hydro16")
    write_replay_input(tmp_path / "replay.in", g, H.initial_state("hydro16"), 1)
    env = dict(os.environ, RAST3DMHD_PRECISION="half")
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120, env=env)
    assert "RAST3DMHD_PRECISION is fp64 or mixed" in r.stdout + r.stderr
    assert not (tmp_path / "replay.out").exists()


@pytest.mark.parametrize("box,data", [("lshr16", "mhdrot16"), ("id16", "hydro16")])
def test_fortran_dropin_refuses_branches_in_the_mixed_form(tmp_path, box, data):
    """The shim passes the optional branches (here LSHR, ID) to the library,
    which has them in the FP64 form only: with RAST3DMHD_PRECISION=mixed,
    mhd_create refuses them (MHD_ERR_SCOPE) before any device call and the
    shim stops with the library's message, so this runs without a GPU."""
    exe = os.path.join(FBUILD, box, "rastreplay")
    assert os.path.exists(exe), "build the Fortran drop-in first (__graft_entry__.build())"
    g = H.load(data)
    write_replay_input(tmp_path / "replay.in", g, H.initial_state(data), 1)
    env = dict(os.environ, RAST3DMHD_PRECISION="mixed")
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120, env=env)
    assert "STEP: mhd_create" in r.stdout and "MHD_FP64 form only" in r.stdout, r.stdout + r.stderr
    assert not os.path.exists(tmp_path / "replay.out")
