"""Writes the golden fixtures in this directory from runs of the reference.

Runs only in the build container, where /root/reference and amdflang exist:
for each box of oracle/refconfigs.py it compiles the reference with the
recipe in oracle/build_ref.py, runs the harness (oracle/refharness.f) under
MPI, stitches the per-rank dumps into single-domain padded arrays and saves

  <name>.npz   scalars, global metric arrays, bare polytrope profiles, and
               the parameters of the optional FLUXES/BCON branches, and
               (small boxes) full initial state, one FLUXES evaluation and the
               states after the first and last dumped steps; every dumped
               step also gets a SHA-256 of its full padded state and its
               DT / UMACH / TIMC / TIMT / NIT.

Large boxes keep only checksums and a few sample planes.

Usage: python tests/golden/make_golden.py [name ...]
"""

from __future__ import annotations

import hashlib
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))


def _load(path, name):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


build_ref = _load(os.path.join(REPO, "oracle", "build_ref.py"), "build_ref")
CONFIGS = build_ref.refconfigs.CONFIGS

FIELDS = ["ru", "rv", "rw", "ro", "tt", "bx", "by", "bz"]
FLUX_FIELDS = ["fu", "fv", "fw", "fr", "ft", "ww1", "ww2", "ww3", "uu", "vv", "ww", "roinv"]
CONST_KEYS = ["hx", "h2x", "hy", "h2y", "hz", "h2z", "c13", "c23", "c43",
              "gam1", "gam2", "gam3", "zeta1", "zeta2",
              "cv", "ocv", "ore", "re", "repr", "theta", "grav", "sf", "gamma",
              "omx", "omz", "orm", "rm", "obeta", "tb", "tp", "dzf", "ampt"]
OPT_KEYS = ["id", "dh", "dv", "tu", "dztu", "dztb", "rlax", "tstart", "toff",
            "xp", "yp", "zp", "tc", "qfh", "hh", "xmax", "ymax", "zmax", "lrem", "lshr"]
# Boxes above this many padded points keep checksums only.
FULL_LIMIT = 40_000


def state_sha(fields: dict) -> str:
    h = hashlib.sha256()
    for k in FIELDS:
        h.update(np.asarray(fields[k], dtype=np.float64).tobytes(order="F"))
    return h.hexdigest()


def _read_rank(path, count):
    a = np.fromfile(path, dtype=np.float64)
    if a.size < count:
        raise ValueError(f"{path}: {a.size} values, expected >= {count}")
    return a


def stitch(run_dir: str, cfg: dict) -> dict:
    p = cfg["params"]
    npx, npy, npz = int(p["NPX"]), int(p["NPY"]), int(p["NPZ"])
    npey, npez = int(p["NPEY"]), int(p["NPEZ"])
    nry, nrz = npy // npey, npz // npez
    nx, nyl, nzl = npx + 2, nry + 2, nrz + 2
    gx, gy, gz = npx + 2, npy + 2, npz + 2
    nloc = nx * nyl * nzl
    out = {"npx": npx, "npy": npy, "npz": npz, "npey": npey, "npez": npez,
           "nx": gx, "ny": gy, "nz": gz,
           "lmag": int(str(p.get("LMAG", ".FALSE.")).upper() == ".TRUE."),
           "lrot": int(str(p.get("LROT", ".FALSE.")).upper() == ".TRUE.")}

    def place(glob, loc, py, pz):
        j0, k0 = py * nry, pz * nrz
        glob[:, j0:j0 + nyl, k0:k0 + nzl] = loc

    ranks = [(r % npey, r // npey, r) for r in range(npey * npez)]
    # metadata and metrics
    mx = {k: np.zeros(n) for k, n in [("exx", gx), ("dxxdx", gx), ("d2xxdx2", gx), ("ddx", gx),
                                       ("wyy", gy), ("dyydy", gy), ("d2yydy2", gy), ("ddy", gy),
                                       ("zee", gz), ("dzzdz", gz), ("d2zzdz2", gz), ("ddz", gz),
                                       ("rkapa", gz), ("dkapa", gz)]}
    prof_t, prof_r, prof_bx = np.zeros(gz), np.zeros(gz), np.zeros(gz)
    for py, pz, r in ranks:
        m = np.fromfile(os.path.join(run_dir, f"meta_r{r:04d}.bin"), dtype=np.float64)
        hdr, m = m[:10], m[10:]
        assert (int(hdr[0]), int(hdr[1]), int(hdr[2])) == (nx, nyl, nzl)
        assert (int(hdr[3]), int(hdr[4])) == (py, pz)
        out["izc"], out["itc"], out["ibc"] = int(hdr[7]), int(hdr[8]), int(hdr[9])
        consts, m = m[:len(CONST_KEYS)], m[len(CONST_KEYS):]
        for k, v in zip(CONST_KEYS, consts):
            out[k] = float(v)
        seg = {}
        for k, n in [("exx", nx), ("dxxdx", nx), ("d2xxdx2", nx), ("ddx", nx),
                     ("wyy", nyl), ("dyydy", nyl), ("d2yydy2", nyl), ("ddy", nyl),
                     ("zee", nzl), ("dzzdz", nzl), ("d2zzdz2", nzl), ("ddz", nzl),
                     ("rkapa", nzl), ("dkapa", nzl)]:
            seg[k], m = m[:n], m[n:]
        assert m.size == 0
        # Overlapping ghost metrics of neighbouring ranks are computed from
        # the same computational coordinate, so plain placement is exact.
        for k in ("exx", "dxxdx", "d2xxdx2", "ddx"):
            mx[k][:] = seg[k]
        for k in ("wyy", "dyydy", "d2yydy2", "ddy"):
            mx[k][py * nry:py * nry + nyl] = seg[k]
        for k in ("zee", "dzzdz", "d2zzdz2", "ddz", "rkapa", "dkapa"):
            mx[k][pz * nrz:pz * nrz + nzl] = seg[k]
    # Profiles: full local columns first, then each rank's interior planes,
    # because z-ghost planes of B are not set by STATIC.
    for pass_ in (0, 1):
        for py, pz, r in ranks:
            if py != 0:
                continue
            pr = np.fromfile(os.path.join(run_dir, f"prof_r{r:04d}.bin"), dtype=np.float64)
            bc = np.fromfile(os.path.join(run_dir, f"bcol_r{r:04d}.bin"), dtype=np.float64)
            lo, sl = (0, slice(None)) if pass_ == 0 else (1, slice(1, -1))
            k0 = pz * nrz + lo
            for dst, src in ((prof_t, pr[:nzl]), (prof_r, pr[nzl:]), (prof_bx, bc)):
                v = src[sl]
                dst[k0:k0 + v.size] = v
    out.update(mx)
    out["prof_t"], out["prof_r"], out["prof_bx"] = prof_t, prof_r, prof_bx
    # parameters of the optional branches (the same on every rank)
    opt = os.path.join(run_dir, "opt_r0000.bin")
    if os.path.exists(opt):
        v = np.fromfile(opt, dtype=np.float64)
        assert v.size == len(OPT_KEYS)
        for k, x in zip(OPT_KEYS, v):
            out[k] = int(x) if k in ("id", "lrem", "lshr") else float(x)

    def fields_from(prefix, names, nextra=0):
        glob = {k: np.zeros((gx, gy, gz), order="F") for k in names}
        extra = None
        # two passes: full local arrays, then interiors, so overlapping ghost
        # layers always end with the owning rank's values
        for pass_ in (0, 1):
            for py, pz, r in ranks:
                a = np.fromfile(os.path.join(run_dir, f"{prefix}_r{r:04d}.bin"), dtype=np.float64)
                if nextra:
                    extra = a[len(names) * nloc:]
                for n, k in enumerate(names):
                    loc = a[n * nloc:(n + 1) * nloc].reshape((nx, nyl, nzl), order="F")
                    if pass_ == 0:
                        place(glob[k], loc, py, pz)
                    else:
                        j0, k0 = py * nry, pz * nrz
                        glob[k][:, j0 + 1:j0 + nyl - 1, k0 + 1:k0 + nzl - 1] = loc[:, 1:-1, 1:-1]
        return glob, extra

    out["init"], _ = fields_from("init", FIELDS)
    out["flux"], _ = fields_from("flux", FLUX_FIELDS)
    out["steps"] = {}
    for n in cfg["dumps"]:
        st, sc = fields_from(f"step{n:04d}", FIELDS, nextra=5)
        out["steps"][n] = (st, sc)
    return out


def save(name: str, cfg: dict, st: dict) -> str:
    rec = {}
    for k, v in st.items():
        if k in ("init", "flux", "steps"):
            continue
        rec[k] = np.asarray(v)
    full = st["nx"] * st["ny"] * st["nz"] <= FULL_LIMIT and not cfg.get("hashes_only")
    rec["full"] = np.asarray(int(full))
    rec["init_sha"] = np.asarray(state_sha(st["init"]))
    rec["dumps"] = np.asarray(cfg["dumps"])
    flux_h = hashlib.sha256()
    for k in FLUX_FIELDS[:8]:
        flux_h.update(np.ascontiguousarray(st["flux"][k][1:-1, 1:-1, 1:-1]).tobytes())
    rec["flux_sha"] = np.asarray(flux_h.hexdigest())
    if full:
        for k in FIELDS:
            rec[f"init_{k}"] = st["init"][k]
        for k in FLUX_FIELDS[:8]:
            rec[f"flux_{k}"] = st["flux"][k]
    else:
        gz = st["nz"]
        planes = [1, gz // 2, gz - 2]
        rec["sample_k"] = np.asarray(planes)
    last = cfg["dumps"][-1]
    for n, (fields, sc) in st["steps"].items():
        rec[f"step{n}_sha"] = np.asarray(state_sha(fields))
        for i, k in enumerate(["dt", "umach", "timc", "timt", "nit"]):
            rec[f"step{n}_{k}"] = np.asarray(sc[i])
        if full and n == cfg["dumps"][0]:
            for k in FIELDS:
                rec[f"step{n}_{k}"] = fields[k]
        elif not full and n == last:
            gz = st["nz"]
            planes = [1, gz // 2, gz - 2]
            for k in FIELDS:
                rec[f"step{n}_{k}_planes"] = fields[k][:, :, planes]
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **rec)
    return path


def main(names):
    if not build_ref.available():
        sys.exit("reference sources or amdflang not available; fixtures are committed")
    for name in names or list(CONFIGS):
        cfg = CONFIGS[name]
        build_ref.build(name, cfg)
        run = build_ref.run_harness(name, cfg)
        st = stitch(run, cfg)
        path = save(name, cfg, st)
        print(f"{name}: {path} ({os.path.getsize(path) / 1024:.0f} KiB)", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
