"""Small box configurations on which the reference is run to write golden
vectors (TEST INFRASTRUCTURE ONLY).

Each entry gives the PARAMETER overrides (3dmhdparam.f names), the PAR
literals (3dmhdset.f slot numbers), the MPI decomposition the reference runs
with, and the harness perturbation.  All boxes use NPEY>=2 so that the
reference's YJACOBI never reads past its local YY array (sub:246-251 with
NPEY=1), and NPEZ>=2 as the reference requires (3dmhdparam.f:11).
"""

# Rotation axis tilted 45 degrees (latitude f-plane); written as a float32
# literal like every PAR value.
ANG45 = "0.7853982E00"

BASELINE_NAME = "bench512"


def baseline_config(npx=512, npy=512, npz=256, npey=8, npez=2):
    """The throughput workload as a reference build: 512x512x256, MHD with a
    45-degree tilted rotation axis, density contrast 30, Ra=1e6, Pr=1, Pm=2,
    Ta=1e6 (parameter mapping of rast-3dmhd_amd/config.physics), run on
    npey*npez MPI ranks (16 host cores by default)."""
    import importlib.util
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    spec = importlib.util.spec_from_file_location(
        "_cfg", os.path.join(os.path.dirname(here), "rast-3dmhd_amd", "config.py"))
    cfg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(cfg)
    ph = cfg.physics(contrast=30.0, ra=1e6, pr=1.0, pm=2.0, ta=1e6, tilt_deg=45.0)
    lit = lambda x: f"{x:.8E}"
    par = {k: lit(v) for k, v in ph.items()}
    par.update({20: "1.0E-04", 21: "1.0E-03", 22: "0.3E00", 23: "0.5E00"})
    return dict(
        params=dict(NPX=npx, NPY=npy, NPZ=npz, NPEY=npey, NPEZ=npez, LMAG=".TRUE.", LROT=".TRUE.",
                    IZCON=1, ITCON=0),
        par=par, ipar={3: 3, 4: 100000},
    )


CONFIGS = {
    # Hydrodynamic, no rotation, stretched z, fixed T top / fixed flux bottom.
    "hydro16": dict(
        params=dict(NPX=16, NPY=16, NPZ=12, NPEY=2, NPEZ=2, IZCON=1, ITCON=0),
        par={1: "1.0E02", 2: "1.0E00", 3: "2.0E00", 4: "4.0E00", 20: "1.0E-03"},
        steps=5, dumps=[1, 2, 5], ampu="2.0D-02", ampb="0.0D0", seed=11,
    ),
    # MHD + tilted rotation, B layer from the reference's own STATIC plus a
    # perturbation, depth-dependent kappa (PZP/SIGMA) so DKAPA != 0.
    "mhdrot16": dict(
        params=dict(NPX=16, NPY=16, NPZ=12, NPEY=2, NPEZ=2, LMAG=".TRUE.", LROT=".TRUE.",
                    IZCON=1, ITCON=0),
        par={1: "1.5E02", 2: "1.0E00", 3: "2.0E00", 4: "4.0E00", 5: "2.0E00", 6: ANG45,
             7: "1.2E02", 8: "5.0E01", 9: "0.6E00", 10: "0.1E00", 11: "3.0E00",
             20: "1.0E-03", 21: "1.0E-01", 22: "0.3E00", 23: "0.5E00"},
        steps=5, dumps=[1, 2, 5], ampu="2.0D-02", ampb="2.0D-02", seed=23,
    ),
    # Same physics, different decomposition: the reference must give the same
    # bits (decomposition invariance of the reference itself).
    "mhdrot16_d43": dict(
        params=dict(NPX=16, NPY=16, NPZ=12, NPEY=4, NPEZ=3, LMAG=".TRUE.", LROT=".TRUE.",
                    IZCON=1, ITCON=0),
        par={1: "1.5E02", 2: "1.0E00", 3: "2.0E00", 4: "4.0E00", 5: "2.0E00", 6: ANG45,
             7: "1.2E02", 8: "5.0E01", 9: "0.6E00", 10: "0.1E00", 11: "3.0E00",
             20: "1.0E-03", 21: "1.0E-01", 22: "0.3E00", 23: "0.5E00"},
        steps=5, dumps=[1, 2, 5], ampu="2.0D-02", ampb="2.0D-02", seed=23, hashes_only=True,
    ),
    # Non-square box, vertical rotation, fixed-flux top (ITCON=1) and fixed T
    # bottom (IZCON=0), stronger stratification.
    "mhdflux24": dict(
        params=dict(NPX=24, NPY=16, NPZ=14, NPEY=2, NPEZ=2, LMAG=".TRUE.", LROT=".TRUE.",
                    IZCON=0, ITCON=1, ZZ2="1.0D+00"),
        par={1: "2.0E02", 2: "5.0E-01", 3: "4.0E00", 4: "8.0E00", 5: "1.0E00",
             7: "2.0E02", 8: "2.0E01", 17: "3.0E00", 18: "2.0E00", 20: "2.0E-03",
             21: "5.0E-02", 22: "0.4E00", 23: "0.3E00"},
        steps=4, dumps=[1, 4], ampu="1.0D-02", ampb="3.0D-02", seed=7,
    ),
    # The CPU-runnable 64x64x32 box (MHD + vertical rotation, stretched z,
    # contrast 10 polytrope, aspect 4), run for the 200 steps of that
    # workload.  Too large to commit in full; the fixture keeps checksums of
    # every dumped step and sample planes of the last.
    "mhdrot64": dict(
        params=dict(NPX=64, NPY=64, NPZ=32, NPEY=4, NPEZ=2, LMAG=".TRUE.", LROT=".TRUE.",
                    IZCON=1, ITCON=0),
        par={1: "3.0E02", 2: "1.0E00", 3: "9.0E00", 4: "1.8E01", 5: "1.0E00",
             7: "3.0E02", 8: "1.0E02", 20: "1.0E-04", 21: "1.0E-02", 22: "0.3E00",
             23: "0.5E00"},
        steps=200, dumps=[1, 3, 50, 100, 200], ampu="1.0D-03", ampb="1.0D-03", seed=5,
    ),
}

# Runs of the reference's own main program (3dmhd.f) with its output files
# (.dat0 per rank, .par, .lis): fresh start from STATIC, a dump after every
# step.  The drop-in STEP and the host driver must write the same bytes.
DRIVER_CONFIGS = {
    # MHD + 45-degree rotation, Bx layer, T noise, depth-dependent kappa.
    "drv12": dict(
        params=dict(NPX=12, NPY=8, NPZ=8, NPEY=2, NPEZ=2, LMAG=".TRUE.", LROT=".TRUE.",
                    IZCON=1, ITCON=0),
        par={1: "1.5E02", 2: "1.0E00", 3: "2.0E00", 4: "4.0E00", 5: "2.0E00", 6: ANG45,
             7: "1.2E02", 8: "5.0E01", 9: "0.6E00", 10: "0.1E00", 11: "3.0E00",
             20: "1.0E-03", 21: "1.0E-01", 22: "0.3E00", 23: "0.5E00"},
        ipar={3: 3, 4: 1}, nsteps=3,
    ),
    # Hydrodynamic, fixed-flux top and fixed-T bottom.
    "drvh12": dict(
        params=dict(NPX=12, NPY=8, NPZ=8, NPEY=2, NPEZ=2, IZCON=0, ITCON=1, ZZ2="1.0D+00"),
        par={1: "2.0E02", 2: "5.0E-01", 3: "4.0E00", 4: "8.0E00", 17: "3.0E00", 18: "2.0E00",
             20: "2.0E-03"},
        ipar={3: 3, 4: 1}, nsteps=3,
    ),
}
