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
    # contrast 10 polytrope, aspect 4).  Too large to commit in full; the
    # fixture keeps checksums and sample planes.
    "mhdrot64": dict(
        params=dict(NPX=64, NPY=64, NPZ=32, NPEY=4, NPEZ=2, LMAG=".TRUE.", LROT=".TRUE.",
                    IZCON=1, ITCON=0),
        par={1: "3.0E02", 2: "1.0E00", 3: "9.0E00", 4: "1.8E01", 5: "1.0E00",
             7: "3.0E02", 8: "1.0E02", 20: "1.0E-04", 21: "1.0E-02", 22: "0.3E00",
             23: "0.5E00"},
        steps=3, dumps=[1, 3], ampu="1.0D-03", ampb="1.0D-03", seed=5,
    ),
}
