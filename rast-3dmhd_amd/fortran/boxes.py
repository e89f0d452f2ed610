"""Writes build/<name>/3dmhdparam.f for the boxes the Fortran drop-in is
built for.  The boxes mirror the small golden-fixture boxes (same physics
switches, sizes and boundary codes) with the single-domain decomposition
NPEY=NPEZ=1 (one process, no MPI); lshr16 and id16 switch on branches that
the GPU step does not have."""

import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import mkparam  # noqa: E402

BOXES = {
    "hydro16": dict(NPX=16, NPY=16, NPZ=12, IZCON=1, ITCON=0),
    "mhdrot16": dict(NPX=16, NPY=16, NPZ=12, LMAG=".TRUE.", LROT=".TRUE.", IZCON=1, ITCON=0),
    # branches the GPU step does not have: the shim must stop before stepping
    "lshr16": dict(NPX=16, NPY=16, NPZ=12, LMAG=".TRUE.", LROT=".TRUE.", IZCON=1, ITCON=0, LSHR=".TRUE."),
    "id16": dict(NPX=16, NPY=16, NPZ=12, IZCON=1, ITCON=0, ID=1, DH="1.0E-02"),
}


def main(names):
    for n in names or list(BOXES):
        d = os.path.join(HERE, "build", n)
        os.makedirs(d, exist_ok=True)
        text = mkparam.param_block(NPEY=1, NPEZ=1, **BOXES[n])
        path = os.path.join(d, "3dmhdparam.f")
        if not os.path.exists(path) or open(path).read() != text:
            with open(path, "w") as f:
                f.write(text)


if __name__ == "__main__":
    main(sys.argv[1:])
