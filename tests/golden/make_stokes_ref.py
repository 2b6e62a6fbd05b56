"""Writes tests/golden/ellipsoid211_in_sphere_ft.npz: the oracle's resistance tensor (bounded form, the
reference's set-up: the body at rest as wall 1, the flow on the far outer sphere as wall 2; LSQR eps
1e-12) of the reference's own mesh run/mesh/ellipsoid211-in-sphere-1024-1041-2117.msh (copied as the
fixture tests/golden/cases/ellipsoid211-in-sphere-1024-1041-2117.msh). The oracle's LSQR is serial,
so this takes tens of minutes on a CPU; the GPU test compares the plain FP64 path with the file.

    python tests/golden/make_stokes_ref.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle  # noqa: E402
from andromeda_amd import mesh as M  # noqa: E402

MESH = os.path.join(HERE, "cases", "ellipsoid211-in-sphere-1024-1041-2117.msh")
OUT = os.path.join(HERE, "ellipsoid211_in_sphere_ft.npz")

if __name__ == "__main__":
    m = M.read_gmsh(MESH)
    t0 = time.time()
    FT, nits, ierr = pyoracle.resistance(m, exterior=False, eps=1e-12)
    np.savez(OUT, FT=FT, nits=nits, ierr=ierr, eps=1e-12, nelem=m.nelem)
    print(f"{m.nelem} triangles: FT diag {np.diag(FT)}, LSQR its {nits.tolist()}, ierr {ierr}, "
          f"{time.time() - t0:.0f} s")
