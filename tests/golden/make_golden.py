"""Writes golden vectors of the reference's own compiled routines (oracle/_ref, built from the
reference's Fortran sources by oracle/Makefile): the expanded Wandzura-Xiao tables
(src/mod_trikotnik.f90:282-781) and the Gauss-Legendre tables (src/mod_gauss.f90:33-391).
Run in the build container: python tests/golden/make_golden.py

The analytic singular integrals of the same build are not written: siNormVector
(src/mod_singInt.f90:1141-1155) reads an uninitialised accumulator and this build returns NaN.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import pyoracle  # noqa: E402

D = C.POINTER(C.c_double)


def main():
    ref = pyoracle.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref not built")
    tabs = {}
    for itri in range(1, 6):
        b = [np.zeros(126) for _ in range(4)]
        n = ref.ref_tri_rule(itri, *[v.ctypes.data_as(D) for v in b])
        tabs[f"rule{itri}"] = np.stack([v[:n] for v in b], 1)
    gi, ome = np.zeros(156), np.zeros(156)
    ng1, ng2 = np.zeros(8, np.int32), np.zeros(8, np.int32)
    ref.ref_gauss(gi.ctypes.data_as(D), ome.ctypes.data_as(D), ng1.ctypes.data_as(C.POINTER(C.c_int32)),
                  ng2.ctypes.data_as(C.POINTER(C.c_int32)))
    tabs["gauss_gi"], tabs["gauss_ome"], tabs["gauss_ng1"], tabs["gauss_ng2"] = gi, ome, ng1, ng2
    np.savez(os.path.join(HERE, "ref_tables.npz"), **tabs)
    print("wrote ref_tables.npz")


if __name__ == "__main__":
    main()
