"""The reference's own Stokes particle meshes (run/mesh, as used by run/stokes/flow-over-ellipsoid and
flow-over-sphere): a prolate spheroid (semi-axes 0.794, 0.397, 0.397; 1540 triangles) and a sphere
(radius 0.5; 3202 triangles) inside an outer sphere of radius 512 (534 and 476 triangles), in the
reference's bounded set-up (the body at rest as wall 1, the flow on the outer wall 2). Mesh files
are copied as fixtures into tests/golden/cases.

* Spheroid: the plain FP64 path against the oracle's tensor on the same mesh, written by
  tests/golden/make_stokes_ref.py (the oracle's serial LSQR takes tens of minutes), at 1e-8: both
  LSQRs stop at the condition limit (istop 4) after ~1100 iterations, a few dozen apart (rounding
  decides where), and the tensors there agree to 2.5e-9; and the
  Oberbeck / Jeffery closed forms of the spheroid in unbounded fluid within 1.5 % (the outer wall
  adds ~0.2 %, the discretisation the rest).
* Sphere: the translation and rotation resistance against the concentric-sphere closed forms
  (Cunningham 1910 / Williams 1915: F = 6 pi mu a U (1 - l^5) / (1 - 9/4 l + 5/2 l^3 - 9/4 l^5 + l^6);
  T = 8 pi mu a^3 Omega / (1 - l^3), l = a / b), within the mesh's 5e-3. The case's and.inp selects
  the LIS solver "-i bicgstab -p jacobi -tol 1.0e-8 -maxiter 1000" instead of LSQR. The bounded
  single-layer system is singular: a constant pressure gives a traction along the normals with no
  velocity, one null direction per closed wall (a 320-triangle sphere of radius 0.5 inside a
  180-triangle sphere of radius 64: the two smallest singular values are 3.6e-7 and 1.7e-9, the
  largest 1.8e3). LSQR returns the
  least-squares solution (istop 4, LSQR's condition limit), while Jacobi-preconditioned BiCGSTAB
  does not converge on it (the oracle's BiCGSTAB, itself pinned by scipy, reaches its iteration
  limit with a residual of 4.5 on that smaller system). The device BiCGSTAB must report that
  failure (status 1 or 2) instead of returning a solution.
"""
import os

import numpy as np
import pytest

from andromeda_amd import api
from andromeda_amd import mesh as M

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = os.path.join(HERE, "golden", "cases")


def _mesh(name):
    return M.read_gmsh(os.path.join(CASES, name))


def test_reference_spheroid_in_sphere_vs_oracle(ctx):
    from test_resistance_problems import ellipsoid_closed_forms
    m = _mesh("ellipsoid211-in-sphere-1024-1041-2117.msh")
    ref = np.load(os.path.join(HERE, "golden", "ellipsoid211_in_sphere_ft.npz"))
    r = ctx.resistance(m, form=api.FORM_BOUNDED, precision=api.PREC_FP64, eps=1e-12)
    rel = np.abs(r.FT - ref["FT"]).max() / np.abs(ref["FT"]).max()
    print(f"spheroid in sphere: vs oracle {rel:.2e}; its {r.nit.tolist()} (oracle {ref['nits'].tolist()}); "
          f"diag {np.diag(r.FT)}")
    assert r.status in (0, 4) and rel < 1e-8
    nd = np.unique(m.con[m.bcid == 1])
    axes = np.abs(m.xyz[nd]).max(axis=0)
    K, Om, _ = ellipsoid_closed_forms(axes)
    assert np.abs(np.diag(r.FT[:3, :3]) / K - 1).max() < 0.015
    assert np.abs(np.diag(r.FT[3:, 3:]) / Om - 1).max() < 0.015


def test_reference_sphere_in_sphere_closed_form_and_bicgstab(ctx, orc):
    m = _mesh("sphere-in-sphere-1024-1843-3678-tri.msh")
    a, b = 0.5, 512.0
    lam = a / b
    Kc = 6 * np.pi * a * (1 - lam ** 5) / (1 - 2.25 * lam + 2.5 * lam ** 3 - 2.25 * lam ** 5 + lam ** 6)
    Oc = 8 * np.pi * a ** 3 / (1 - lam ** 3)
    r = ctx.resistance(m, form=api.FORM_BOUNDED, precision=api.PREC_FP64, eps=1e-12)
    K, Om = np.diag(r.FT[:3, :3]), np.diag(r.FT[3:, 3:])
    print(f"sphere in sphere: K/closed {K / Kc}, Omega/closed {Om / Oc}, its {r.nit.tolist()}, status {r.status}")
    assert r.status in (0, 4)
    assert np.all(np.abs(K / Kc - 1) < 5e-3) and np.all(np.abs(Om / Oc - 1) < 5e-3)
    # the same system with BiCGSTAB + Jacobi: centroid rows of the integrals, the reference's
    # bounded system (formQrow), the outer wall moving with e_x
    nn, ne = m.nnodes, m.nelem
    T, G, _, _, _ = ctx.assemble_stokes_rows(m, nn, nn + ne)
    sysM, rhsM = orc.stokes_dirichlet_system(m, T, G, row_offset=0)
    del T, G
    node_wall = np.zeros(nn, dtype=np.int32)
    for e in range(ne):
        node_wall[m.con[e]] = np.maximum(node_wall[m.con[e]], m.bcid[e])
    ub = np.zeros(3 * nn)
    ub[:nn] = (node_wall == 2).astype(np.float64)
    x, info = ctx.bicgstab_dense(sysM, rhsM @ ub, tol=1e-8, maxit=1000)
    _, area, _ = orc.geometry(m)
    Fx = np.sum(area[m.bcid == 1] * x[:ne, 0][m.bcid == 1])
    print(f"BiCGSTAB: its {info['nit']}, status {info['status']}, Fx {Fx:.10f} vs LSQR {r.FT[0, 0]:.10f}")
    assert info["status"][0] in (1, 2)
