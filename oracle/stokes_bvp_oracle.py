"""Test infrastructure, not product code: a CPU restatement of the reference's general Stokes solve
(one subdomain, velocity or traction known per wall and per component), used only as the checker of
the HIP path.

It follows the reference routine by routine:
  lists / assembly   SetUpSubdomains + sdFormIntegralMatricesStokes (src/lib_Stokes.f90:1391-1555):
                     the C oracle's 18 matrices of the mesh of the subdomain's walls (as in
                     laplace_oracle.py), bounded form (no exterior free term)
  classify           sdSetUpXandB per component (src/lib_sdBem.f90:4-107), then StokesBigSystemXB
                     (src/lib_Stokes.f90:1180-1276): the three component systems stacked, unknowns of
                     component k after those of components < k, knowns likewise
  system             stokesFormCRSsysMrhsM with formUrow / formQrow (src/lib_Stokes.f90:475-637,
                     src/lib_StokesFull.f90:4-160): rows in order (nodes, then elements; components
                     inside), -T and -(normmul/diff) G for unknowns, +T and +(normmul/diff) G for knowns
  solve              formStokesLeftRightHandSideVectors, CRSxV (b = rhsM b), lsqr_precon, lsqr_solve
                     with the hot start, StokesDistributeUnknowns (src/lib_StokesFull.f90:379-501,
                     src/lib_Stokes.f90:417-469)
Arrays u (3, nnodes) and q (3, nelem) are per component.
"""
from __future__ import annotations

import numpy as np

import laplace_oracle as LO
import pyoracle

SYM = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]  # (component, component) -> {xx, xy, xz, yy, yz, zz}


def assemble(mesh, sd, L, itri=5):
    sm, nm = LO.submesh(mesh, sd, L)
    T, G, _, _, err = pyoracle.assemble_stokes(sm, normmul=nm, itri=itri)
    return T, G, err


def classify(mesh, sd, L, wall_bc3):
    """Per component: (col, qcol, nx, nb, neq); then the stacked columns."""
    per = [LO.classify(mesh, [sd], [L], wall_bc3[k]) for k in range(3)]
    off_x = np.cumsum([0] + [p[2] for p in per])
    off_b = np.cumsum([0] + [p[3] for p in per])
    return per, off_x, off_b


def solve(mesh, sd, L, T, G, wall_bc3, u, q, eps=1e-15, maxit=500):
    per, off_x, off_b = classify(mesh, sd, L, wall_bc3)
    nx, nb = int(off_x[-1]), int(off_b[-1])
    nodes, elems = L["nodes"], L["elems"]
    nn = nodes.size
    bcn = np.array([[wall_bc3[k][w - 1] for w in L["node_wall"]] for k in range(3)])
    bcq = np.array([[wall_bc3[k][w - 1] for w in L["qwall"]] for k in range(3)])
    # stacked column of every (component, node) and (component, element)
    cn = np.zeros((3, nn), dtype=np.int64)
    cq = np.zeros((3, elems.size), dtype=np.int64)
    for k in range(3):
        col, qcol = per[k][0], per[k][1]
        cn[k] = np.where(bcn[k] == LO.FUNCTION, col[nodes] + off_b[k], col[nodes] + off_x[k])
        cq[k] = np.where(bcq[k] == LO.FLUX, qcol[elems] + off_b[k], qcol[elems] + off_x[k])
    wpos = {w: i for i, w in enumerate(sd.walls)}
    nm = sd.normmul if sd.normmul is not None else [1.0] * len(sd.walls)
    gmul = np.array([nm[wpos[int(w)]] for w in L["qwall"]]) / sd.diff
    rows = []
    for r in range(nn):
        for een in range(3):
            if bcn[een][r] != LO.FUNCTION:
                rows.append((een, r))
    for r in range(elems.size):
        for een in range(3):
            if bcq[een][r] != LO.FLUX:
                rows.append((een, nn + r))
    sysM = np.zeros((len(rows), nx))
    rhsM = np.zeros((len(rows), nb))
    for row, (een, r) in enumerate(rows):
        for en in range(3):
            tv = T[SYM[een][en]][r]
            gv = gmul * G[SYM[een][en]][r]
            kn = bcn[en] == LO.FUNCTION
            rhsM[row, cn[en][kn]] = tv[kn]
            sysM[row, cn[en][~kn]] = -tv[~kn]
            kq = bcq[en] == LO.FLUX
            rhsM[row, cq[en][kq]] = gv[kq]
            sysM[row, cq[en][~kq]] = -gv[~kq]
    u = np.array(u, dtype=np.float64)
    q = np.array(q, dtype=np.float64)
    b = np.zeros(nb)
    x = np.zeros(nx)
    for k in range(3):
        kn = bcn[k] == LO.FUNCTION
        b[cn[k][kn]] = u[k][nodes[kn]]
        x[cn[k][~kn]] = u[k][nodes[~kn]]
        kq = bcq[k] == LO.FLUX
        b[cq[k][kq]] = q[k][elems[kq]]
        x[cq[k][~kq]] = q[k][elems[~kq]]
    rhs = rhsM @ b
    ia, ja, a = pyoracle.csr_from_dense(sysM)
    qs = pyoracle.lsqr_precon(ia, ja, a, len(rows), nx)
    xs, info = pyoracle.lsqr_solve(ia, ja, a, len(rows), nx, qs, rhs, x0=x, maxit=maxit, eps=eps)
    for k in range(3):
        kn = bcn[k] == LO.FUNCTION
        u[k][nodes[~kn]] = xs[cn[k][~kn]]
        kq = bcq[k] == LO.FLUX
        q[k][elems[~kq]] = xs[cq[k][~kq]]
    info.update({"neq": len(rows), "nx": nx, "nb": nb})
    return u, q, info


def wall_force(mesh, q, nwalls):
    """(3, nwalls): sum of A_e q_e over the elements of every wall, per component."""
    return np.array([LO.wall_flux(mesh, q[k], nwalls) for k in range(3)])
