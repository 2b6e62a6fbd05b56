"""Test infrastructure, not product code: a CPU restatement of the reference's multi-subdomain Laplace
solve, used only as the checker of the HIP path (tests and smoke()).

It follows the reference routine by routine:
  subdomains(...)       SetUpSubdomains          src/lib_sdBem.f90:319-430 (element, q-node and node
                        lists; a node takes the wall id of the last listed element touching it)
                        (node wall ids given by the EDGE / CORNER rules of the input override it)
  assemble(...)         sdFormIntegralMatrices   src/lib_int.f90:225-323 with sdIntRowU / sdIntRowQ,
                        src/lib_sdBem.f90:438-733: the subdomain's H and G, computed by the C oracle
                        (orc_assemble_laplace) on the mesh of the subdomain's walls, whose node and
                        element order is the reference's node and q-node list order
  classify(...)         sdSetUpXandB             src/lib_sdBem.f90:4-107
  system(...)           sdSetUpSysMrhsM_CRS      src/lib_CRS.f90:148-343 (dense instead of CSR: the
                        reference's CSR holds every entry of the dense blocks)
  solve(...)            sdSolveCRS               src/lib_CRS.f90:4-142 (b = rhsM (-b), lsqr_precon,
                        lsqr_solve with the hot start x, then the scatter back to u and q)
  wall_flux(...)        CalFluxIntegralWalls     src/lib_mesh.f90:851-882
Boundary condition kinds: 1 function (u known), 2 flux (q known), 3 contact (both unknown).
"""
from __future__ import annotations

import numpy as np

import pyoracle

FUNCTION, FLUX, CONTACT = 1, 2, 3
NOT_SET = -999


class Subdomain:
    """walls (ids), normmul per wall, diff, node_wall (nnodes wall-id overrides, 0 = none, or None)."""

    def __init__(self, walls, normmul=None, diff=1.0, node_wall=None):
        self.walls = [int(w) for w in walls]
        self.normmul = [1.0] * len(self.walls) if normmul is None else [float(v) for v in normmul]
        self.diff = float(diff)
        self.node_wall = node_wall


def subdomains(mesh, sds):
    """Element lists (walls in the subdomain's order, elements ascending within a wall), q-node lists
    and node lists (ascending node ids) with the node wall ids, as SetUpSubdomains builds them."""
    out = []
    bcid = np.asarray(mesh.bcid)
    con = np.asarray(mesh.con)
    for sd in sds:
        elems = np.concatenate([np.nonzero(bcid == w)[0] for w in sd.walls]).astype(np.int64)
        tmp = np.zeros(mesh.nnodes, dtype=np.int64)
        for e in elems:  # last writer wins
            for v in con[e]:
                if v >= 0:
                    tmp[v] = bcid[e]
        nodes = np.nonzero(tmp)[0]
        node_wall = tmp[nodes].copy()
        if getattr(sd, "node_wall", None) is not None:  # EDGE / CORNER overrides of the input
            ov = np.asarray(sd.node_wall)[nodes]
            node_wall = np.where(ov != 0, ov, node_wall)
        out.append({"elems": elems, "qwall": bcid[elems].copy(), "nodes": nodes, "node_wall": node_wall})
    return out


def submesh(mesh, sd, L):
    """The mesh of one subdomain's walls: nodes in node-list order, elements in element-list order,
    wall ids renumbered 1..len(walls) in the subdomain's order (with its NORMMUL factors)."""
    from andromeda_amd.mesh import Mesh
    remap = -np.ones(mesh.nnodes, dtype=np.int64)
    remap[L["nodes"]] = np.arange(L["nodes"].size)
    con = np.asarray(mesh.con)[L["elems"]]
    con_s = np.where(con >= 0, remap[np.maximum(con, 0)], -1).astype(np.int32)
    wid = {w: k + 1 for k, w in enumerate(sd.walls)}
    bc = np.array([wid[int(w)] for w in L["qwall"]], dtype=np.int32)
    nm = np.ones(len(sd.walls)) if sd.normmul is None else np.array(sd.normmul, dtype=np.float64)
    return Mesh(np.asarray(mesh.xyz)[L["nodes"]].copy(), con_s, bc, "subdomain"), nm


def assemble(mesh, sds, lists, itri=5, iquad=4):
    """H (nsp x nnodes_sd) and G (nsp x nelem_sd) of every subdomain, and its (fspCerr, QspCerr)."""
    out = []
    for sd, L in zip(sds, lists):
        sm, nm = submesh(mesh, sd, L)
        H, G, err = pyoracle.assemble_laplace(sm, normmul=nm, itri=itri, iquad=iquad)
        out.append((H, G, err))
    return out


def classify(mesh, sds, lists, wall_bc):
    """col (per node), qcol (per element), nx unknowns, nb knowns, neq equations (sdSetUpXandB);
    the column numbers are 0-based here."""
    col = np.full(mesh.nnodes, NOT_SET, dtype=np.int64)
    qcol = np.full(mesh.nelem, NOT_SET, dtype=np.int64)
    nx = nb = neq = 0
    for L in lists:
        for isp, w in zip(L["nodes"], L["node_wall"]):
            bc = wall_bc[w - 1]
            if bc == FUNCTION:
                col[isp] = nb
                nb += 1
            elif bc == FLUX:
                col[isp] = nx
                nx += 1
                neq += 1
            elif bc == CONTACT:
                neq += 1
                if col[isp] == NOT_SET:
                    col[isp] = nx
                    nx += 1
        for isp, w in zip(L["elems"], L["qwall"]):
            bc = wall_bc[w - 1]
            if bc == FUNCTION:
                qcol[isp] = nx
                nx += 1
                neq += 1
            elif bc == FLUX:
                qcol[isp] = nb
                nb += 1
            elif bc == CONTACT:
                neq += 1
                if qcol[isp] == NOT_SET:
                    qcol[isp] = nx
                    nx += 1
    return col, qcol, nx, nb, neq


def system(mesh, sds, lists, mats, wall_bc, cls):
    """Dense sysM (neq x nx) and rhsM (neq x nb) in the reference's row order (sdSetUpSysMrhsM_CRS)."""
    col, qcol, nx, nb, neq = cls
    sysM = np.zeros((neq, nx))
    rhsM = np.zeros((neq, nb))
    row = 0
    for sd, L, (H, G, _) in zip(sds, lists, mats):
        nn = L["nodes"].size
        bcn = np.array([wall_bc[w - 1] for w in L["node_wall"]])
        bcq = np.array([wall_bc[w - 1] for w in L["qwall"]])
        cn = col[L["nodes"]]
        cq = qcol[L["elems"]]
        wpos = {w: k for k, w in enumerate(sd.walls)}
        multi = np.array([list(sd.normmul)[wpos[int(w)]] if sd.normmul is not None else 1.0 for w in L["qwall"]])
        gmul = multi / sd.diff
        rows = [r for r in range(nn) if bcn[r] in (FLUX, CONTACT)]
        rows += [nn + r for r in range(L["elems"].size) if bcq[r] in (FUNCTION, CONTACT)]
        for r in rows:
            hv, gv = H[r], gmul * G[r]
            kn = bcn == FUNCTION
            rhsM[row, cn[kn]] = hv[kn]
            sysM[row, cn[~kn]] = hv[~kn]
            kq = bcq == FLUX
            rhsM[row, cq[kq]] = gv[kq]
            sysM[row, cq[~kq]] = gv[~kq]
            row += 1
    assert row == neq
    return sysM, rhsM


def solve(mesh, sds, lists, mats, wall_bc, u, q, eps=1e-15, maxit=500):
    """sdSolveCRS: returns (u, q, info) with the unknowns replaced by the LSQR solution."""
    cls = classify(mesh, sds, lists, wall_bc)
    col, qcol, nx, nb, neq = cls
    sysM, rhsM = system(mesh, sds, lists, mats, wall_bc, cls)
    u = np.array(u, dtype=np.float64)
    q = np.array(q, dtype=np.float64)
    b = np.zeros(nb)
    x = np.zeros(nx)
    for L in lists:
        for isp, w in zip(L["nodes"], L["node_wall"]):
            bc = wall_bc[w - 1]
            if bc == FUNCTION:
                b[col[isp]] = u[isp]
            else:
                x[col[isp]] = u[isp]
        for isp, w in zip(L["elems"], L["qwall"]):
            bc = wall_bc[w - 1]
            if bc == FLUX:
                b[qcol[isp]] = q[isp]
            else:
                x[qcol[isp]] = q[isp]
    rhs = rhsM @ (-b)
    ia, ja, a = pyoracle.csr_from_dense(sysM)
    qs = pyoracle.lsqr_precon(ia, ja, a, neq, nx)
    xs, info = pyoracle.lsqr_solve(ia, ja, a, neq, nx, qs, rhs, x0=x, maxit=maxit, eps=eps)
    for L in lists:
        for isp, w in zip(L["nodes"], L["node_wall"]):
            bc = wall_bc[w - 1]
            u[isp] = b[col[isp]] if bc == FUNCTION else xs[col[isp]]
        for isp, w in zip(L["elems"], L["qwall"]):
            bc = wall_bc[w - 1]
            q[isp] = b[qcol[isp]] if bc == FLUX else xs[qcol[isp]]
    info.update({"neq": neq, "nx": nx, "nb": nb})
    return u, q, info


def wall_flux(mesh, q, nwalls):
    """Sum of area * q over the elements of every wall (CalFluxIntegralWalls)."""
    area = pyoracle.geometry(mesh)[1]
    bcid = np.asarray(mesh.bcid)
    return np.array([float(np.sum(area[bcid == w] * q[bcid == w])) for w in range(1, nwalls + 1)])
