"""The reference's input files for the Laplace driver: `and.inp` (mesh, BIC file, quadrature) and the
BIC file (equations, subdomains, NORMMUL, SOLVER, BOUNDARY / INITIAL conditions, EDGE / CORNER
rules), read as the reference reads them, and the driver itself on the device
(`solve_laplace_case`).

Reference routines restated here (host-side input handling; the solve is andr_laplace_solve):
  read_lines        rOneTL                 src/lib_mesh.f90:1804-1826 (a line with '#' anywhere is skipped)
  read_bic          ReadsdBIC + ReadBIC    src/lib_mesh.f90:990-1237, RBiC :1577-1624; EDGE / CORNER
                    from HandleCornersEdges src/lib_sdBem.f90:112-300
  bic_value         GetBiCValue            src/lib_mesh.f90:1770-1800
  conditions        sdApplyInitalCond      src/lib_mesh.f90:1628-1692
The Stokes driver's general problem (three equations, the velocity components) runs through
`solve_stokes_case` (andr_stokes_solve).
"""
from __future__ import annotations

import os
import shlex
from dataclasses import dataclass, field

import numpy as np

from . import api
from .mesh import Mesh

KINDS = {"function": api.BC_FUNCTION, "flux": api.BC_FLUX, "contact": api.BC_CONTACT}
NPARAMS = {"const": 1, "lin": 4, "quad": 10}


def read_lines(path: str) -> list:
    """Non-empty lines without a '#' character, in file order (rOneTL)."""
    out = []
    with open(path) as fh:
        for ln in fh:
            ln = ln.rstrip("\n")
            if ln.strip() and "#" not in ln:
                out.append(ln)
    return out


def _words(line: str) -> list:
    return shlex.split(line.replace(",", " "))


@dataclass
class Condition:
    kind: int          # api.BC_*
    form: str          # const / lin / quad
    params: np.ndarray


@dataclass
class Solver:
    type: int = 0
    pret: int = 2
    prep: int = 2
    maxit: int = 500   # ReadBIC defaults, src/lib_mesh.f90:1128-1133
    stopt: int = 5
    eps: float = 1e-15


@dataclass
class BicProblem:
    equations: list
    sd_names: list
    subdomains: list                       # api.Subdomain per subdomain (walls in wall-index order)
    boundary: dict                         # (equation, wall id) -> Condition
    initial: dict                          # (equation, wall id) -> Condition
    solver: dict                           # equation -> Solver
    rules: list = field(default_factory=list)  # (subdomain, [wall ids], target wall id), file order


def read_bic(path: str, mesh: Mesh) -> BicProblem:
    """The BIC file of a mesh whose wall names are mesh.walls."""
    if not mesh.walls:
        raise ValueError("the mesh has no wall names (a gmsh file with $PhysicalNames is needed)")
    walls = list(mesh.walls)
    wid = {n: k + 1 for k, n in enumerate(walls)}
    lines = read_lines(path)
    equations, sd_names, sds = [], [], []
    i = 0
    while i < len(lines):
        w = _words(lines[i])
        key = w[0].upper()
        if key == "EQUATIONS":
            n = int(w[1])
            equations = [_words(lines[i + 1 + k])[0] for k in range(n)]
            i += n
        elif key == "SUBDOMAINS":
            n = int(w[1])
            for k in range(n):
                ln = lines[i + 1 + k]
                t = _words(ln)
                # a wall belongs to the subdomain when its name occurs anywhere in the line, walls
                # taken in wall-index order (ReadsdBIC, src/lib_mesh.f90:1027-1043)
                members = [j + 1 for j, nm in enumerate(walls) if nm in ln]
                sd_names.append(t[0])
                sds.append(api.Subdomain(members, [1.0] * len(members), float(t[1])))
            i += n
        i += 1
    if not equations:
        raise ValueError("BIC: EQUATIONS missing")
    if not sds:
        raise ValueError("BIC: SUBDOMAINS missing")
    boundary, initial = {}, {}
    solver = {e: Solver() for e in equations}
    for e in equations:  # default: zero flux on every wall, initial function 0 (ReadBIC)
        for k in range(1, len(walls) + 1):
            boundary[(e, k)] = Condition(api.BC_FLUX, "const", np.array([0.0]))
            initial[(e, k)] = Condition(api.BC_FUNCTION, "const", np.array([0.0]))
    rules = []
    i = 0
    while i < len(lines):
        w = _words(lines[i])
        key = w[0].upper()
        if key == "NORMMUL":
            sd, wall, mult = w[1], w[2], float(w[3])
            for s, nm in enumerate(sd_names):
                if nm == sd:
                    sub = sds[s]
                    for k, wl in enumerate(sub.walls):
                        if walls[wl - 1] == wall:
                            sub.normmul[k] = mult
        elif key == "SOLVER":
            e = w[1]
            if e not in solver:
                raise ValueError(f"BIC: SOLVER of unknown equation {e}")
            solver[e] = Solver(int(w[2]), int(w[3]), int(w[4]), int(w[5]), int(w[6]), float(w[7]))
        elif key in ("BOUNDARY", "INITIAL"):
            e, wall, kind, form = w[1], w[2], w[3], w[4]
            if e not in solver or wall not in wid or kind not in KINDS or form not in NPARAMS:
                raise ValueError(f"BIC: bad condition line: {lines[i]}")
            vals = [float(x) for x in _words(lines[i + 1])[:NPARAMS[form]]]
            (boundary if key == "BOUNDARY" else initial)[(e, wid[wall])] = Condition(KINDS[kind], form,
                                                                                    np.array(vals))
            i += 1
        elif key in ("EDGE", "CORNER"):
            nw = 2 if key == "EDGE" else 3
            sd = sd_names.index(w[1])
            rules.append((sd, [wid[x] for x in w[2:2 + nw]], wid[w[2 + nw]]))
        i += 1
    prob = BicProblem(equations, sd_names, sds, boundary, initial, solver, rules)
    _apply_rules(mesh, prob)
    return prob


def _apply_rules(mesh: Mesh, prob: BicProblem):
    """EDGE and CORNER rules in file order: the subdomain's nodes shared by elements of every listed
    wall (of that subdomain) take the target wall's condition (HandleCornersEdges)."""
    con = np.asarray(mesh.con)
    bcid = np.asarray(mesh.bcid)
    for sd, wl, target in prob.rules:
        sub = prob.subdomains[sd]
        if sub.node_wall is None:
            sub.node_wall = np.zeros(mesh.nnodes, dtype=np.int32)
        members = set(sub.walls)
        sets = []
        for w in wl:
            if w not in members:
                sets.append(set())
                continue
            c = con[bcid == w]
            sets.append(set(int(v) for v in c.reshape(-1) if v >= 0))
        for nd in set.intersection(*sets):
            sub.node_wall[nd] = target


def bic_value(c: Condition, x: np.ndarray) -> np.ndarray:
    """GetBiCValue at points x (n, 3)."""
    p = c.params
    if c.form == "const":
        return np.full(x.shape[0], p[0])
    v = p[0] + p[1] * x[:, 0] + p[2] * x[:, 1] + p[3] * x[:, 2]
    if c.form == "quad":
        v = v + (p[4] * x[:, 0] * x[:, 1] + p[5] * x[:, 0] * x[:, 2] + p[6] * x[:, 1] * x[:, 2] +
                 p[7] * x[:, 0] ** 2 + p[8] * x[:, 1] ** 2 + p[9] * x[:, 2] ** 2)
    return v


def qnode_points(mesh: Mesh) -> np.ndarray:
    """Flux points: element centroids (the mean of the corner nodes, CalQMesh)."""
    con = np.asarray(mesh.con)
    pts = np.zeros((mesh.nelem, 3))
    for e in range(mesh.nelem):
        vs = [v for v in con[e] if v >= 0]
        pts[e] = mesh.xyz[vs].mean(axis=0)
    return pts


def conditions(mesh: Mesh, prob: BicProblem, eq: str):
    """(wall_bc, u, q) of one equation: the kind of every wall and the known and initial values
    (sdApplyInitalCond: per subdomain, nodes then flux points)."""
    nw = len(mesh.walls)
    wall_bc = np.array([prob.boundary[(eq, w)].kind for w in range(1, nw + 1)], dtype=np.int32)
    node_wall = api.subdomain_node_walls(mesh, prob.subdomains)
    u = np.zeros(mesh.nnodes)
    q = np.zeros(mesh.nelem)
    cq = qnode_points(mesh)
    bcid = np.asarray(mesh.bcid)
    for s, sub in enumerate(prob.subdomains):
        nodes = np.nonzero(node_wall[s])[0]
        for i in nodes:
            w = int(node_wall[s][i])
            for c, kinds in ((prob.initial[(eq, w)], (api.BC_FUNCTION,)),
                             (prob.boundary[(eq, w)], (api.BC_FUNCTION, api.BC_CONTACT))):
                if c.kind in kinds:
                    u[i] = bic_value(c, mesh.xyz[i:i + 1])[0]
        elems = np.concatenate([np.nonzero(bcid == w)[0] for w in sub.walls])
        for e in elems:
            w = int(bcid[e])
            for c, kinds in ((prob.initial[(eq, w)], (api.BC_FLUX,)),
                             (prob.boundary[(eq, w)], (api.BC_FLUX, api.BC_CONTACT))):
                if c.kind in kinds:
                    q[e] = bic_value(c, cq[e:e + 1])[0]
    return wall_bc, u, q


def read_inp(path: str) -> dict:
    """and.inp keys of the Laplace driver: MDIR, LGEO (mesh), LBIC, ITRI, IQUD (src/mod_par.f90)."""
    out = {"MDIR": "./", "ITRI": 5, "IQUD": 4}
    for ln in read_lines(path):
        w = ln.split()
        key = w[0].upper()
        if key in ("MDIR", "LGEO", "LBIC") and len(w) > 1:
            out[key] = w[1]
        elif key in ("ITRI", "IQUD") and len(w) > 1:
            out[key] = int(w[1])
    return out


def load_case(inp_path: str):
    """(mesh, problem, inp) of a Laplace case directory's and.inp (paths relative to the case)."""
    from .mesh import read_gmsh
    inp = read_inp(inp_path)
    base = os.path.join(os.path.dirname(os.path.abspath(inp_path)), inp["MDIR"])
    mesh = read_gmsh(os.path.normpath(os.path.join(base, inp["LGEO"])))
    prob = read_bic(os.path.normpath(os.path.join(base, inp["LBIC"])), mesh)
    return mesh, prob, inp


def solve_laplace_case(ctx: "api.Context", mesh: Mesh, prob: BicProblem, itri: int = 5, iquad: int = 4,
                       eps=None, maxit=None, precision: int = 0):
    """Every equation of the problem on the device (andr_laplace_solve; precision 1 = api.PREC_MIXED:
    andr_laplace_solve_mixed); eps and maxit override the SOLVER settings. Returns
    {equation: api.LaplaceResult}."""
    bcs, us, qs = [], [], []
    for e in prob.equations:
        bc, u, q = conditions(mesh, prob, e)
        bcs.append(bc)
        us.append(u)
        qs.append(q)
    ep = [prob.solver[e].eps if eps is None else eps for e in prob.equations]
    mi = [prob.solver[e].maxit if maxit is None else maxit for e in prob.equations]
    res = ctx.laplace_solve(mesh, prob.subdomains, np.array(bcs), np.array(us), np.array(qs), eps=ep, maxit=mi,
                            itri=itri, iquad=iquad, precision=precision)
    return dict(zip(prob.equations, res))


def solve_stokes_case(ctx: "api.Context", mesh: Mesh, prob: BicProblem, itri: int = 5, eps=None, maxit=None):
    """The Stokes driver's general problem on the device (andr_stokes_solve): the three equations of
    the BIC file are the velocity components (EQUATIONS 3: X, Y, Z), one subdomain, the solver limits
    of the first equation (src/lib_Stokes.f90:626-629)."""
    if len(prob.equations) != 3 or len(prob.subdomains) != 1:
        raise ValueError("a Stokes case has three equations (the velocity components) and one subdomain")
    bcs, us, qs = zip(*[conditions(mesh, prob, e) for e in prob.equations])
    slv = prob.solver[prob.equations[0]]
    return ctx.stokes_solve(mesh, prob.subdomains[0], np.array(bcs), np.array(us), np.array(qs),
                            eps=slv.eps if eps is None else eps, maxit=slv.maxit if maxit is None else maxit,
                            itri=itri)
