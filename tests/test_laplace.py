"""The Laplace boundary-value problem: the reference's input files (BIC parsing, conditions), the oracle
pinned by exact solutions and the reference's own analytic values, and the device path
(andr_laplace_solve, andr_assemble_*_rows) against the oracle.

Pins of the oracle (CPU):
* the unit-sphere Dirichlet problem u = x (icosphere subdivision 3, 1280 triangles): the flux of the
  linear field is constant on every flat element, so the discrete solution is exact up to
  quadrature and solver tolerance: q = -n_x, sum_e A_e q_e = 0 and sum_e A_e q_e x_e = -V of the
  polyhedron (the divergence theorem, exact for a linear integrand on flat triangles);
* the reference's cube case (run/laplace/cube-triangles): u = x - 1/2 between the walls at x = 0 and
  x = 1 with zero flux elsewhere, wall fluxes (0, 0, 0, 0, 1, -1), and the second equation's
  fluxes +-2;
* the reference's two joined cylinders (run/laplace/cylinder-2joined), diffusivities 2 and 5,
  u = 3 and 1 at the ends: the contact value 11/7 that the case file states as the analytic
  solution, and flux conservation through the contact wall;
* the reference's three nested spheres (run/laplace/sphere-onion): the fluxes through the two
  Dirichlet patches cancel.
The fixtures under tests/golden/cases are the reference's mesh, BIC and and.inp files of those
cases (the mesh paths of and.inp point into the fixture folder).
"""
import os

import numpy as np
import pytest

from andromeda_amd import api, bic
from andromeda_amd import mesh as M

CASES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases")


def _case(name):
    return bic.load_case(os.path.join(CASES, name, "and.inp"))


def _oracle_solve(mesh, prob, eq, eps=1e-14, maxit=5000):
    import laplace_oracle as LO
    sds = [LO.Subdomain(s.walls, s.normmul, s.diff, s.node_wall) for s in prob.subdomains]
    L = LO.subdomains(mesh, sds)
    mats = LO.assemble(mesh, sds, L)
    bc, u, q = bic.conditions(mesh, prob, eq)
    u2, q2, info = LO.solve(mesh, sds, L, mats, bc, u, q, eps=eps, maxit=maxit)
    return u2, q2, LO.wall_flux(mesh, q2, len(mesh.walls)), info


def _sphere_problem():
    m = M.icosphere(3)
    m = M.Mesh(m.xyz, m.con, m.bcid, m.name, ["krogla"])
    prob = bic.BicProblem(["u"], ["sd"], [api.Subdomain([1], [1.0], 1.0)],
                          {("u", 1): bic.Condition(api.BC_FUNCTION, "lin", np.array([0.0, 1.0, 0.0, 0.0]))},
                          {("u", 1): bic.Condition(api.BC_FUNCTION, "const", np.array([0.0]))},
                          {"u": bic.Solver(maxit=0, eps=1e-12)})
    return m, prob


def _poly_volume(m):
    x1, x2, x3 = (m.xyz[m.con[:, k]] for k in range(3))
    return np.einsum("ij,ij->", np.cross(x2 - x1, x3 - x1), x1) / 6.0


# ---- input files (CPU) ---------------------------------------------------------------------------
def test_bic_reader_cylinder_case():
    mesh, prob, inp = _case("cylinder-2joined")
    assert mesh.walls == ["in1", "in2", "wall1", "wall2", "contact"]
    assert inp["ITRI"] == 5 and inp["IQUD"] == 4
    # walls of a SUBDOMAINS line are matched by name anywhere in the line and taken in wall order
    assert [s.walls for s in prob.subdomains] == [[1, 3, 5], [2, 4, 5]]
    assert [s.diff for s in prob.subdomains] == [2.0, 5.0]
    assert prob.subdomains[0].normmul == [1.0, 1.0, -1.0] and prob.subdomains[1].normmul == [1.0, 1.0, 1.0]
    assert prob.solver["temperatura"].maxit == 500 and prob.solver["temperatura"].eps == 1e-15
    bc, u, q = bic.conditions(mesh, prob, "temperatura")
    assert bc.tolist() == [1, 1, 2, 2, 3]
    # sdApplyInitalCond: the initial value, then the boundary value of the same kind over it
    assert np.all(q[mesh.bcid == 4] == 0.0)  # INITIAL flux 1.0, then BOUNDARY flux 0.0 on wall2
    nw = api.subdomain_node_walls(mesh, prob.subdomains)
    assert np.all(u[nw[0] == 3] == -3.3)     # INITIAL function -3.3: the hot start of wall1's unknown u
    assert np.all(u[nw[0] == 1] == 3.0) and np.all(u[nw[1] == 2] == 1.0)


def test_bic_edge_rules_cube():
    """EDGE and CORNER rules: every node shared by elements of the listed walls takes the target
    wall (checked against a brute-force pass over the elements), later rules over earlier ones."""
    mesh, prob, _ = _case("cube-triangles")
    nw = api.subdomain_node_walls(mesh, prob.subdomains)[0]
    expect = np.zeros(mesh.nnodes, dtype=np.int32)
    for e in range(mesh.nelem):  # default: the wall of the last element touching the node
        expect[mesh.con[e]] = mesh.bcid[e]
    for _, walls, target in prob.rules:
        for i in range(mesh.nnodes):
            if all(np.any(mesh.con[mesh.bcid == w] == i) for w in walls):
                expect[i] = target
    assert np.array_equal(nw, expect)
    levo, desno = mesh.walls.index("levo") + 1, mesh.walls.index("desno") + 1
    assert np.all(np.isin(nw[np.isclose(mesh.xyz[:, 0], 0.0)], [levo]))
    assert np.all(np.isin(nw[np.isclose(mesh.xyz[:, 0], 1.0)], [desno]))


def test_bic_value_forms():
    x = np.array([[1.0, 2.0, 3.0]])
    assert bic.bic_value(bic.Condition(1, "const", np.array([4.0])), x)[0] == 4.0
    assert bic.bic_value(bic.Condition(1, "lin", np.array([1.0, 2.0, 3.0, 4.0])), x)[0] == 1 + 2 + 6 + 12
    p = np.arange(1.0, 11.0)
    assert bic.bic_value(bic.Condition(1, "quad", p), x)[0] == 1 + 2 + 6 + 12 + 5 * 2 + 6 * 3 + 7 * 6 + 8 + 36 + 90


def test_subdomain_node_walls_rejects_bad_input():
    m = M.icosphere(1)
    with pytest.raises(api.AndromedaError, match="wall id 2"):
        api.subdomain_node_walls(m, [api.Subdomain([2])])
    with pytest.raises(api.AndromedaError, match="diff"):
        api.subdomain_node_walls(m, [api.Subdomain([1], diff=0.0)])


# ---- oracle pins (CPU) ---------------------------------------------------------------------------
def test_oracle_sphere_dirichlet_exact():
    m, prob = _sphere_problem()
    u, q, F, info = _oracle_solve(m, prob, "u", eps=1e-12, maxit=0)
    import pyoracle
    nrm, area, cen = pyoracle.geometry(m)
    assert info["neq"] == info["nx"] == 1280 and info["nb"] == 642
    assert np.abs(q + nrm[:, 0]).max() < 1e-7
    assert abs(F[0]) < 1e-10
    assert abs(np.sum(area * q * cen[:, 0]) + _poly_volume(m)) < 1e-7
    assert np.array_equal(u, m.xyz[:, 0])


def test_oracle_cube_case_linear_solution():
    mesh, prob, _ = _case("cube-triangles")
    u, q, F, info = _oracle_solve(mesh, prob, "temperatura")
    assert np.abs(u - (mesh.xyz[:, 0] - 0.5)).max() < 1e-7
    assert np.abs(F - [0, 0, 0, 0, 1, -1]).max() < 1e-6
    u, q, F, info = _oracle_solve(mesh, prob, "koncentracija")
    assert np.abs(F - [2, -2, 0, 0, 0, 0]).max() < 1e-5


def test_oracle_joined_cylinders_contact_value():
    mesh, prob, _ = _case("cylinder-2joined")
    u, q, F, info = _oracle_solve(mesh, prob, "temperatura")
    contact = np.unique(mesh.con[mesh.bcid == 5])
    assert np.abs(u[contact] - 11.0 / 7.0).max() < 1e-4  # the case file's analytic value
    assert abs(F[0] + F[1]) < 1e-4 * abs(F[0]) and abs(F[0] - F[4]) < 1e-4 * abs(F[0])


def test_oracle_nested_spheres_conservation():
    mesh, prob, _ = _case("sphere-onion")
    u, q, F, info = _oracle_solve(mesh, prob, "temperatura")
    assert abs(F[0] + F[1]) < 1e-3 * abs(F[0]) and abs(F[2]) == 0.0


# ---- the device path -----------------------------------------------------------------------------
def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["sphere", "cube-triangles", "cylinder-2joined", "sphere-onion"])
def test_gpu_laplace_solve_vs_oracle(ctx, case):
    if case == "sphere":
        mesh, prob = _sphere_problem()
    else:
        mesh, prob, _ = _case(case)
    res = bic.solve_laplace_case(ctx, mesh, prob, eps=1e-14, maxit=5000)
    for eq, r in res.items():
        u, q, F, info = _oracle_solve(mesh, prob, eq)
        print(f"{case} {eq}: neq {r.neq} nx {r.nx} nb {r.nb} its {r.nit} (oracle {info['nit']}) istop {r.istop} "
              f"u {_rel(r.u, u):.1e} q {_rel(r.q, q):.1e} flux {_rel(r.wall_flux, F):.1e}")
        assert (r.neq, r.nx, r.nb) == (info["neq"], info["nx"], info["nb"])
        assert r.istop in (1, 2) and abs(r.nit - info["nit"]) <= max(3, info["nit"] // 20)
        assert _rel(r.u, u) < 1e-9 and _rel(r.q, q) < 1e-9
        assert np.abs(r.wall_flux - F).max() < 1e-9 * max(np.abs(F).max(), 1.0)


@pytest.mark.gpu
def test_gpu_laplace_sphere_flux_results(ctx):
    """C1 end to end on the device: u = x on the unit icosphere (s = 3), LSQR eps 1e-12; the fluxes
    reported per wall and the flux moment against their exact values."""
    m, prob = _sphere_problem()
    r = bic.solve_laplace_case(ctx, m, prob)["u"]
    import pyoracle
    nrm, area, cen = pyoracle.geometry(m)
    assert r.istop in (1, 2)
    assert np.abs(r.q + nrm[:, 0]).max() < 1e-7
    assert abs(r.wall_flux[0]) < 1e-10
    assert abs(np.sum(area * r.q * cen[:, 0]) + _poly_volume(m)) < 1e-7
    assert r.err[0] < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["sphere", "cube-triangles", "cylinder-2joined", "sphere-onion"])
def test_gpu_laplace_mixed_precision(ctx, case):
    """The mixed-precision form (FP32 regular integrals and system storage, FP64 singular integrals,
    products, LSQR vectors and update) against the plain FP64 path on the same cases, the error
    stated per equation: u within 1e-6, q within 1e-5 (2.2e-6 on the joined cylinders, whose contact
    system needs ~260 iterations; 7e-7 on C1) and the wall fluxes within 1e-6 of the largest
    (relative); on C1 (the sphere, u = x) the flux and its moment against the exact values as well."""
    if case == "sphere":
        mesh, prob = _sphere_problem()
    else:
        mesh, prob, _ = _case(case)
    plain = bic.solve_laplace_case(ctx, mesh, prob, eps=1e-14, maxit=5000)
    mixed = bic.solve_laplace_case(ctx, mesh, prob, eps=1e-14, maxit=5000, precision=api.PREC_MIXED)
    for eq, r in mixed.items():
        p = plain[eq]
        eu, eq_, ef = _rel(r.u, p.u), _rel(r.q, p.q), np.abs(r.wall_flux - p.wall_flux).max() / max(
            np.abs(p.wall_flux).max(), 1.0)
        print(f"{case} {eq}: mixed vs FP64: u {eu:.1e} q {eq_:.1e} flux {ef:.1e}; its {r.nit} (FP64 {p.nit}) "
              f"istop {r.istop}")
        assert r.istop in (1, 2) and eu < 1e-6 and eq_ < 1e-5 and ef < 1e-6
    if case == "sphere":
        import pyoracle
        r = mixed["u"]
        nrm, area, cen = pyoracle.geometry(mesh)
        assert np.abs(r.q + nrm[:, 0]).max() < 1e-5
        assert abs(np.sum(area * r.q * cen[:, 0]) + _poly_volume(mesh)) < 1e-6


@pytest.mark.gpu
def test_gpu_assembly_row_ranges_bitwise(ctx):
    """andr_assemble_laplace_rows / andr_assemble_stokes_rows: row ranges (inside the node rows,
    across the node/element boundary, at the end) equal those rows of the full assembly bit for bit;
    the self-check errors cover the range's rows."""
    m = M.icosphere(2)
    H, G, err = ctx.assemble_laplace(m)
    nn, nsp = m.nnodes, m.nnodes + m.nelem
    for r0, r1 in ((0, 17), (nn - 5, nn + 9), (nsp - 3, nsp)):
        Hr, Gr, er = ctx.assemble_laplace_rows(m, r0, r1)
        assert np.array_equal(Hr, H[r0:r1]) and np.array_equal(Gr, G[r0:r1])
        Hc, Gc, _ = ctx.assemble_laplace_rows(m, r0, r1, layout=api.COL_MAJOR)
        assert np.array_equal(Hc, Hr) and np.array_equal(Gc, Gr)
        assert er[1] <= err[1]
    ms = M.geodesic_sphere(2)
    T, Gs, pT, pG, _ = ctx.assemble_stokes(ms, pressure=True)
    nn = ms.nnodes
    Tr, Gsr, pTr, pGr, _ = ctx.assemble_stokes_rows(ms, nn - 4, nn + 6)
    for k in range(6):
        assert np.array_equal(Tr[k], T[k][nn - 4:nn + 6]) and np.array_equal(Gsr[k], Gs[k][nn - 4:nn + 6])
    for k in range(3):
        assert np.array_equal(pTr[k], pT[k][nn - 4:nn + 6]) and np.array_equal(pGr[k], pG[k][nn - 4:nn + 6])
    with pytest.raises(api.AndromedaError, match=r"\(-1\)"):
        ctx.assemble_laplace_rows(m, 5, 5)
