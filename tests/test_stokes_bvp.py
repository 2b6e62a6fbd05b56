"""The general Stokes problem of the reference's Stokes driver (velocity or traction known per wall and
component, one subdomain) on the reference's Couette case (run/stokes/Couette: a unit cube of 624
triangles, kocka-000-111-314-624_tri.msh, copied with its BIC files into tests/golden/cases).

Pins of the oracle (oracle/stokes_bvp_oracle.py), from the exact flow u = (z, 0, 0), stress
sigma_xz = sigma_zx = mu = 1 (traction q = -sigma n in the reference's convention):
* stokesCouette-X.bic (velocity known everywhere): the system is singular along the constant-pressure
  mode (q + c n), LSQR stops at its condition limit; q = -sigma n up to that mode (1.5e-3), the shear
  forces on the top and bottom walls -1 and +1;
* stokesCouette-X-fluxBC.bic (traction +-3 on the side walls): the pressure is fixed at 3; the
  velocity on the traction walls is x-independent Couette (u_x = z to 1e-4) and every wall force is
  the exact one (shear 1, pressure 3).
The device path (andr_stokes_solve) against the oracle at 1e-9 on both.
"""
import os

import numpy as np
import pytest

from andromeda_amd import api, bic

CASE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases", "couette")


def _case(bicfile):
    mesh, _, _ = bic.load_case(os.path.join(CASE, "and.inp"))
    return mesh, bic.read_bic(os.path.join(CASE, bicfile), mesh)


def _oracle(mesh, prob, eps=1e-14, maxit=5000):
    import laplace_oracle as LO
    import stokes_bvp_oracle as SO
    sd = prob.subdomains[0]
    s = LO.Subdomain(sd.walls, sd.normmul, sd.diff, sd.node_wall)
    L = LO.subdomains(mesh, [s])[0]
    T, G, _ = SO.assemble(mesh, s, L)
    bcs, us, qs = zip(*[bic.conditions(mesh, prob, e) for e in prob.equations])
    u, q, info = SO.solve(mesh, s, L, T, G, np.array(bcs), np.array(us), np.array(qs), eps=eps, maxit=maxit)
    return u, q, SO.wall_force(mesh, q, len(mesh.walls)), info


def _exact(mesh):
    import pyoracle
    nrm, area, _ = pyoracle.geometry(mesh)
    sig = np.zeros((3, 3))
    sig[0, 2] = sig[2, 0] = 1.0
    return -(nrm @ sig.T).T, nrm, area  # (3, nelem)


def test_oracle_couette_velocity_known():
    mesh, prob = _case("stokesCouette-X.bic")
    u, q, F, info = _oracle(mesh, prob)
    qe, nrm, _ = _exact(mesh)
    c = np.sum((q - qe) * nrm.T) / np.sum(nrm.T * nrm.T)  # the constant-pressure mode
    assert info["istop"] == 4
    assert np.abs(q - qe - c * nrm.T).max() < 2e-3
    top, bottom = mesh.walls.index("top"), mesh.walls.index("bottom")
    assert abs(F[0, top] + 1) < 1e-4 and abs(F[0, bottom] - 1) < 1e-4


def test_oracle_couette_traction_walls():
    mesh, prob = _case("stokesCouette-X-fluxBC.bic")
    u, q, F, info = _oracle(mesh, prob)
    assert info["istop"] in (1, 2)
    assert np.abs(u[0] - mesh.xyz[:, 2]).max() < 1e-4 and np.abs(u[1:]).max() == 0.0
    W = {n: k for k, n in enumerate(mesh.walls)}
    exact = np.zeros((3, len(mesh.walls)))
    exact[0, W["top"]], exact[0, W["bottom"]] = -1, 1
    exact[0, W["left"]], exact[0, W["right"]] = -3, 3
    exact[1, W["back"]], exact[1, W["front"]] = 3, -3
    exact[2, W["top"]], exact[2, W["bottom"]] = 3, -3
    exact[2, W["left"]], exact[2, W["right"]] = 1, -1
    assert np.abs(F - exact).max() < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("bicfile", ["stokesCouette-X.bic", "stokesCouette-X-fluxBC.bic"])
def test_gpu_stokes_solve_vs_oracle(ctx, bicfile):
    mesh, prob = _case(bicfile)
    r = bic.solve_stokes_case(ctx, mesh, prob, eps=1e-14, maxit=5000)
    u, q, F, info = _oracle(mesh, prob)
    rel = lambda a, b: np.abs(a - b).max() / np.abs(b).max()
    print(f"{bicfile}: neq {r.neq} nx {r.nx} nb {r.nb} its {r.nit} (oracle {info['nit']}) istop {r.istop} "
          f"(oracle {info['istop']}) u {rel(r.u, u):.1e} q {rel(r.q, q):.1e} F {rel(r.wall_flux, F):.1e}")
    assert (r.neq, r.nx, r.nb) == (info["neq"], info["nx"], info["nb"])
    assert r.istop == info["istop"]
    if info["istop"] == 4:
        # singular system (velocity known everywhere): the component along the constant-pressure mode
        # q = c n depends on where LSQR stops; the rest is the solution and must agree
        _, nrm, _ = _exact(mesh)
        n = nrm.T.reshape(-1)

        def proj(x):
            x = x.reshape(-1)
            return x - (x @ n) / (n @ n) * n
        print(f"  along the null mode removed: q {rel(proj(r.q), proj(q)):.1e}")
        assert rel(r.u, u) < 1e-9 and rel(proj(r.q), proj(q)) < 1e-7
    else:
        assert rel(r.u, u) < 1e-9 and rel(r.q, q) < 1e-9 and rel(r.wall_flux, F) < 1e-9
