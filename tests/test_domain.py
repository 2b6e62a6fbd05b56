"""Velocity and pressure at points of the flow domain (GetStokesDomainValue,
src/lib_Stokes.f90:3043-3390) and on the surface (pressureStokes, :4-94). CPU: the oracle pinned by exact Stokes flows inside a sphere — a
uniform flow (reproduced up to the reference's single-precision 8 pi) and u = (x^2, -2xy, 0),
p = 2 mu x (second-order convergence of velocity and pressure under refinement), with the
closure error telling points inside from points outside; the surface pressure is zero for a
uniform flow and converges at first order to 2 mu x for the quadratic one. GPU: the HIP kernels
against the oracle on the same inputs."""
import numpy as np
import pytest

from andromeda_amd import mesh as M

INSIDE = np.array([[0.0, 0.0, 0.0], [0.3, -0.2, 0.1], [0.0, 0.5, 0.0], [-0.4, 0.2, 0.45]])
OUTSIDE = np.array([[2.0, 0.0, 0.0], [0.0, -1.5, 1.0]])


def _sigma(x):  # stress of u = (x^2, -2xy, 0), p = 2x, mu = 1
    return np.array([[2 * x[0], -2 * x[1], 0.0], [-2 * x[1], -6 * x[0], 0.0], [0.0, 0.0, -2 * x[0]]])


def _quadratic_flow(orc, m):
    nrm, _, cen = orc.geometry(m)
    X = m.xyz
    u = np.stack([X[:, 0] ** 2, -2 * X[:, 0] * X[:, 1], 0 * X[:, 0]], 1)
    # the traction of the boundary data is -sigma n for the outward normal of the mesh
    q = np.array([-_sigma(c) @ n for c, n in zip(cen, nrm)])
    ex = np.array([[x * x, -2 * x * y, 0.0] for x, y, _ in INSIDE])
    return u, q, ex, 2 * INSIDE[:, 0]


def test_oracle_domain_uniform_flow(orc):
    m = M.icosphere(3)
    U = np.array([1.0, -2.0, 0.5])
    pts = np.vstack([INSIDE, OUTSIDE])
    vel, p, err = orc.stokes_domain_values(m, pts, np.tile(U, (m.nnodes, 1)), np.zeros((m.nelem, 3)))
    # inside: the closure holds up to the reference's REAL(4) 8 pi (relative 2.8e-8)
    assert np.abs(vel[:4] - U).max() < 1e-7 and np.abs(p[:4]).max() < 1e-12
    assert err[:4].max() < 1e-7
    # outside: the stresslet integral vanishes, the velocity with it
    assert np.abs(vel[4:]).max() < 1e-10 and np.all(np.abs(err[4:] - 1.0) < 1e-10)


def test_oracle_domain_quadratic_flow_converges(orc):
    ev, ep = [], []
    for s in (2, 3, 4):
        m = M.icosphere(s)
        u, q, ex, pe = _quadratic_flow(orc, m)
        vel, p, err = orc.stokes_domain_values(m, INSIDE, u, q)
        ev.append(np.abs(vel - ex).max())
        ep.append(np.abs(p - pe).max())
        assert err.max() < 1e-7
    # second order in the element size: a factor ~4 per subdivision
    assert ev[0] / ev[1] > 3.5 and ev[1] / ev[2] > 3.5 and ev[2] < 1e-3, ev
    assert ep[0] / ep[1] > 3.5 and ep[1] / ep[2] > 3.5 and ep[2] < 1e-3, ep


def test_oracle_domain_viscosity_scaling(orc):
    """With viscosity diff the tractions scale with it and the velocity does not change; the
    pressure scales with it."""
    m = M.icosphere(2)
    u, q, _, _ = _quadratic_flow(orc, m)
    v1, p1, _ = orc.stokes_domain_values(m, INSIDE, u, q, diff=1.0)
    v3, p3, _ = orc.stokes_domain_values(m, INSIDE, u, 3.0 * q, diff=3.0)
    assert np.abs(v3 - v1).max() < 1e-13 and np.abs(p3 - 3.0 * p1).max() < 1e-12


def _mesh(kind):
    if kind == "sphere":
        return M.icosphere(3), None
    if kind == "superellipsoid":
        return M.random_superellipsoid(np.random.default_rng(9), 6)[0], None
    # a particle in a far sphere, two walls, the outer one's normals turned (NORMMUL -1)
    return M.body_in_sphere(M.geodesic_sphere(3, 0.5), 4.0, 2), np.array([1.0, -1.0])


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["sphere", "superellipsoid", "two_walls"])
def test_gpu_domain_values_vs_oracle(ctx, orc, kind):
    rng = np.random.default_rng(4)
    m, nm = _mesh(kind)
    u = rng.normal(size=(m.nnodes, 3))
    q = rng.normal(size=(m.nelem, 3))
    pts = np.vstack([INSIDE * 0.5, OUTSIDE * 2.0, rng.uniform(-0.3, 0.3, size=(57, 3))])
    if kind == "two_walls":  # between the particle and the outer sphere
        pts = np.vstack([pts[:6] * 1.5 + np.array([0.0, 0.0, 1.2]), rng.uniform(0.8, 2.5, size=(20, 3))])
    vel, p, err = ctx.stokes_domain_values(m, pts, u, q, diff=1.7, normmul=nm)
    vo, po, eo = orc.stokes_domain_values(m, pts, u, q, diff=1.7, normmul=nm)
    scale = np.abs(vo).max()
    assert np.abs(vel - vo).max() < 1e-10 * scale
    assert np.abs(p - po).max() < 1e-10 * np.abs(po).max()
    assert np.abs(err - eo).max() < 1e-10


@pytest.mark.gpu
def test_gpu_domain_quadratic_flow(ctx, orc):
    m = M.icosphere(4)
    u, q, ex, pe = _quadratic_flow(orc, m)
    vel, p, err = ctx.stokes_domain_values(m, INSIDE, u, q)
    assert np.abs(vel - ex).max() < 1e-3 and np.abs(p - pe).max() < 1e-3 and err.max() < 1e-7


def test_oracle_surface_pressure_pins(orc):
    U = np.array([1.0, -2.0, 0.5])
    rms = []
    for sub in (2, 3):
        m = M.icosphere(sub)
        assert np.abs(orc.stokes_surface_pressure(m, np.tile(U, (m.nnodes, 1)), np.zeros((m.nelem, 3)))).max() < 1e-12
        u, q, _, _ = _quadratic_flow(orc, m)
        _, _, cen = orc.geometry(m)
        p = orc.stokes_surface_pressure(m, u, q)
        rms.append(np.sqrt(np.mean((p - 2 * cen[:, 0]) ** 2)))
    # first order in the element size (the flat-element factor 2 of the reference)
    assert rms[0] / rms[1] > 1.8 and rms[1] < 0.035, rms


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["sphere", "superellipsoid", "two_walls"])
def test_gpu_surface_pressure_vs_oracle(ctx, orc, kind):
    rng = np.random.default_rng(6)
    m, nm = _mesh(kind)
    u = rng.normal(size=(m.nnodes, 3))
    q = rng.normal(size=(m.nelem, 3))
    p = ctx.stokes_surface_pressure(m, u, q, diff=0.8, normmul=nm)
    po = orc.stokes_surface_pressure(m, u, q, diff=0.8, normmul=nm)
    assert np.abs(p - po).max() < 1e-10 * np.abs(po).max()
