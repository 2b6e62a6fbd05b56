"""Oracle assembly, LSQR and resistance checked against what the mathematics fixes:
discrete identities of the collocation rows, the exact Laplace solution u = x, the sphere's
Stokes resistance 6 pi mu a / 8 pi mu a^3, and LSQR against dense least squares."""
import numpy as np
import pytest

from andromeda_amd import mesh as M


def test_laplace_rows_identities(orc):
    m = M.icosphere(2)
    H, G, err = orc.assemble_laplace(m)
    nn, ne = m.nnodes, m.nelem
    nrm, area, cen = orc.geometry(m)
    # rigid-body trick: every row of H sums to zero after the free term
    assert np.abs(H.sum(1)).max() < 1e-13
    # node rows reproduce u = x, q = n_x exactly (src/lib_sdBem.f90:566-581)
    assert err[0] < 1e-13
    r = H[:nn] @ m.xyz - G[:nn] @ nrm
    assert np.abs(r).max() < 1e-13
    # centroid rows: free term close to 1/2 on a smooth surface
    assert err[1] < 0.05


def test_laplace_sphere_solution(orc):
    m = M.icosphere(2)
    H, G, _ = orc.assemble_laplace(m)
    nrm, area, cen = orc.geometry(m)
    sysM, rhsM = orc.laplace_dirichlet_system(m, H, G)
    b = rhsM @ (-m.xyz[:, 0])
    ia, ja, a = orc.csr_from_dense(sysM)
    q = orc.lsqr_precon(ia, ja, a, m.nelem, m.nelem)
    x, info = orc.lsqr_solve(ia, ja, a, m.nelem, m.nelem, q, b, eps=1e-12)
    assert info["ierr"] == 0 and info["istop"] in (1, 2)
    # the reference's flux is -diff du/dn (b = rhs * (-u), src/lib_CRS.f90:66): q = -n_x
    assert np.linalg.norm(x + nrm[:, 0]) / np.linalg.norm(nrm[:, 0]) < 1e-6
    assert abs((area * x).sum()) < 1e-10


def test_lsqr_against_lstsq(orc):
    rng = np.random.default_rng(3)
    for m_, n_ in [(60, 40), (50, 50)]:
        A = rng.normal(size=(m_, n_)) + 3 * np.eye(m_, n_)
        b = rng.normal(size=m_)
        ia, ja, a = orc.csr_from_dense(A)
        q = orc.lsqr_precon(ia, ja, a, m_, n_)
        assert np.allclose(q, 1 / np.linalg.norm(A, axis=0), rtol=1e-14)
        x, info = orc.lsqr_solve(ia, ja, a, m_, n_, q, b, eps=1e-14)
        xr = np.linalg.lstsq(A, b, rcond=None)[0]
        assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-10
        assert info["maxit"] == 10 * (m_ + n_ + 50)  # maxit = 0 is written back


def test_lsqr_iteration_limit(orc):
    rng = np.random.default_rng(4)
    A = rng.normal(size=(40, 40))
    ia, ja, a = orc.csr_from_dense(A)
    q = orc.lsqr_precon(ia, ja, a, 40, 40)
    x, info = orc.lsqr_solve(ia, ja, a, 40, 40, q, rng.normal(size=40), maxit=3, eps=1e-15)
    assert info["istop"] == 5 and info["ierr"] == 5 and info["nit"] == 3


def test_lsqr_hot_start(orc):
    rng = np.random.default_rng(5)
    A = rng.normal(size=(30, 30)) + 5 * np.eye(30)
    b = rng.normal(size=30)
    xr = np.linalg.solve(A, b)
    ia, ja, a = orc.csr_from_dense(A)
    q = orc.lsqr_precon(ia, ja, a, 30, 30)
    x, info = orc.lsqr_solve(ia, ja, a, 30, 30, q, b, x0=xr, eps=1e-14)
    assert info["nit"] <= 1 and np.allclose(x, xr, rtol=1e-12, atol=1e-12)


def test_stokes_rows_identities(orc):
    m = M.geodesic_sphere(3)
    T, G, prT, prG, err = orc.assemble_stokes(m, pressure=True)
    nn = m.nnodes
    # sum n . G = 0 on node rows (Pozrikidis 7.2.19, src/lib_Stokes.f90:2152-2178)
    assert err[0] < 1e-12
    # after the free term every diagonal T row sums to zero (rigid translation)
    for k in (0, 3, 5):
        assert np.abs(T[k].sum(1)).max() < 1e-12
    # pressure rows on centroid rows sum to zero after the correction (:1851-1863)
    for k in range(3):
        assert np.abs(prT[k][nn:].sum(1)).max() < 1e-11


def test_stokes_sphere_resistance(orc):
    m = M.geodesic_sphere(4)
    FT, nits, ierr = orc.resistance(m, exterior=True, eps=1e-12)
    assert ierr == 0
    K = FT[:3, :3] / np.pi
    Om = FT[3:, 3:] / np.pi
    C = FT[:3, 3:]
    assert np.abs(K - np.diag(np.diag(K))).max() < 1e-8 and np.abs(C).max() < 1e-8
    assert np.allclose(np.diag(K), 6.0, rtol=0.02)   # 6 pi mu a U, flat-element discretisation
    assert np.allclose(np.diag(Om), 8.0, rtol=0.05)  # 8 pi mu a^3 Omega


def test_bounded_formulation(orc):
    """The reference's own set-up (body at rest inside a far sphere that carries the flow)."""
    inner = M.geodesic_sphere(3, 0.5)
    m = M.body_in_sphere(inner, 32.0, 3)
    FT, nits, ierr = orc.resistance(m, exterior=False, eps=1e-12)
    assert ierr == 0
    F = np.diag(FT[:3, :3])
    # Stokes drag on a sphere of radius 0.5 held in a uniform stream, within the wall and
    # discretisation effects of this coarse set-up
    assert np.allclose(F, 6 * np.pi * 0.5, rtol=0.1)
