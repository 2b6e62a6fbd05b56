"""Quadrilateral elements: Gauss-Legendre tables, regular sub-cell integration of bilinear
elements, singular integrals split into triangles, and the assembled matrices on quadrilateral and
mixed meshes.

Oracle pins (CPU): the tables against numpy's Gauss-Legendre nodes; regular integrals against an
independent tensor Gauss-Legendre integration of the bilinear map on a warped (non-planar) element;
singular integrals against closed forms over squares; assembled matrices against the rigid-body and
u = x identities and the sphere drag. GPU: the HIP assembly against the oracle at 1e-10.
"""
import math

import numpy as np
import pytest

from andromeda_amd import mesh as M

L1 = math.log(1.0 + math.sqrt(2.0))  # int over the unit square of 1/r from a corner = 2 L1
K4PI = 1.0 / (4.0 * math.pi)
SQUARE = [np.array(v, float) for v in ([0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0])]
WARPED = [np.array(v, float) for v in ([0, 0, 0], [1.3, 0, 0.1], [1.2, 1.1, -0.1], [-0.1, 0.9, 0.05])]


@pytest.mark.parametrize("iq,n", [(1, 2), (2, 4), (3, 6), (4, 8), (5, 16), (6, 32), (7, 40), (8, 48)])
def test_gauss_tables(orc, iq, n):
    x, w = orc.gauss_rule(iq)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert x.size == n
    assert np.abs(x - xr).max() < 1e-14 and np.abs(w - wr).max() < 1e-14


def _bilinear(x, xi, et):
    """point, shape functions and x_xi x x_eta of the bilinear element (independent of the oracle)."""
    phi = 0.25 * np.array([(1 - xi) * (1 - et), (1 + xi) * (1 - et), (1 + xi) * (1 + et), (1 - xi) * (1 + et)])
    dxi = 0.25 * np.array([-(1 - et), (1 - et), (1 + et), -(1 + et)])
    det = 0.25 * np.array([-(1 - xi), -(1 + xi), (1 + xi), (1 - xi)])
    X = np.array(x)
    return phi @ X, phi, np.cross(dxi @ X, det @ X)


def _brute(x, fn, n=40, cells=4):
    """tensor Gauss-Legendre over cells x cells sub-squares of [-1, 1]^2, fn(f, phi, an) -> array."""
    g, w = np.polynomial.legendre.leggauss(n)
    tot = 0.0
    h = 2.0 / cells
    for a in range(cells):
        for b in range(cells):
            for i in range(n):
                for j in range(n):
                    xi = -1 + h * (a + 0.5 * (g[i] + 1))
                    et = -1 + h * (b + 0.5 * (g[j] + 1))
                    f, phi, an = _bilinear(x, xi, et)
                    tot = tot + fn(f, phi, an) * w[i] * w[j] * (h / 2) ** 2
    return tot


def test_quad_laplace_regular_vs_brute_force(orc):
    s = np.array([0.5, 0.4, 1.5])

    def lap(f, phi, an):  # G with dA = |an|; H_k = (r.n) / (4 pi r^3) phi_k dA, r = s - f
        r = s - f
        d = np.linalg.norm(r)
        return np.concatenate([[K4PI / d * np.linalg.norm(an)], K4PI * (r @ an) / d ** 3 * phi])

    ref = _brute(WARPED, lap)
    for iq in (4, 8):
        out = orc.quad_lap_int(*WARPED, s, 1.0, iq)
        assert np.abs(out - ref).max() < 1e-9 * np.abs(ref).max(), iq
    # the wall factor multiplies the normal, so only H changes sign
    outm = orc.quad_lap_int(*WARPED, s, -1.0, 4)
    assert np.allclose(outm[0], ref[0], rtol=1e-9) and np.allclose(outm[1:], -ref[1:], rtol=1e-9)


def test_quad_laplace_cells_scale_with_size(orc):
    """elements larger than one unit get more sub-cells (minedge = 1): a 6 x 6 element near the
    source stays accurate where 2 x 2 cells of 8 points would not be"""
    big = [6.0 * v for v in SQUARE]
    s = np.array([3.1, 2.9, 0.7])

    def lap(f, phi, an):
        r = s - f
        d = np.linalg.norm(r)
        return np.concatenate([[K4PI / d * np.linalg.norm(an)], K4PI * (r @ an) / d ** 3 * phi])

    ref = _brute(big, lap, n=48, cells=12)
    out = orc.quad_lap_int(*big, s, 1.0, 4)
    assert np.abs(out - ref).max() < 1e-6 * np.abs(ref).max()


def test_quad_stokes_regular_vs_brute_force(orc):
    s = np.array([0.6, 0.3, -1.4])

    def stk(f, phi, an):  # r = f - s; G, prG with dA; T, prT with the pointwise normal
        r = f - s
        d = np.linalg.norm(r)
        J = np.linalg.norm(an)
        n = an / J
        G = (np.eye(3) / d + np.outer(r, r) / d ** 3)[[0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]
        T = (-6.0 * np.outer(r, r) * (r @ n) / d ** 5)[[0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]
        pG = 2.0 * r / d ** 3
        pT = 4.0 * (-n / d ** 3 + 3.0 * r * (r @ n) / d ** 5)
        return np.concatenate([G * J, pG * J, (np.outer(T, phi) * J).ravel(), (np.outer(pT, phi) * J).ravel()])

    ref = _brute(WARPED, stk)
    acc = orc.quad_stokes_int(*WARPED, s, 1.0, 4)[:45]
    assert np.abs(acc - ref).max() < 1e-9 * np.abs(ref).max()


@pytest.mark.parametrize("isrc,sign", [(1, 1.0), (2, -1.0), (3, 1.0), (4, -1.0)])
def test_quad_singular_vertex_closed_forms(orc, isrc, sign):
    """unit square, source at a corner: int 1/r = 2 ln(1 + sqrt 2); Stokeslet xx = yy = 3 ln(1 + sqrt 2),
    zz = 2 ln(1 + sqrt 2), xy = +-(2 - sqrt 2) by the quadrant the square lies in"""
    assert orc.si_qua_lap(*SQUARE, isrc) == pytest.approx(2 * L1 * K4PI, rel=1e-13)
    g = orc.si_qua_stk(*SQUARE, isrc)
    want = [3 * L1, sign * (2 - math.sqrt(2)), 0.0, 3 * L1, 0.0, 2 * L1]
    assert np.allclose(g, want, rtol=1e-13, atol=1e-14)


def test_quad_singular_centre_closed_forms(orc):
    sq2 = [2.0 * v - np.array([1.0, 1.0, 0.0]) for v in SQUARE]  # side 2, centred at the origin
    assert orc.si_qua_lap(*sq2, 5) == pytest.approx(8 * L1 * K4PI, rel=1e-13)
    g = orc.si_qua_stk(*sq2, 5)
    assert np.allclose(g, [12 * L1, 0, 0, 12 * L1, 0, 8 * L1], rtol=1e-13, atol=1e-13)
    # rigid rotation of the element rotates the block
    c, s_ = math.cos(0.7), math.sin(0.7)
    R = np.array([[c, -s_, 0], [s_, c, 0], [0, 0, 1]]) @ np.array([[1, 0, 0], [0, c, -s_], [0, s_, c]])
    gr = orc.si_qua_stk(*[R @ v for v in sq2], 5)
    full = lambda v: np.array([[v[0], v[1], v[2]], [v[1], v[3], v[4]], [v[2], v[4], v[5]]])
    assert np.allclose(full(gr), R @ full(g) @ R.T, rtol=1e-12, atol=1e-12)


def test_cubed_sphere_geometry(orc):
    deficit = []
    for n in (6, 12):
        m = M.cubed_sphere(n)
        nrm, area, cen = orc.geometry(m)
        deficit.append(4 * math.pi - area.sum())  # bilinear facets lie inside the sphere
        assert np.all(np.einsum("ij,ij->i", nrm, cen) > 0)  # outward
        assert np.allclose(cen, m.xyz[m.con].mean(axis=1))
    for k in (0, 7, 100):  # area = integral of |x_xi x x_eta| (8 x 8 Gauss-Legendre)
        x = [m.xyz[v] for v in m.con[k]]
        assert area[k] == pytest.approx(_brute(x, lambda f, phi, an: np.linalg.norm(an), n=8, cells=1), rel=1e-14)
    assert 0 < deficit[1] < 0.3 * deficit[0]  # second order in the element size


def test_cubed_sphere_laplace_identities(orc):
    """u = x gives q = n_x: the oracle's own check (returned err) falls with refinement; node rows
    of H sum to zero by the free term"""
    errs = []
    for n in (3, 6):
        m = M.cubed_sphere(n)
        H, G, err = orc.assemble_laplace(m)
        assert np.abs(H[: m.nnodes].sum(axis=1)).max() < 1e-12
        errs.append(err[0])
    assert errs[1] < 0.75 * errs[0] and errs[1] < 1e-2


def test_mixed_mesh_centroid_rows_spread(orc):
    """centroid rows: the free term goes 1/4 to each node of a quadrilateral and 1/3 to each node of
    a triangle, so every centroid row of H sums to zero as well"""
    m = M.split_some_quads(M.cubed_sphere(3), 4)
    H, G, err = orc.assemble_laplace(m)
    assert np.abs(H.sum(axis=1)).max() < 1e-12
    T, _, _, _, _ = orc.assemble_stokes(m)
    for k in (0, 3, 5):
        assert np.abs(T[k].sum(axis=1)).max() < 1e-10


def test_cubed_sphere_drag(orc):
    m = M.cubed_sphere(6)
    FT, nit, ierr = orc.resistance(m, exterior=True)
    assert FT[0, 0] == pytest.approx(6 * math.pi, rel=0.03)
    assert FT[3, 3] == pytest.approx(8 * math.pi, rel=0.1)


# ---- GPU: the HIP assembly against the oracle ---------------------------------------------------
def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


QUAD_MESHES = {
    "cubed4": lambda: M.cubed_sphere(4),
    "mixed": lambda: M.split_some_quads(M.cubed_sphere(3), 3),
    "large": lambda: M.cubed_sphere(2, radius=6.0),  # sides ~4.7: 5 x 5 sub-cells in Laplace pairs
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(QUAD_MESHES))
def test_gpu_quad_laplace(ctx, orc, name):
    m = QUAD_MESHES[name]()
    H, G, err = ctx.assemble_laplace(m)
    Ho, Go, erro = orc.assemble_laplace(m)
    assert _rel(H, Ho) < 1e-10 and _rel(G, Go) < 1e-10
    assert abs(err[0] - erro[0]) < 1e-9 * max(1.0, abs(erro[0]))


@pytest.mark.gpu
@pytest.mark.parametrize("name,iquad", [("cubed4", 4), ("mixed", 4), ("mixed", 2)])
def test_gpu_quad_stokes(ctx, orc, name, iquad):
    m = QUAD_MESHES[name]()
    T, G, pT, pG, err = ctx.assemble_stokes(m, pressure=True, iquad=iquad)
    To, Go, pTo, pGo, erro = orc.assemble_stokes(m, pressure=True, iquad=iquad)
    for k in range(6):
        assert _rel(T[k], To[k]) < 1e-10 and _rel(G[k], Go[k]) < 1e-10, k
    for k in range(3):
        assert _rel(pT[k], pTo[k]) < 1e-10 and _rel(pG[k], pGo[k]) < 1e-10, k


@pytest.mark.gpu
def test_gpu_quad_iquad_changes_rule(ctx, orc):
    m = M.cubed_sphere(3)
    H2, _, _ = ctx.assemble_laplace(m, iquad=1)
    H4, _, _ = ctx.assemble_laplace(m, iquad=4)
    Ho, _, _ = orc.assemble_laplace(m, iquad=1)
    assert _rel(H2, Ho) < 1e-10 and _rel(H2, H4) > 1e-8


@pytest.mark.gpu
def test_gpu_quad_mesh_rejected_by_resistance(ctx):
    from andromeda_amd.api import AndromedaError
    with pytest.raises(AndromedaError, match="triangle meshes"):
        ctx.resistance(M.cubed_sphere(3))
