"""Pins of the oracle's point kernels and regular triangle integrals that do not restate them.

Point kernels (BemKernel, src/lib_sdBem.f90:911-936; StokesKernel, src/lib_Stokes.f90:2450-2544) are
checked against the differential relations that define them, by central finite differences in the
field point f:
* Laplace: H = n . grad_f G (with r = s - f, G = 1 / (4 pi |r|));
* Stokes (r = f - s): the Stokeslet is divergence-free, solves the Stokes equations with the pressure
  prG (lap_f G_ij = d_i prG_j), its stress is the stresslet, T_ij = (-prG_j delta_ik + d_k G_ij +
  d_i G_kj) n_k, and the stresslet pressure is prT_j = -2 n_k d_k prG_j.
Regular triangle integrals (Triangle_3DLapInt, src/lib_sdBem.f90:859-905; Triangle_3DStokesInt,
src/lib_Stokes.f90:2341-2446) are checked against brute force: the same point kernels integrated with
a rule of its own (the triangle cut into 4^k similar triangles, each integrated by a 16 x 16
Gauss-Legendre rule on the Duffy square), for sources at 3, 1 and 0.4 element sizes from the
element, with the shape-function weights of H and T.
"""
import ctypes as C

import numpy as np
import pytest


def _lap(orc, s, f, n):
    g, h = C.c_double(), C.c_double()
    orc.lib().orc_bem_kernel(orc._p(np.ascontiguousarray(s)), orc._p(np.ascontiguousarray(f)),
                             orc._p(np.ascontiguousarray(n)), C.byref(g), C.byref(h))
    return g.value, h.value


def _stk(orc, s, f, n):
    G, T, pg, pt = np.zeros(9), np.zeros(9), np.zeros(3), np.zeros(3)
    orc.lib().orc_stokes_kernel(orc._p(np.ascontiguousarray(s)), orc._p(np.ascontiguousarray(f)),
                                orc._p(np.ascontiguousarray(n)), orc._p(G), orc._p(T), orc._p(pg), orc._p(pt))
    return G.reshape(3, 3), T.reshape(3, 3), pg, pt


def _points(seed, k=12):
    rng = np.random.default_rng(seed)
    for _ in range(k):
        s, f, n = rng.standard_normal(3), rng.standard_normal(3), rng.standard_normal(3)
        yield s, f, n / np.linalg.norm(n)


def test_laplace_kernel_normal_derivative(orc):
    h = 1e-5
    for s, f, n in _points(1):
        G, H = _lap(orc, s, f, n)
        assert abs(G - 1.0 / (4 * np.pi * np.linalg.norm(s - f))) < 1e-15 * abs(G)
        fd = (_lap(orc, s, f + h * n, n)[0] - _lap(orc, s, f - h * n, n)[0]) / (2 * h)
        assert abs(H - fd) < 1e-8 * max(abs(H), abs(G))


def _dG(orc, s, f, n, k, h=1e-5):
    e = np.zeros(3)
    e[k] = h
    Gp, _, pp, _ = _stk(orc, s, f + e, n)
    Gm, _, pm, _ = _stk(orc, s, f - e, n)
    return (Gp - Gm) / (2 * h), (pp - pm) / (2 * h)


def test_stokes_kernel_differential_relations(orc):
    for s, f, n in _points(2):
        G, T, pg, pt = _stk(orc, s, f, n)
        scale = np.abs(G).max() / np.linalg.norm(f - s)
        dG = [_dG(orc, s, f, n, k) for k in range(3)]  # dG[k] = (d_k G, d_k prG)
        # symmetric and divergence-free
        assert np.abs(G - G.T).max() < 1e-15 * np.abs(G).max()
        div = sum(dG[k][0][:, k] for k in range(3))
        assert np.abs(div).max() < 1e-8 * scale
        # stress of the Stokeslet contracted with n is the stresslet
        Tfd = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                for k in range(3):
                    Tfd[i, j] += (-pg[j] * (i == k) + dG[k][0][i, j] + dG[i][0][k, j]) * n[k]
        assert np.abs(T - Tfd).max() < 1e-7 * np.abs(T).max()
        # stresslet pressure
        ptfd = -2.0 * sum(n[k] * dG[k][1] for k in range(3))
        assert np.abs(pt - ptfd).max() < 1e-7 * np.abs(pt).max()
        # momentum: lap_f G_ij = d_i prG_j (second differences, h = 1e-3)
        h = 1e-3
        lap = np.zeros((3, 3))
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            lap += (_stk(orc, s, f + e, n)[0] - 2 * G + _stk(orc, s, f - e, n)[0]) / h ** 2
        gradp = np.array([dG[i][1] for i in range(3)])  # gradp[i, j] = d_i prG_j
        assert np.abs(lap - gradp).max() < 1e-4 * np.abs(gradp).max()


# ---- regular triangle integrals against brute force -----------------------------------------------
def _subdivide(tri, k):
    tris = [tri]
    for _ in range(k):
        nxt = []
        for a, b, c in tris:
            ab, bc, ca = (a + b) / 2, (b + c) / 2, (c + a) / 2
            nxt += [(a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca)]
        tris = nxt
    return tris


def _brute(tri, k, fn):
    """sum over sub-triangles of the Duffy-mapped 16 x 16 Gauss-Legendre rule; fn(point, l) with l the
    barycentric coordinates of the point in the parent triangle."""
    x, w = np.polynomial.legendre.leggauss(16)
    x, w = 0.5 * (x + 1), 0.5 * w
    A = np.array(tri[:3])
    area = 0.5 * np.linalg.norm(np.cross(A[1] - A[0], A[2] - A[0]))
    M = np.array([A[1] - A[0], A[2] - A[0]]).T
    acc = 0.0
    for a, b, c in _subdivide(tuple(A), k):
        sub_area = 0.5 * np.linalg.norm(np.cross(b - a, c - a))
        for ui, wu in zip(x, w):
            for vi, wv in zip(x, w):
                l2, l3 = ui * (1 - vi), ui * vi  # Duffy: the square onto the triangle, jacobian u
                p = a + l2 * (b - a) + l3 * (c - a)
                lp = np.linalg.lstsq(M, p - A[0], rcond=None)[0]
                bary = np.array([1 - lp[0] - lp[1], lp[0], lp[1]])
                acc = acc + fn(p, bary) * (wu * wv * ui * 2 * sub_area)
    return acc, area


TRI = (np.array([0.0, 0.0, 0.0]), np.array([1.0, 0.1, 0.0]), np.array([0.3, 0.9, 0.05]))


# at 3 and 1 element sizes the 126-point rule is exact to rounding; at 0.4 its own quadrature error
# (~1e-6 relative, the same in the reference) is the difference; the brute force itself agrees with
# one subdivision level less to ~4e-15
@pytest.mark.parametrize("dist,tol", [(3.0, 1e-13), (1.0, 1e-12), (0.4, 3e-6)])
def test_triangle_laplace_integral_brute_force(orc, dist, tol):
    x1, x2, x3 = TRI
    nrm = np.cross(x2 - x1, x3 - x1)
    area = 0.5 * np.linalg.norm(nrm)
    nrm = nrm / np.linalg.norm(nrm)
    h = np.sqrt(2 * area)
    s = (x1 + x2 + x3) / 3 + dist * h * (0.6 * nrm + 0.8 * (x2 - x1) / np.linalg.norm(x2 - x1))
    got = orc.tri_lap_int(x1, x2, x3, s, nrm, area)

    def fn(p, l):
        G, H = _lap(orc, s, p, nrm)
        return np.array([G, H * l[0], H * l[1], H * l[2]])

    ref, _ = _brute(TRI, 3, fn)
    assert np.abs(got - ref).max() < tol * np.abs(ref).max(), (got, ref)


# at 3 and 1 element sizes the 126-point rule is exact to rounding; at 0.4 its own quadrature error
# (~1e-6 relative, the same in the reference) is the difference; the brute force itself agrees with
# one subdivision level less to ~4e-15
@pytest.mark.parametrize("dist,tol", [(3.0, 1e-13), (1.0, 1e-12), (0.4, 3e-6)])
def test_triangle_stokes_integral_brute_force(orc, dist, tol):
    x1, x2, x3 = TRI
    nrm = np.cross(x2 - x1, x3 - x1)
    area = 0.5 * np.linalg.norm(nrm)
    nrm = nrm / np.linalg.norm(nrm)
    h = np.sqrt(2 * area)
    s = (x1 + x2 + x3) / 3 + dist * h * (0.8 * nrm - 0.6 * (x3 - x1) / np.linalg.norm(x3 - x1))
    got = orc.tri_stokes_int(x1, x2, x3, s, nrm, area)
    sym = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]

    def fn(p, l):
        G, T, pg, pt = _stk(orc, s, p, nrm)
        out = [G[i, j] for i, j in sym] + list(pg)
        for i, j in sym:
            out += [T[i, j] * l[0], T[i, j] * l[1], T[i, j] * l[2]]
        for j in range(3):
            out += [pt[j] * l[0], pt[j] * l[1], pt[j] * l[2]]
        return np.array(out)

    ref, _ = _brute(TRI, 3, fn)
    assert got.shape == ref.shape == (36,)
    for blk in (slice(0, 6), slice(6, 9), slice(9, 27), slice(27, 36)):  # G, prG, T, prT on their own scales
        assert np.abs(got[blk] - ref[blk]).max() < tol * np.abs(ref[blk]).max(), blk
