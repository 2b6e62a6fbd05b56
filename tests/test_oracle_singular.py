"""Pins the analytic singular integrals of the oracle (Ren-Chan forms restated from
src/mod_singInt.f90) against (a) a 1-D reduction of the surface integral integrated by adaptive
quadrature, (b) rotation/translation/permutation invariance; the reference's own closed-form value
ln(1 + sqrt 2)/(2 pi) is checked in test_oracle_quadrature.

The reference's compiled si_tri_lap / si_tri_stk cannot serve as golden vectors here: siNormVector
(src/mod_singInt.f90:1141-1155) accumulates into a local that is never initialised, and the
amdflang build returns NaN at -O2 and stack-dependent values at -O0. The oracle starts the
accumulator at zero."""
import os

import numpy as np
import pytest
from scipy import integrate

HERE = os.path.dirname(os.path.abspath(__file__))


def _polar_reference(x1, x2, x3, P, isrc):
    """int 1/r dA and int (I + e e^T)/r dA over the triangle, as sums over the sub-triangles
    (P, A, B): with X(t) = A + t (B - A) the area element is h L / rho(t)^2 dt rho drho."""
    edges = [(x2, x3, 1), (x3, x1, 2), (x1, x2, 3)]
    I0 = 0.0
    M = np.zeros((3, 3))
    for A, B, k in edges:
        if isrc != 4 and k != isrc:  # a vertex source sees only the opposite edge
            continue
        A, B = np.asarray(A, float), np.asarray(B, float)
        L = np.linalg.norm(B - A)
        t_ = (B - A) / L
        h = np.linalg.norm(np.cross(A - P, t_))

        def rho(t):
            return np.linalg.norm(A + t * (B - A) - P)

        I0 += integrate.quad(lambda t: h * L / rho(t), 0, 1, epsabs=0, epsrel=1e-13, limit=200)[0]
        for i in range(3):
            for j in range(i, 3):
                v = integrate.quad(lambda t: h * L * (A[i] + t * (B[i] - A[i]) - P[i]) *
                                   (A[j] + t * (B[j] - A[j]) - P[j]) / rho(t) ** 3, 0, 1,
                                   epsabs=0, epsrel=1e-13, limit=200)[0]
                M[i, j] += v
                if i != j:
                    M[j, i] += v
    return I0, I0 * np.eye(3) + M


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_polar_reduction(orc, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(3, 3))
    area = 0.5 * np.linalg.norm(np.cross(x[1] - x[0], x[2] - x[0]))
    cen = x.mean(0)
    for isrc in range(1, 5):
        P = cen if isrc == 4 else x[isrc - 1]
        I0, M = _polar_reference(x[0], x[1], x[2], P, isrc)
        lap = orc.si_tri_lap(x[0], x[1], x[2], area, isrc)
        assert abs(lap - I0 / (4 * np.pi)) < 1e-12 * abs(lap)
        s = orc.si_tri_stk(x[0], x[1], x[2], area, isrc)
        S = np.array([[s[0], s[1], s[2]], [s[1], s[3], s[4]], [s[2], s[4], s[5]]])
        assert np.abs(S - M).max() < 1e-11 * np.abs(M).max()


def test_invariance(orc):
    rng = np.random.default_rng(7)
    x = rng.normal(size=(3, 3))
    area = 0.5 * np.linalg.norm(np.cross(x[1] - x[0], x[2] - x[0]))
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    y = x @ Q.T + np.array([0.3, -1.0, 2.0])
    for isrc in range(1, 5):
        a = orc.si_tri_lap(x[0], x[1], x[2], area, isrc)
        b = orc.si_tri_lap(y[0], y[1], y[2], area, isrc)
        assert abs(a - b) < 1e-13 * abs(a)
        s = orc.si_tri_stk(x[0], x[1], x[2], area, isrc)
        S = np.array([[s[0], s[1], s[2]], [s[1], s[3], s[4]], [s[2], s[4], s[5]]])
        t = orc.si_tri_stk(y[0], y[1], y[2], area, isrc)
        T = np.array([[t[0], t[1], t[2]], [t[1], t[3], t[4]], [t[2], t[4], t[5]]])
        assert np.abs(Q @ S @ Q.T - T).max() < 1e-12 * np.abs(S).max()
    # vertex permutation: source at vertex 2 of (x1,x2,x3) == source at vertex 1 of (x2,x3,x1)
    a = orc.si_tri_lap(x[0], x[1], x[2], area, 2)
    b = orc.si_tri_lap(x[1], x[2], x[0], area, 1)
    assert abs(a - b) < 1e-13 * abs(a)
