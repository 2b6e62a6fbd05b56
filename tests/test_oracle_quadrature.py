"""Pins the quadrature tables of the oracle (shared with the product): polynomial exactness to
the rule degree, the closed forms of Triangle_TestIntegration (src/mod_trikotnik.f90:108-191), and
point-set equality with the reference's own compiled table (oracle/_ref) when it is built."""
import math

import numpy as np
import pytest

DEG = {1: 5, 2: 10, 3: 15, 4: 20, 5: 25}
NPTS = {1: 7, 2: 25, 3: 54, 4: 85, 5: 126}


@pytest.mark.parametrize("itri", [1, 2, 3, 4, 5])
def test_wx_exactness(orc, itri):
    l1, l2, l3, w = orc.tri_rule(itri)
    assert len(w) == NPTS[itri]
    assert abs(w.sum() - 1.0) < 1e-14
    assert np.abs(l1 + l2 + l3 - 1.0).max() < 1e-15
    worst = 0.0
    for i in range(DEG[itri] + 1):
        for j in range(DEG[itri] + 1 - i):
            exact = 2.0 * math.factorial(i) * math.factorial(j) / math.factorial(i + j + 2)
            worst = max(worst, abs(np.sum(w * l1 ** i * l2 ** j) - exact) / exact)
    assert worst < 1e-13
    # one degree higher is not integrated exactly: the table is not a different, higher rule
    d = DEG[itri] + 1
    ex = [2.0 * math.factorial(i) * math.factorial(d - i) / math.factorial(d + 2) for i in range(d + 1)]
    errs = [abs(np.sum(w * l1 ** i * l2 ** (d - i)) - ex[i]) / ex[i] for i in range(d + 1)]
    assert max(errs) > 1e-9


def _tri_integral(itri, f, x1, x2, x3, orc):
    l1, l2, l3, w = orc.tri_rule(itri)
    p = np.outer(l1, x1) + np.outer(l2, x2) + np.outer(l3, x3)
    area = 0.5 * np.linalg.norm(np.cross(np.subtract(x2, x1), np.subtract(x3, x1)))
    return np.sum(w * f(p)) * area


def test_reference_closed_forms(orc):
    # src/mod_trikotnik.f90:124-155: triangle (0,0,1), (2,0,1), (2,2,1)
    x1, x2, x3 = (0.0, 0.0, 1.0), (2.0, 0.0, 1.0), (2.0, 2.0, 1.0)
    v = _tri_integral(5, lambda p: p[:, 0] ** 10 * p[:, 1] ** 10, x1, x2, x3, orc)
    assert abs(v - 2097152.0 / 121.0) / (2097152.0 / 121.0) < 1e-12
    v = _tri_integral(5, lambda p: np.sin(p[:, 0] * p[:, 1]), x1, x2, x3, orc)
    assert abs(v - 1.0522458619541769455) < 1e-12
    v = _tri_integral(5, lambda p: np.exp(-p[:, 0] * p[:, 1]), x1, x2, x3, orc)
    assert abs(v - 0.98364468921563619296) < 1e-12
    # 1/(4 pi r) with the source at vertex 1: the analytic singular integral, ln(1+sqrt 2)/(2 pi)
    g = orc.si_tri_lap(x1, x2, x3, 2.0, 1)
    assert abs(g - 0.14027496308479503178) < 1e-15
    assert abs(g - math.log(1.0 + math.sqrt(2.0)) / (2.0 * math.pi)) < 1e-15


@pytest.mark.parametrize("itri", [1, 2, 3, 4, 5])
def test_wx_matches_reference_table(orc, itri):
    import ctypes as C
    ref = orc.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    b = [np.zeros(126) for _ in range(4)]
    n = ref.ref_tri_rule(itri, *[x.ctypes.data_as(C.POINTER(C.c_double)) for x in b])
    B = np.stack([x[:n] for x in b], 1)
    A = np.stack(orc.tri_rule(itri), 1)
    assert A.shape == B.shape
    used = set()
    for r in A:
        d = np.abs(B - r).max(1)
        j = int(np.argmin(d))
        assert d[j] < 1e-15 * 8
        used.add(j)
    assert len(used) == n


def test_dunavant_exactness():
    import os
    import re
    src = open(os.path.join(os.path.dirname(__file__), "..", "andromeda_amd", "csrc", "andr_dunavant_rules.h")).read()
    rows = re.findall(r"\{([-0-9.e]+), ([-0-9.e]+), ([-0-9.e]+), ([-0-9.e]+)\}", src)
    P = np.array([[float(x) for x in r] for r in rows])
    assert P.shape == (64, 4)
    offs = [0, 3, 7, 13, 20, 32, 45, 64]
    degs = [2, 3, 4, 5, 6, 7, 9]
    for r in range(7):
        q = P[offs[r]:offs[r + 1]]
        assert abs(q[:, 3].sum() - 1.0) < 1e-15
        for i in range(degs[r] + 1):
            for j in range(degs[r] + 1 - i):
                exact = 2.0 * math.factorial(i) * math.factorial(j) / math.factorial(i + j + 2)
                assert abs(np.sum(q[:, 3] * q[:, 0] ** i * q[:, 1] ** j) - exact) < 1e-15


def test_wx_matches_reference_golden(orc):
    """Golden tables written from the reference's compiled module (tests/golden/make_golden.py)."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_tables.npz"))
    for itri in range(1, 6):
        B = g[f"rule{itri}"]
        A = np.stack(orc.tri_rule(itri), 1)
        assert A.shape == B.shape
        used = set()
        for r in A:
            d = np.abs(B - r).max(1)
            j = int(np.argmin(d))
            assert d[j] < 1e-14
            used.add(j)
        assert len(used) == B.shape[0]


def test_gauss_legendre_reference_golden():
    """src/mod_gauss.f90 tables (used by quadrilaterals) against numpy's Golub-Welsch nodes."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_tables.npz"))
    for r in range(8):
        a, b = int(g["gauss_ng1"][r]) - 1, int(g["gauss_ng2"][r])
        x, w = np.polynomial.legendre.leggauss(b - a)
        gi, om = g["gauss_gi"][a:b], g["gauss_ome"][a:b]
        o = np.argsort(gi)
        assert np.abs(gi[o] - x).max() < 1e-14 and np.abs(om[o] - w).max() < 1e-14
