"""CPU oracle of the learned rule selection's data (test infrastructure only: imported by tests/
and nothing else).

The reference has no learned part; it integrates every regular pair with one fixed rule
(Triangle_3DLapInt, src/lib_sdBem.f90:859-905; Triangle_3DStokesInt, src/lib_Stokes.f90:2341-2446).
This module restates, in float64 numpy, what the product's label generator and classifier
compute for one (collocation point = element centroid, element) pair:

* features: log(d / L), log(L^2 / (2 A)), cos(n_s, n_e), one-hot kernel (Laplace, Stokes-G,
  Stokes-H); d = distance from the collocation point to the element centroid, L = longest edge,
  A = area;
* kernel values with a given barycentric rule (l1, l2, l3, w), sum w = 1:
  Laplace (r = s - f, BemKernel src/lib_sdBem.f90:911-936) A sum w {G, H l1, H l2, H l3};
  Stokes-G (r = f - s, StokesKernel src/lib_Stokes.f90:2450-2544) A sum w G_ij (6 components);
  Stokes-H A sum w T_ij l_v (6 x 3);
* label: the first Dunavant rule (3, 4, 6, 7, 12, 13, 19 points) whose values agree with the
  fixed-rule values to rel_tol * max|fixed values| in max norm, else 7;
* far bin: floor(2 log2(d / (L near_ratio))) clamped to 0..15.

The point kernels are checked against the C oracle's BemKernel/StokesKernel in the tests.
"""
import numpy as np

FOUR_PI = 4.0 * np.pi
DUN_POINTS = (3, 4, 6, 7, 12, 13, 19)
FAR_BINS = 16


def lap_point(s, f, n):
    """G, H of BemKernel at many points: s (3,), f (..., 3), n (3,)."""
    r = s - f
    ra = np.sqrt((r * r).sum(-1))
    G = 1.0 / (FOUR_PI * ra)
    H = (r @ n) / ra ** 2 * G
    return G, H


def stokes_point(s, f, n):
    """Stokeslet G (..., 3, 3) and stresslet-normal T (..., 3, 3) of StokesKernel (no 1/8pi)."""
    r = f - s
    ra = np.sqrt((r * r).sum(-1))
    rr = r[..., :, None] * r[..., None, :]
    G = np.eye(3) / ra[..., None, None] + rr / ra[..., None, None] ** 3
    T = -6.0 * rr * (r @ n)[..., None, None] / ra[..., None, None] ** 5
    return G, T


SYM = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


def kernel_values(kernel, s, x, n, area, rule):
    """Values of one kernel family for one element (x: 3 vertices) with rule (nq, 4)."""
    l = rule[:, :3]
    w = rule[:, 3]
    f = l @ x
    if kernel == 0:
        G, H = lap_point(s, f, n)
        return area * np.array([w @ G, w @ (H * l[:, 0]), w @ (H * l[:, 1]), w @ (H * l[:, 2])])
    G, T = stokes_point(s, f, n)
    if kernel == 1:
        return area * np.array([w @ G[:, i, j] for i, j in SYM])
    return area * np.array([w @ (T[:, i, j] * l[:, v]) for i, j in SYM for v in range(3)])


def geometry(xyz, con):
    x = xyz[con]
    c = np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0])
    area = 0.5 * np.linalg.norm(c, axis=1)
    nrm = c / np.linalg.norm(c, axis=1)[:, None]
    return x, nrm, area, x.mean(axis=1)


def pair_geo(s, ns, x, ne, area):
    d = np.linalg.norm(s - x.mean(axis=0))
    L2 = max(((x[0] - x[1]) ** 2).sum(), ((x[1] - x[2]) ** 2).sum(), ((x[2] - x[0]) ** 2).sum())
    return d / np.sqrt(L2), L2 / (2.0 * area), float(ns @ ne)


def features(dh, aspect, cosn, kernel):
    f = np.zeros(6)
    f[0], f[1], f[2] = np.log(dh), np.log(aspect), cosn
    f[3 + kernel] = 1.0
    return f


def label(kernel, s, x, n, area, fixed_rule, dun_rules, rel_tol):
    ref = kernel_values(kernel, s, x, n, area, fixed_rule)
    scale = np.abs(ref).max()
    for c, rule in enumerate(dun_rules):
        if np.abs(kernel_values(kernel, s, x, n, area, rule) - ref).max() <= rel_tol * scale:
            return c
    return len(dun_rules)


def far_bin(dh, near_ratio):
    b = int(np.floor(2.0 * np.log2(dh / near_ratio)))
    return min(max(b, 0), FAR_BINS - 1)


def near_records(xyz, con, fixed_rule, dun_rules, near_ratio, rel_tol):
    """(features, label) of every near pair and kernel, in (row, element, kernel) order."""
    x, nrm, area, cen = geometry(xyz, con)
    F, L = [], []
    for r in range(len(con)):
        for e in range(len(con)):
            if e == r:
                continue
            dh, asp, cs = pair_geo(cen[r], nrm[r], x[e], nrm[e], area[e])
            if dh >= near_ratio:
                continue
            for k in range(3):
                F.append(features(dh, asp, cs, k))
                L.append(label(k, cen[r], x[e], nrm[e], area[e], fixed_rule, dun_rules, rel_tol))
    return np.array(F), np.array(L, dtype=np.int64)
