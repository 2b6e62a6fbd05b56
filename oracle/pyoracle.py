"""ctypes bindings of the CPU oracle (oracle/liborc.so) and of the reference's own compiled
modules (oracle/_ref/libandr_ref.so). TEST INFRASTRUCTURE ONLY: imported by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the product."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)
_L = C.POINTER(C.c_int64)


def _p(a):
    return a.ctypes.data_as(_D)


def _pi(a):
    return a.ctypes.data_as(_I)


class _Body(C.Structure):
    _fields_ = [("nnodes", C.c_int), ("nelem", C.c_int), ("nwalls", C.c_int),
                ("xyz", _D), ("con", _I), ("bcid", _I), ("normmul", _D), ("itri", C.c_int),
                ("nvert", C.c_int), ("iquad", C.c_int)]


class _Csr(C.Structure):
    _fields_ = [("m", C.c_int64), ("n", C.c_int64), ("nnz", C.c_int64),
                ("ia", _L), ("ja", _I), ("a", _D)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liborc.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _lib = C.CDLL(path)
        _lib.orc_si_tri_lap.restype = C.c_double
        _lib.orc_si_qua_lap.restype = C.c_double
    return _lib


def ref_lib():
    """The reference's modules built from its sources (None when not built)."""
    path = os.path.join(HERE, "_ref", "libandr_ref.so")
    if not os.path.exists(path):
        return None
    r = C.CDLL(path)
    r.ref_si_tri_lap.restype = C.c_double
    r.ref_tri_area.restype = C.c_double
    return r


class Body:
    """Keeps the numpy arrays alive for the C struct."""

    def __init__(self, mesh, normmul=None, itri: int = 5, iquad: int = 4):
        m = mesh.contiguous()
        self.mesh = m
        self.normmul = np.ascontiguousarray(normmul if normmul is not None else np.ones(m.nwalls), dtype=np.float64)
        self.s = _Body(m.nnodes, m.nelem, m.nwalls, _p(m.xyz), _pi(m.con), _pi(m.bcid), _p(self.normmul), itri,
                       m.con.shape[1], iquad)


def tri_rule(itri: int):
    a = [np.zeros(126) for _ in range(4)]
    n = lib().orc_tri_rule(itri, *[_p(x) for x in a])
    return [x[:n].copy() for x in a]


def geometry(mesh):
    m = mesh.contiguous()
    nrm = np.zeros((m.nelem, 3)); area = np.zeros(m.nelem); cen = np.zeros((m.nelem, 3))
    lib().orc_geometry_n(m.nnodes, m.nelem, _p(m.xyz), _pi(m.con), m.con.shape[1], _p(nrm), _p(area), _p(cen))
    return nrm, area, cen


def gauss_rule(iq: int):
    x = np.zeros(48); w = np.zeros(48)
    n = lib().orc_gauss_rule(iq, _p(x), _p(w))
    return x[:n].copy(), w[:n].copy()


def _q4(x1, x2, x3, x4):
    return [np.ascontiguousarray(v, dtype=np.float64) for v in (x1, x2, x3, x4)]


def quad_lap_int(x1, x2, x3, x4, s, mul=1.0, iq=4):
    x = _q4(x1, x2, x3, x4)
    out = np.zeros(5)
    lib().orc_quad_lap_int(*[_p(v) for v in x], _p(np.ascontiguousarray(s, dtype=np.float64)), C.c_double(mul), iq,
                           _p(out))
    return out


def si_qua_lap(x1, x2, x3, x4, isrc):
    return lib().orc_si_qua_lap(*[_p(v) for v in _q4(x1, x2, x3, x4)], isrc)


def quad_stokes_int(x1, x2, x3, x4, s, mul=1.0, iq=4):
    x = _q4(x1, x2, x3, x4)
    acc = np.zeros(48)
    lib().orc_quad_stokes_int(*[_p(v) for v in x], _p(np.ascontiguousarray(s, dtype=np.float64)), C.c_double(mul),
                              iq, _p(acc))
    return acc


def si_qua_stk(x1, x2, x3, x4, isrc):
    g = np.zeros(6)
    lib().orc_si_qua_stk(*[_p(v) for v in _q4(x1, x2, x3, x4)], isrc, _p(g))
    return g


def si_tri_lap(x1, x2, x3, area, isrc):
    x = [np.ascontiguousarray(v, dtype=np.float64) for v in (x1, x2, x3)]
    return lib().orc_si_tri_lap(_p(x[0]), _p(x[1]), _p(x[2]), C.c_double(area), isrc)


def si_tri_stk(x1, x2, x3, area, isrc):
    x = [np.ascontiguousarray(v, dtype=np.float64) for v in (x1, x2, x3)]
    g = np.zeros(6)
    lib().orc_si_tri_stk(_p(x[0]), _p(x[1]), _p(x[2]), C.c_double(area), isrc, _p(g))
    return g


def tri_lap_int(x1, x2, x3, s, n, area):
    x = [np.ascontiguousarray(v, dtype=np.float64) for v in (x1, x2, x3, s, n)]
    out = np.zeros(4)
    lib().orc_tri_lap_int(_p(x[0]), _p(x[1]), _p(x[2]), _p(x[3]), _p(x[4]), C.c_double(area), _p(out))
    return out


def tri_stokes_int(x1, x2, x3, s, n, area):
    x = [np.ascontiguousarray(v, dtype=np.float64) for v in (x1, x2, x3, s, n)]
    out = np.zeros(36)
    lib().orc_tri_stokes_int(_p(x[0]), _p(x[1]), _p(x[2]), _p(x[3]), _p(x[4]), C.c_double(area), _p(out))
    return out


def assemble_laplace(mesh, normmul=None, itri=5, iquad=4):
    b = Body(mesh, normmul, itri, iquad)
    nsp = b.mesh.nnodes + b.mesh.nelem
    H = np.zeros((nsp, b.mesh.nnodes)); G = np.zeros((nsp, b.mesh.nelem)); err = np.zeros(2)
    lib().orc_assemble_laplace(C.byref(b.s), _p(H), _p(G), _p(err))
    return H, G, err


def assemble_stokes(mesh, normmul=None, rows=None, extra=0.0, itri=5, pressure=False, iquad=4):
    b = Body(mesh, normmul, itri, iquad)
    nn, ne = b.mesh.nnodes, b.mesh.nelem
    nsp = nn + ne
    rows_a = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
    nr = nsp if rows_a is None else rows_a.size
    T = [np.zeros((nr, nn)) for _ in range(6)]
    G = [np.zeros((nr, ne)) for _ in range(6)]
    prT = [np.zeros((nr, nn)) for _ in range(3)] if pressure else None
    prG = [np.zeros((nr, ne)) for _ in range(3)] if pressure else None
    err = np.zeros(2)
    Tp = (_D * 6)(*[_p(x) for x in T]); Gp = (_D * 6)(*[_p(x) for x in G])
    prTp = (_D * 3)(*[_p(x) for x in prT]) if pressure else None
    prGp = (_D * 3)(*[_p(x) for x in prG]) if pressure else None
    lib().orc_assemble_stokes(C.byref(b.s), nr, _pi(rows_a) if rows_a is not None else None,
                              C.c_double(extra), Tp, Gp, prTp, prGp, _p(err))
    return T, G, prT, prG, err


def stokes_domain_values(mesh, pts, u, q, diff=1.0, normmul=None, itri=5):
    """Velocity (npts, 3), pressure (npts,) and closure error (npts,) at domain points."""
    b = Body(mesh, normmul, itri)
    P = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    U = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 3)
    Q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
    vel = np.zeros((P.shape[0], 3)); p = np.zeros(P.shape[0]); err = np.zeros(P.shape[0])
    lib().orc_stokes_domain_values(C.byref(b.s), P.shape[0], _p(P), _p(U), _p(Q), C.c_double(diff), _p(vel), _p(p),
                                   _p(err))
    return vel, p, err


def stokes_surface_pressure(mesh, u, q, diff=1.0, normmul=None, itri=5):
    """pressureStokes (src/lib_Stokes.f90:4-94): pressure at the element centres from the centroid
    rows of prT (with the row-sum correction) and prG, p = -2 (diff prT.(u - u_c) - prG.q) / 8pi,
    u_c the velocity interpolated to the centre (InterpolateToQMesh, src/lib_mesh.f90:635-690:
    the mean of the three nodes), 8pi from a REAL(4) ATAN(1.0)."""
    m = mesh.contiguous()
    nn, ne = m.nnodes, m.nelem
    _, _, prT, prG, _ = assemble_stokes(m, normmul, rows=np.arange(nn, nn + ne), itri=itri, pressure=True)
    U = np.asarray(u, dtype=np.float64).reshape(nn, 3)
    Q = np.asarray(q, dtype=np.float64).reshape(ne, 3)
    uc = U[m.con].mean(axis=1)
    osem = float(np.arctan(np.float32(1.0))) * 4.0 * 8.0
    p = np.zeros(ne)
    for k in range(3):
        p += diff * (prT[k] @ U[:, k] - prT[k].sum(axis=1) * uc[:, k])
        p -= prG[k] @ Q[:, k]
    return -p / osem * 2.0


def csr_from_dense(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    m, n = A.shape
    ia = np.arange(0, m * n + 1, n, dtype=np.int64)
    ja = np.tile(np.arange(n, dtype=np.int32), m)
    return ia, ja, A.reshape(-1).copy()


def lsqr_precon(ia, ja, a, m, n):
    csr = _Csr(m, n, a.size, ia.ctypes.data_as(_L), _pi(ja), _p(a))
    q = np.zeros(n)
    lib().orc_lsqr_precon(C.byref(csr), _p(q))
    return q


def lsqr_solve(ia, ja, a, m, n, q, b, x0=None, maxit=0, eps=1e-15):
    csr = _Csr(m, n, a.size, ia.ctypes.data_as(_L), _pi(ja), _p(a))
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    mi = C.c_int(maxit); nit = C.c_int(0); ist = C.c_int(0)
    b = np.ascontiguousarray(b, dtype=np.float64)
    ierr = lib().orc_lsqr_solve(C.byref(csr), C.byref(mi), C.c_double(eps), C.byref(nit), C.byref(ist),
                                _p(q), _p(b), _p(x))
    return x, {"ierr": ierr, "nit": nit.value, "istop": ist.value, "maxit": mi.value}


def laplace_dirichlet_system(mesh, H, G, diff=1.0, normmul=None):
    b = Body(mesh, normmul)
    ne, nn = b.mesh.nelem, b.mesh.nnodes
    sys = np.zeros((ne, ne)); rhs = np.zeros((ne, nn))
    lib().orc_laplace_dirichlet_system(C.byref(b.s), _p(np.ascontiguousarray(H)), _p(np.ascontiguousarray(G)),
                                       C.c_double(diff), _p(sys), _p(rhs))
    return sys, rhs


def stokes_dirichlet_system(mesh, T, G, row_offset=0, diff=1.0, normmul=None):
    b = Body(mesh, normmul)
    ne, nn = b.mesh.nelem, b.mesh.nnodes
    sys = np.zeros((3 * ne, 3 * ne)); rhs = np.zeros((3 * ne, 3 * nn))
    Tp = (_D * 6)(*[_p(np.ascontiguousarray(x)) for x in T])
    Gp = (_D * 6)(*[_p(np.ascontiguousarray(x)) for x in G])
    lib().orc_stokes_dirichlet_system(C.byref(b.s), Tp, Gp, row_offset, C.c_double(diff), _p(sys), _p(rhs))
    return sys, rhs


def resistance(mesh, exterior=True, mu=1.0, eps=1e-12, maxit=0, normmul=None, itri=5, problems=0):
    """Returns FT (6, 6): row k = (F, T) for problem k, and the LSQR iteration counts; problems = 1:
    the three strain fields (rows 0..2, rows 3..5 zero)."""
    b = Body(mesh, normmul, itri)
    FT = np.zeros(36)
    nits = np.zeros(6, dtype=np.int32)
    ierr = lib().orc_resistance_problems(C.byref(b.s), 1 if exterior else 0, C.c_double(mu), C.c_double(eps), maxit,
                                         problems, _p(FT), _pi(nits))
    return FT.reshape(6, 6), nits, ierr
