"""Python host interface over libandromeda_hip (ctypes), mirroring the reference's operator
boundaries: integral matrices (sdFormIntegralMatrices / sdFormIntegralMatricesStokes), the LSQR
routines of src/lib_lsqr.f, and the resistance drivers of src/lib_particle.f90.

The native library is required: every call goes to the HIP implementation, and a missing or
unloadable library raises immediately (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from .mesh import Mesh

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libandromeda_hip.so")

ROW_MAJOR, COL_MAJOR = 0, 1
FORM_EXTERIOR, FORM_BOUNDED = 0, 1
PREC_FP64, PREC_MIXED, PREC_LEARNED = 0, 1, 2
PROBLEMS_RIGID, PROBLEMS_STRAIN = 0, 1
MLP_NPARAM = 9288

_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int32)
_L = C.POINTER(C.c_int64)


class _Mesh(C.Structure):
    _fields_ = [("nnodes", C.c_int32), ("nelem", C.c_int32), ("nwalls", C.c_int32),
                ("xyz", _D), ("con", _I), ("bcid", _I), ("normmul", _D),
                ("index_base", C.c_int32), ("itri", C.c_int32), ("nvert", C.c_int32), ("iquad", C.c_int32)]


class _LsqrOpts(C.Structure):
    _fields_ = [("maxit", C.c_int32), ("eps", C.c_double), ("precon", C.c_int32)]


class _LsqrStats(C.Structure):
    _fields_ = [("nit", C.c_int32 * 16), ("istop", C.c_int32 * 16), ("seconds", C.c_double)]


class _Subdomain(C.Structure):
    _fields_ = [("nwalls", C.c_int32), ("walls", _I), ("normmul", _D), ("diff", C.c_double), ("node_wall", _I)]


class _LapReport(C.Structure):
    _fields_ = [("neq", C.c_int32), ("nx", C.c_int32), ("nb", C.c_int32), ("nit", C.c_int32), ("istop", C.c_int32),
                ("err", C.c_double * 2), ("t_assemble", C.c_double), ("t_solve", C.c_double)]


class _ResOpts(C.Structure):
    _fields_ = [("form", C.c_int32), ("precision", C.c_int32), ("mu", C.c_double), ("eps", C.c_double),
                ("maxit", C.c_int32), ("center", C.c_double * 3), ("restart", C.c_int32),
                ("problems", C.c_int32)]


class _ResReport(C.Structure):
    _fields_ = [("nit", C.c_int32 * 6), ("status", C.c_int32), ("err", C.c_double * 2),
                ("t_assemble", C.c_double), ("t_solve", C.c_double), ("quad_points", C.c_int64),
                ("pairs_learned", C.c_int64), ("pairs_fallback", C.c_int64)]


class AndromedaError(RuntimeError):
    pass


_lib = None


def load_library(path: str = LIB_PATH):
    """Loads libandromeda_hip.so; raises if it is absent (build with andromeda_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise AndromedaError(f"{path} is missing: build it with `python -m andromeda_amd.build`")
    lib = C.CDLL(path)
    lib.andr_version.restype = C.c_char_p
    lib.andr_last_error.restype = C.c_char_p
    lib.andr_create.argtypes = [C.POINTER(C.c_void_p), C.c_int32]
    lib.andr_destroy.argtypes = [C.c_void_p]
    lib.andr_last_error.argtypes = [C.c_void_p]
    lib.andr_assemble_laplace.argtypes = [C.c_void_p, C.POINTER(_Mesh), C.c_int32, _D, _D, _D]
    lib.andr_assemble_stokes.argtypes = [C.c_void_p, C.POINTER(_Mesh), C.c_int32, C.c_double,
                                         C.POINTER(_D), C.POINTER(_D), C.POINTER(_D), C.POINTER(_D), _D]
    lib.andr_stokes_domain_values.argtypes = [C.c_void_p, C.POINTER(_Mesh), C.c_int32, _D, _D, _D, C.c_double,
                                              _D, _D, _D]
    lib.andr_integrals_write_stokes.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, _D, C.POINTER(_D),
                                                C.POINTER(_D), C.POINTER(_D), C.POINTER(_D)]
    lib.andr_integrals_read_stokes.argtypes = lib.andr_integrals_write_stokes.argtypes
    lib.andr_integrals_write_laplace.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, _D, _D, _D]
    lib.andr_integrals_read_laplace.argtypes = lib.andr_integrals_write_laplace.argtypes
    lib.andr_bicgstab_dense.argtypes = [C.c_void_p, C.c_int64, _D, C.c_int32, _D, _D, C.c_double, C.c_int32,
                                        C.POINTER(_LsqrStats)]
    lib.andr_stokes_surface_pressure.argtypes = [C.c_void_p, C.POINTER(_Mesh), _D, _D, C.c_double, _D]
    lib.andr_lsqr_dense.argtypes = [C.c_void_p, C.c_int64, C.c_int64, _D, C.c_int32, _D, _D,
                                    C.POINTER(_LsqrOpts), C.POINTER(_LsqrStats)]
    lib.andr_resistance.argtypes = [C.c_void_p, C.POINTER(_Mesh), C.POINTER(_ResOpts), _D, C.POINTER(_ResReport)]
    lib.andr_resistance_batch.argtypes = [C.c_void_p, C.c_int32, C.POINTER(_Mesh), C.POINTER(_ResOpts), _D,
                                          C.POINTER(_ResReport)]
    lib.lsqr_precon_.argtypes = [_I, _I, _L, _D, _L, _I, _D]
    lib.lsqr_solve_.argtypes = [_I, _I, _L, _I, _D, _I, _I, _D, _L, _I, _D, _D, _D]
    lib.andr_quad_training_pairs.argtypes = [C.c_void_p, C.POINTER(_Mesh), C.c_double, C.c_double, C.c_int64,
                                             _F, _I, _L]
    lib.andr_mlp_train.argtypes = [C.c_void_p, C.c_int64, _F, _I, C.c_int32, C.c_int32, C.c_float, C.c_uint64,
                                   _F, _F]
    lib.andr_mlp_infer.argtypes = [C.c_void_p, C.c_int64, _F, _F, _F]
    lib.andr_mlp_loss_grad.argtypes = [C.c_void_p, C.c_int64, _F, _I, _F, _F, _F]
    lib.andr_set_quadrature_model.argtypes = [C.c_void_p, _F, C.c_float, C.c_double, _I]
    lib.andr_quad_far_calibration.argtypes = [C.c_void_p, C.POINTER(_Mesh), C.c_double, C.c_double, _I]
    lib.andr_profile.argtypes = [C.c_void_p, C.c_int32]
    lib.andr_profile_read.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_int32, _L, _D, _D, _D]
    lib.andr_bodies_upload.argtypes = [C.c_void_p, C.c_int32, C.POINTER(_Mesh), C.POINTER(_ResOpts),
                                       C.POINTER(C.c_void_p)]
    lib.andr_bodies_count.argtypes = [C.c_void_p]
    lib.andr_resistance_batch_resident.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(_ResOpts), _D,
                                                   C.POINTER(_ResReport)]
    lib.andr_bodies_free.argtypes = [C.c_void_p, C.c_void_p]
    lib.andr_mesh_validate.argtypes = [C.POINTER(_Mesh), C.c_char_p, C.c_int32]
    lib.andr_assemble_laplace_rows.argtypes = [C.c_void_p, C.POINTER(_Mesh), C.c_int32, C.c_int32, C.c_int32, _D, _D,
                                               _D]
    lib.andr_assemble_stokes_rows.argtypes = [C.c_void_p, C.POINTER(_Mesh), C.c_int32, C.c_int32, C.c_int32,
                                              C.c_double, C.POINTER(_D), C.POINTER(_D), C.POINTER(_D), C.POINTER(_D),
                                              _D]
    lib.andr_laplace_solve.argtypes = [C.c_void_p, C.POINTER(_Mesh), C.c_int32, C.POINTER(_Subdomain), C.c_int32, _I,
                                       C.POINTER(_LsqrOpts), _D, _D, _D, C.POINTER(_LapReport)]
    lib.andr_laplace_solve_mixed.argtypes = lib.andr_laplace_solve.argtypes
    lib.andr_stokes_solve.argtypes = [C.c_void_p, C.POINTER(_Mesh), C.POINTER(_Subdomain), _I, C.POINTER(_LsqrOpts),
                                      _D, _D, _D, C.POINTER(_LapReport)]
    lib.andr_subdomain_node_walls.argtypes = [C.POINTER(_Mesh), C.c_int32, C.POINTER(_Subdomain), _I, C.c_char_p,
                                              C.c_int32]
    lib.andr_bicgstab_dense_noalloc.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p,
                                                C.c_void_p, C.c_double, C.c_int32, C.POINTER(_LsqrStats)]
    lib.andr_autotune.argtypes = [C.c_void_p, C.c_int32]
    lib.andr_autotune_read.argtypes = [C.c_void_p, C.c_int32, _I, _I, _I, _D]
    lib.andr_autotune_read_krylov.argtypes = [C.c_void_p, C.c_int32, _I, _I, _D]
    _lib = lib
    return lib


EXPORTED = ["andr_create", "andr_destroy", "andr_last_error", "andr_version", "andr_assemble_laplace",
            "andr_assemble_stokes", "lsqr_precon_", "lsqr_solve_", "andr_lsqr_forget", "andr_lsqr_dense",
            "andr_resistance", "andr_resistance_batch", "andr_quad_training_pairs", "andr_mlp_train",
            "andr_mlp_infer", "andr_mlp_loss_grad", "andr_set_quadrature_model", "andr_profile",
            "andr_profile_read", "andr_quad_far_calibration", "andr_autotune", "andr_autotune_read", "andr_autotune_read_krylov",
            "andr_bodies_upload", "andr_bodies_count", "andr_resistance_batch_resident", "andr_bodies_free",
            "andr_stokes_domain_values", "andr_stokes_surface_pressure", "andr_integrals_write_stokes",
            "andr_integrals_read_stokes", "andr_integrals_write_laplace", "andr_integrals_read_laplace",
            "andr_bicgstab_dense", "andr_mesh_validate", "andr_bicgstab_dense_noalloc", "andr_assemble_laplace_rows",
            "andr_assemble_stokes_rows", "andr_laplace_solve", "andr_laplace_solve_mixed", "andr_subdomain_node_walls",
            "andr_stokes_solve"]
BC_FUNCTION, BC_FLUX, BC_CONTACT = 1, 2, 3
FAR_BINS = 16


def _p(a):
    return a.ctypes.data_as(_D)


def _pf(a):
    return a.ctypes.data_as(_F)


def _pi(a):
    return a.ctypes.data_as(_I)


class _MeshView:
    """Holds the arrays referenced by the C struct."""

    def __init__(self, mesh: Mesh, normmul=None, itri: int = 5, iquad: int = 4):
        m = mesh.contiguous()
        self.m = m
        self.normmul = np.ascontiguousarray(normmul if normmul is not None else np.ones(max(m.nwalls, 1)),
                                            dtype=np.float64)
        self.s = _Mesh(m.nnodes, m.nelem, max(m.nwalls, 1), _p(m.xyz), _pi(m.con), _pi(m.bcid),
                       _p(self.normmul), 0, itri, int(m.con.shape[1]), iquad)


def validate_mesh(mesh: Mesh, nwalls=None, itri: int = 5, iquad: int = 4, con=None, bcid=None):
    """andr_mesh_validate on the host (no device): (status, message). con and bcid override the
    mesh's arrays (to check raw inputs), nwalls its wall count."""
    mv = _MeshView(mesh, None, itri, iquad)
    keep = []
    if con is not None:
        c = np.ascontiguousarray(con, dtype=np.int32)
        keep.append(c)
        mv.s.con = _pi(c)
    if bcid is not None:
        b = np.ascontiguousarray(bcid, dtype=np.int32)
        keep.append(b)
        mv.s.bcid = _pi(b)
    if nwalls is not None:
        mv.s.nwalls = nwalls
    msg = C.create_string_buffer(256)
    rc = load_library().andr_mesh_validate(C.byref(mv.s), msg, 256)
    return rc, msg.value.decode()


@dataclass
class Subdomain:
    """Walls (ids 1..nwalls, in the order of the BIC file's SUBDOMAINS line), their NORMMUL factors,
    the diffusivity, and optional per-node wall overrides (EDGE / CORNER rules: nnodes entries, 0 =
    the default rule)."""
    walls: list
    normmul: list | None = None
    diff: float = 1.0
    node_wall: np.ndarray | None = None


class _SdView:
    def __init__(self, sds, nnodes):
        self.keep = []
        arr = []
        for sd in sds:
            w = np.ascontiguousarray(sd.walls, dtype=np.int32)
            nm = np.ascontiguousarray(sd.normmul if sd.normmul is not None else np.ones(w.size), dtype=np.float64)
            nw = None
            if sd.node_wall is not None:
                nw = np.ascontiguousarray(sd.node_wall, dtype=np.int32)
                if nw.size != nnodes:
                    raise ValueError("node_wall needs one entry per node")
            self.keep += [w, nm, nw]
            arr.append(_Subdomain(w.size, _pi(w), _p(nm), float(sd.diff), _pi(nw) if nw is not None else None))
        self.arr = (_Subdomain * len(arr))(*arr)
        self.n = len(arr)


def subdomain_node_walls(mesh: Mesh, subdomains, itri: int = 5):
    """(nsd, nnodes) wall id each node of each subdomain takes its condition from (0: not in the
    subdomain), as andr_laplace_solve uses it (host only)."""
    mv = _MeshView(mesh, None, itri)
    sv = _SdView(subdomains, mv.m.nnodes)
    out = np.zeros((sv.n, mv.m.nnodes), dtype=np.int32)
    msg = C.create_string_buffer(256)
    rc = load_library().andr_subdomain_node_walls(C.byref(mv.s), sv.n, sv.arr, _pi(out), msg, 256)
    if rc != 0:
        raise AndromedaError(f"andr_subdomain_node_walls failed ({rc}): {msg.value.decode()}")
    return out


@dataclass
class LaplaceResult:
    u: np.ndarray          # (nnodes,) function values
    q: np.ndarray          # (nelem,) fluxes
    wall_flux: np.ndarray  # (nwalls,) sum of A_e q_e per wall
    neq: int
    nx: int
    nb: int
    nit: int
    istop: int
    err: tuple
    t_assemble: float
    t_solve: float


@dataclass
class ResistanceResult:
    FT: np.ndarray           # (6, 6): row k = (force, torque) of rigid-body problem k
    nit: np.ndarray          # (6,) iterations
    status: int
    err: tuple
    t_assemble: float
    t_solve: float
    quad_points: int
    pairs_learned: int
    pairs_fallback: int


def _rows(mats):
    return [np.ascontiguousarray(a, dtype=np.float64) for a in mats]


def integrals_write_stokes(path: str, T, G, prT, prG, err=(0.0, 0.0)):
    """The reference's integrals file (Stokes) from row-major nsp x nnodes / nsp x nqnodes matrices,
    T and G in the order xx, xy, xz, yy, yz, zz as Context.assemble_stokes returns them."""
    T, G, prT, prG = _rows(T), _rows(G), _rows(prT), _rows(prG)
    nsp, nn = T[0].shape
    nq = G[0].shape[1]
    e = np.asarray(err, dtype=np.float64)
    rc = load_library().andr_integrals_write_stokes(
        path.encode(), nsp, nn, nq, _p(e), (_D * 6)(*[_p(a) for a in T]), (_D * 6)(*[_p(a) for a in G]),
        (_D * 3)(*[_p(a) for a in prT]), (_D * 3)(*[_p(a) for a in prG]))
    if rc != 0:
        raise AndromedaError(f"andr_integrals_write_stokes failed ({rc}): {path}")


def integrals_read_stokes(path: str, nsp: int, nnodes: int, nqnodes: int):
    """(T, G, prT, prG, err) from the reference's integrals file (Stokes); sizes must match."""
    T = [np.zeros((nsp, nnodes)) for _ in range(6)]
    G = [np.zeros((nsp, nqnodes)) for _ in range(6)]
    prT = [np.zeros((nsp, nnodes)) for _ in range(3)]
    prG = [np.zeros((nsp, nqnodes)) for _ in range(3)]
    e = np.zeros(2)
    rc = load_library().andr_integrals_read_stokes(
        path.encode(), nsp, nnodes, nqnodes, _p(e), (_D * 6)(*[_p(a) for a in T]), (_D * 6)(*[_p(a) for a in G]),
        (_D * 3)(*[_p(a) for a in prT]), (_D * 3)(*[_p(a) for a in prG]))
    if rc != 0:
        raise AndromedaError(f"andr_integrals_read_stokes failed ({rc}): {path}")
    return T, G, prT, prG, e


def integrals_write_laplace(path: str, H, G, err=(0.0, 0.0)):
    H, G = _rows([H, G])
    e = np.asarray(err, dtype=np.float64)
    rc = load_library().andr_integrals_write_laplace(path.encode(), H.shape[0], H.shape[1], G.shape[1], _p(e), _p(H),
                                                     _p(G))
    if rc != 0:
        raise AndromedaError(f"andr_integrals_write_laplace failed ({rc}): {path}")


def integrals_read_laplace(path: str, nsp: int, nnodes: int, nqnodes: int):
    H, G, e = np.zeros((nsp, nnodes)), np.zeros((nsp, nqnodes)), np.zeros(2)
    rc = load_library().andr_integrals_read_laplace(path.encode(), nsp, nnodes, nqnodes, _p(e), _p(H), _p(G))
    if rc != 0:
        raise AndromedaError(f"andr_integrals_read_laplace failed ({rc}): {path}")
    return H, G, e


def _results(FT, reps):
    return [ResistanceResult(FT[i].reshape(6, 6).copy(), np.array(list(r.nit)), r.status, (r.err[0], r.err[1]),
                             r.t_assemble, r.t_solve, r.quad_points, r.pairs_learned, r.pairs_fallback)
            for i, r in enumerate(reps)]


class Bodies:
    """Device-resident bodies of a context (andr_bodies_upload / andr_bodies_free)."""

    def __init__(self, ctx, h, form, mu):
        self.ctx, self.h, self.form, self.mu = ctx, h, form, mu

    def __len__(self):
        return self.ctx.lib.andr_bodies_count(self.h) if self.h else 0

    def free(self):
        if self.h and getattr(self.ctx, "h", None):
            self.ctx.lib.andr_bodies_free(self.ctx.h, self.h)
        self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One GPU + one HIP stream (andr_create / andr_destroy)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.andr_create(C.byref(h), device)
        if rc != 0:
            raise AndromedaError(f"andr_create failed ({rc}): no usable HIP device {device}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.andr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc: int, what: str):
        if rc < 0:
            msg = self.lib.andr_last_error(self.h)
            raise AndromedaError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
        return rc

    # ---- seam 1 ---------------------------------------------------------------------------------
    def assemble_laplace(self, mesh: Mesh, normmul=None, itri: int = 5, layout: int = ROW_MAJOR, iquad: int = 4):
        mv = _MeshView(mesh, normmul, itri, iquad)
        nn, ne = mv.m.nnodes, mv.m.nelem
        nsp = nn + ne
        shapeH = (nsp, nn) if layout == ROW_MAJOR else (nn, nsp)
        shapeG = (nsp, ne) if layout == ROW_MAJOR else (ne, nsp)
        H = np.zeros(shapeH)
        G = np.zeros(shapeG)
        err = np.zeros(2)
        self._check(self.lib.andr_assemble_laplace(self.h, C.byref(mv.s), layout, _p(H), _p(G), _p(err)),
                    "andr_assemble_laplace")
        if layout == COL_MAJOR:
            H, G = H.T, G.T
        return H, G, err

    def assemble_laplace_rows(self, mesh: Mesh, row_begin: int, row_end: int, normmul=None, itri: int = 5,
                              layout: int = ROW_MAJOR, iquad: int = 4):
        """Rows row_begin .. row_end - 1 of H and G (andr_assemble_laplace_rows)."""
        mv = _MeshView(mesh, normmul, itri, iquad)
        nn, ne, nr = mv.m.nnodes, mv.m.nelem, row_end - row_begin
        H = np.zeros((nr, nn) if layout == ROW_MAJOR else (nn, nr))
        G = np.zeros((nr, ne) if layout == ROW_MAJOR else (ne, nr))
        err = np.zeros(2)
        self._check(self.lib.andr_assemble_laplace_rows(self.h, C.byref(mv.s), row_begin, row_end, layout, _p(H),
                                                        _p(G), _p(err)), "andr_assemble_laplace_rows")
        if layout == COL_MAJOR:
            H, G = H.T, G.T
        return H, G, err

    def assemble_stokes_rows(self, mesh: Mesh, row_begin: int, row_end: int, normmul=None, itri: int = 5,
                             extra: float = 0.0, iquad: int = 4):
        """Rows row_begin .. row_end - 1 of the 18 Stokes matrices (andr_assemble_stokes_rows)."""
        mv = _MeshView(mesh, normmul, itri, iquad)
        nn, ne, nr = mv.m.nnodes, mv.m.nelem, row_end - row_begin
        T = [np.zeros((nr, nn)) for _ in range(6)]
        G = [np.zeros((nr, ne)) for _ in range(6)]
        prT = [np.zeros((nr, nn)) for _ in range(3)]
        prG = [np.zeros((nr, ne)) for _ in range(3)]
        err = np.zeros(2)
        self._check(self.lib.andr_assemble_stokes_rows(
            self.h, C.byref(mv.s), row_begin, row_end, ROW_MAJOR, extra, (_D * 6)(*[_p(x) for x in T]),
            (_D * 6)(*[_p(x) for x in G]), (_D * 3)(*[_p(x) for x in prT]), (_D * 3)(*[_p(x) for x in prG]),
            _p(err)), "andr_assemble_stokes_rows")
        return T, G, prT, prG, err

    def laplace_solve(self, mesh: Mesh, subdomains, wall_bc, u, q, eps=1e-15, maxit=500, itri: int = 5,
                      iquad: int = 4, precision: int = PREC_FP64):
        """The Laplace boundary-value problem (andr_laplace_solve) for one or more equations:
        wall_bc (neqn, nwalls) of BC_* kinds, u (neqn, nnodes) and q (neqn, nelem) the known and
        initial values; eps and maxit scalars or per equation. precision PREC_MIXED: the
        mixed-precision form (andr_laplace_solve_mixed). Returns a LaplaceResult per equation."""
        if precision not in (PREC_FP64, PREC_MIXED):
            raise ValueError("laplace_solve takes PREC_FP64 or PREC_MIXED")
        mv = _MeshView(mesh, None, itri, iquad)
        nn, ne, nw = mv.m.nnodes, mv.m.nelem, mv.s.nwalls
        bc = np.ascontiguousarray(np.atleast_2d(wall_bc), dtype=np.int32)
        neqn = bc.shape[0]
        if bc.shape[1] != nw:
            raise ValueError(f"wall_bc needs {nw} entries per equation")
        U = np.ascontiguousarray(np.array(u, dtype=np.float64).reshape(neqn, nn))
        Q = np.ascontiguousarray(np.array(q, dtype=np.float64).reshape(neqn, ne))
        eps_a = np.broadcast_to(np.asarray(eps, dtype=np.float64), (neqn,))
        mit_a = np.broadcast_to(np.asarray(maxit, dtype=np.int64), (neqn,))
        opts = (_LsqrOpts * neqn)(*[_LsqrOpts(int(mit_a[k]), float(eps_a[k]), 1) for k in range(neqn)])
        flux = np.zeros((neqn, nw))
        reps = (_LapReport * neqn)()
        sv = _SdView(subdomains, nn)
        fn = self.lib.andr_laplace_solve_mixed if precision == PREC_MIXED else self.lib.andr_laplace_solve
        self._check(fn(self.h, C.byref(mv.s), sv.n, sv.arr, neqn, _pi(bc), opts, _p(U), _p(Q), _p(flux), reps),
                    "andr_laplace_solve")
        return [LaplaceResult(U[k].copy(), Q[k].copy(), flux[k].copy(), r.neq, r.nx, r.nb, r.nit, r.istop,
                              (r.err[0], r.err[1]), r.t_assemble, r.t_solve) for k, r in enumerate(reps)]

    def stokes_solve(self, mesh: Mesh, subdomain: "Subdomain", wall_bc, u, q, eps=1e-15, maxit=500, itri: int = 5):
        """The general Stokes problem (andr_stokes_solve): wall_bc (3, nwalls) of BC_* kinds per velocity
        component, u (3, nnodes) and q (3, nelem) the known and initial values. Returns a
        LaplaceResult with u (3, nnodes), q (3, nelem) and wall_flux (3, nwalls) the wall forces."""
        mv = _MeshView(mesh, None, itri)
        nn, ne, nw = mv.m.nnodes, mv.m.nelem, mv.s.nwalls
        bc = np.ascontiguousarray(wall_bc, dtype=np.int32).reshape(3, nw)
        U = np.ascontiguousarray(np.array(u, dtype=np.float64).reshape(3, nn))
        Q = np.ascontiguousarray(np.array(q, dtype=np.float64).reshape(3, ne))
        opts = _LsqrOpts(int(maxit), float(eps), 1)
        force = np.zeros((3, nw))
        rep = _LapReport()
        sv = _SdView([subdomain], nn)
        self._check(self.lib.andr_stokes_solve(self.h, C.byref(mv.s), sv.arr, _pi(bc), C.byref(opts), _p(U), _p(Q),
                                               _p(force), C.byref(rep)), "andr_stokes_solve")
        return LaplaceResult(U, Q, force, rep.neq, rep.nx, rep.nb, rep.nit, rep.istop, (rep.err[0], rep.err[1]),
                             rep.t_assemble, rep.t_solve)

    def assemble_stokes(self, mesh: Mesh, normmul=None, itri: int = 5, extra: float = 0.0, pressure: bool = True,
                        iquad: int = 4):
        mv = _MeshView(mesh, normmul, itri, iquad)
        nn, ne = mv.m.nnodes, mv.m.nelem
        nsp = nn + ne
        T = [np.zeros((nsp, nn)) for _ in range(6)]
        G = [np.zeros((nsp, ne)) for _ in range(6)]
        prT = [np.zeros((nsp, nn)) for _ in range(3)] if pressure else None
        prG = [np.zeros((nsp, ne)) for _ in range(3)] if pressure else None
        err = np.zeros(2)
        Tp = (_D * 6)(*[_p(x) for x in T])
        Gp = (_D * 6)(*[_p(x) for x in G])
        pTp = (_D * 3)(*[_p(x) for x in prT]) if pressure else (_D * 3)()
        pGp = (_D * 3)(*[_p(x) for x in prG]) if pressure else (_D * 3)()
        self._check(self.lib.andr_assemble_stokes(self.h, C.byref(mv.s), ROW_MAJOR, extra, Tp, Gp, pTp, pGp, _p(err)),
                    "andr_assemble_stokes")
        return T, G, prT, prG, err

    def stokes_domain_values(self, mesh: Mesh, pts, u, q, diff: float = 1.0, normmul=None, itri: int = 5):
        """Velocity (npts, 3), pressure (npts,) and closure error (npts,) at points of the flow domain
        from nodal velocities u (nnodes, 3) and element tractions q (nelem, 3)."""
        mv = _MeshView(mesh, normmul, itri)
        P = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        U = np.ascontiguousarray(u, dtype=np.float64).reshape(mv.m.nnodes, 3)
        Q = np.ascontiguousarray(q, dtype=np.float64).reshape(mv.m.nelem, 3)
        vel = np.zeros((P.shape[0], 3))
        p = np.zeros(P.shape[0])
        err = np.zeros(P.shape[0])
        self._check(self.lib.andr_stokes_domain_values(self.h, C.byref(mv.s), P.shape[0], _p(P), _p(U), _p(Q), diff,
                                                       _p(vel), _p(p), _p(err)), "andr_stokes_domain_values")
        return vel, p, err

    def stokes_surface_pressure(self, mesh: Mesh, u, q, diff: float = 1.0, normmul=None, itri: int = 5):
        """Pressure at the element centres (nelem,) from nodal velocities u and element tractions q."""
        mv = _MeshView(mesh, normmul, itri)
        U = np.ascontiguousarray(u, dtype=np.float64).reshape(mv.m.nnodes, 3)
        Q = np.ascontiguousarray(q, dtype=np.float64).reshape(mv.m.nelem, 3)
        p = np.zeros(mv.m.nelem)
        self._check(self.lib.andr_stokes_surface_pressure(self.h, C.byref(mv.s), _p(U), _p(Q), diff, _p(p)),
                    "andr_stokes_surface_pressure")
        return p

    # ---- seam 2 ---------------------------------------------------------------------------------
    def lsqr_dense(self, A, B, X0=None, maxit: int = 0, eps: float = 1e-15, precon: bool = True):
        A = np.ascontiguousarray(A, dtype=np.float64)
        m, n = A.shape
        B = np.asarray(B, dtype=np.float64)
        if B.ndim == 1:
            B = B[:, None]
        nrhs = B.shape[1]
        Bc = np.ascontiguousarray(B.T)  # column k contiguous
        Xc = np.zeros((nrhs, n)) if X0 is None else np.ascontiguousarray(np.asarray(X0, dtype=np.float64).reshape(n, nrhs).T)
        opts = _LsqrOpts(maxit, eps, 1 if precon else 0)
        st = _LsqrStats()
        self._check(self.lib.andr_lsqr_dense(self.h, m, n, _p(A), nrhs, _p(Bc), _p(Xc), C.byref(opts), C.byref(st)),
                    "andr_lsqr_dense")
        return Xc.T.copy(), {"nit": list(st.nit)[:nrhs], "istop": list(st.istop)[:nrhs], "seconds": st.seconds}

    def bicgstab_dense(self, A, B, X0=None, tol: float = 1e-8, maxit: int = 1000):
        """Jacobi-preconditioned BiCGSTAB on a dense square system (the LIS path's default solver)."""
        A = np.ascontiguousarray(A, dtype=np.float64)
        n = A.shape[0]
        B = np.asarray(B, dtype=np.float64)
        if B.ndim == 1:
            B = B[:, None]
        nrhs = B.shape[1]
        Bc = np.ascontiguousarray(B.T)
        Xc = np.zeros((nrhs, n)) if X0 is None else np.ascontiguousarray(np.asarray(X0, dtype=np.float64).reshape(n, nrhs).T)
        st = _LsqrStats()
        self._check(self.lib.andr_bicgstab_dense(self.h, n, _p(A), nrhs, _p(Bc), _p(Xc), tol, maxit, C.byref(st)),
                    "andr_bicgstab_dense")
        return Xc.T.copy(), {"nit": list(st.nit)[:nrhs], "status": list(st.istop)[:nrhs], "seconds": st.seconds}

    def bicgstab_dense_noalloc(self, n: int, A_dev: int, lda: int, nrhs: int, B_dev: int, X_dev: int,
                               tol: float = 1e-8, maxit: int = 1000):
        """andr_bicgstab_dense_noalloc on caller-owned device arrays (raw device addresses): A n x n
        row-major with leading dimension lda, B and X n x nrhs with column k contiguous."""
        st = _LsqrStats()
        self._check(self.lib.andr_bicgstab_dense_noalloc(self.h, n, C.c_void_p(A_dev), lda, nrhs, C.c_void_p(B_dev),
                                                         C.c_void_p(X_dev), tol, maxit, C.byref(st)),
                    "andr_bicgstab_dense_noalloc")
        return {"nit": list(st.nit)[:nrhs], "status": list(st.istop)[:nrhs], "seconds": st.seconds}

    # ---- seam 3 ---------------------------------------------------------------------------------
    def resistance_batch(self, meshes, form: int = FORM_EXTERIOR, precision: int = PREC_FP64, mu: float = 1.0,
                         eps: float = 0.0, maxit: int = 0, center=(0.0, 0.0, 0.0), itri: int = 5, normmul=None,
                         restart: int = 0, problems: int = PROBLEMS_RIGID):
        views = [_MeshView(m, normmul, itri) for m in meshes]
        arr = (_Mesh * len(views))(*[v.s for v in views])
        opts = _ResOpts(form, precision, mu, eps, maxit, (C.c_double * 3)(*center), restart, problems)
        FT = np.zeros((len(views), 36))
        reps = (_ResReport * len(views))()
        self._check(self.lib.andr_resistance_batch(self.h, len(views), arr, C.byref(opts), _p(FT), reps),
                    "andr_resistance_batch")
        return _results(FT, reps)

    def upload_bodies(self, meshes, form: int = FORM_EXTERIOR, mu: float = 1.0, itri: int = 5,
                      normmul=None) -> "Bodies":
        """Device-resident bodies for repeated resistance batches (andr_bodies_upload)."""
        views = [_MeshView(m, normmul, itri) for m in meshes]
        arr = (_Mesh * len(views))(*[v.s for v in views])
        opts = _ResOpts(form, PREC_FP64, mu, 0.0, 0, (C.c_double * 3)(0.0, 0.0, 0.0), 0, PROBLEMS_RIGID)
        h = C.c_void_p()
        self._check(self.lib.andr_bodies_upload(self.h, len(views), arr, C.byref(opts), C.byref(h)),
                    "andr_bodies_upload")
        return Bodies(self, h, form, mu)

    def resistance_batch_resident(self, bodies: "Bodies", precision: int = PREC_FP64, eps: float = 0.0,
                                  maxit: int = 0, center=(0.0, 0.0, 0.0), restart: int = 0,
                                  problems: int = PROBLEMS_RIGID):
        """andr_resistance_batch on device-resident bodies (form and mu of the upload)."""
        n = self.lib.andr_bodies_count(bodies.h)
        opts = _ResOpts(bodies.form, precision, bodies.mu, eps, maxit, (C.c_double * 3)(*center), restart, problems)
        FT = np.zeros((n, 36))
        reps = (_ResReport * n)()
        self._check(self.lib.andr_resistance_batch_resident(self.h, bodies.h, C.byref(opts), _p(FT), reps),
                    "andr_resistance_batch_resident")
        return _results(FT, reps)

    def resistance(self, mesh, **kw) -> ResistanceResult:
        return self.resistance_batch([mesh], **kw)[0]

    # ---- measurement ----------------------------------------------------------------------------
    def profile(self, enable: bool = True):
        self._check(self.lib.andr_profile(self.h, 1 if enable else 0), "andr_profile")

    def profile_read(self) -> dict:
        out = {}
        i = 0
        name = C.create_string_buffer(128)
        n, ms, by, fl = C.c_int64(), C.c_double(), C.c_double(), C.c_double()
        while self.lib.andr_profile_read(self.h, i, name, 128, C.byref(n), C.byref(ms), C.byref(by), C.byref(fl)) == 1:
            out[name.value.decode()] = {"launches": n.value, "ms": ms.value, "bytes": by.value, "flops": fl.value}
            i += 1
        return out

    def autotune(self, mode=True):
        """True/1: tune; False/0: 64 rows per wave; 32, 64 or 128: that many rows per wave."""
        self._check(self.lib.andr_autotune(self.h, int(mode)), "andr_autotune")

    def autotune_table(self) -> list:
        """Autotuner choices: rows per wave of the plain solve's FP64 product per size bucket."""
        out = []
        i = 0
        lr, lc, rpw, ms = C.c_int32(), C.c_int32(), C.c_int32(), C.c_double()
        while self.lib.andr_autotune_read(self.h, i, C.byref(lr), C.byref(lc), C.byref(rpw), C.byref(ms)) == 1:
            out.append({"log2_rows": lr.value, "log2_cols": lc.value, "rows_per_wave": rpw.value, "ms": ms.value})
            i += 1
        return out

    def autotune_table_krylov(self) -> list:
        """Autotuner choices of the fast forms' Krylov product: the waves its column segments aim at,
        per bucket of live 32-row waves."""
        out = []
        i = 0
        lw, tw, ms = C.c_int32(), C.c_int32(), C.c_double()
        while self.lib.andr_autotune_read_krylov(self.h, i, C.byref(lw), C.byref(tw), C.byref(ms)) == 1:
            out.append({"log2_row_waves": lw.value, "target_waves": tw.value, "ms": ms.value})
            i += 1
        return out

    # ---- learned quadrature ---------------------------------------------------------------------
    def quad_training_pairs(self, mesh, near_ratio: float = 4.0, rel_tol: float = 1e-8, max_pairs: int = 200000,
                            itri: int = 5):
        mv = _MeshView(mesh, None, itri)
        feats = np.zeros((max_pairs, 6), dtype=np.float32)
        labels = np.zeros(max_pairs, dtype=np.int32)
        npairs = C.c_int64(0)
        self._check(self.lib.andr_quad_training_pairs(self.h, C.byref(mv.s), near_ratio, rel_tol, max_pairs,
                                                      _pf(feats), _pi(labels), C.byref(npairs)),
                    "andr_quad_training_pairs")
        k = npairs.value
        return feats[:k].copy(), labels[:k].copy()

    def mlp_train(self, feats, labels, params, epochs: int = 10, batch: int = 1024, lr: float = 1e-3, seed: int = 0):
        feats = np.ascontiguousarray(feats, dtype=np.float32)
        labels = np.ascontiguousarray(labels, dtype=np.int32)
        p = np.ascontiguousarray(params, dtype=np.float32).copy()
        loss = np.zeros(epochs, dtype=np.float32)
        self._check(self.lib.andr_mlp_train(self.h, feats.shape[0], _pf(feats), _pi(labels), epochs, batch, lr, seed,
                                            _pf(p), _pf(loss)), "andr_mlp_train")
        return p, loss

    def mlp_infer(self, feats, params):
        feats = np.ascontiguousarray(feats, dtype=np.float32)
        params = np.ascontiguousarray(params, dtype=np.float32)
        probs = np.zeros((feats.shape[0], 8), dtype=np.float32)
        self._check(self.lib.andr_mlp_infer(self.h, feats.shape[0], _pf(feats), _pf(params), _pf(probs)),
                    "andr_mlp_infer")
        return probs

    def mlp_loss_grad(self, feats, labels, params):
        feats = np.ascontiguousarray(feats, dtype=np.float32)
        labels = np.ascontiguousarray(labels, dtype=np.int32)
        params = np.ascontiguousarray(params, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        grad = np.zeros(MLP_NPARAM, dtype=np.float32)
        self._check(self.lib.andr_mlp_loss_grad(self.h, feats.shape[0], _pf(feats), _pi(labels), _pf(params),
                                                _pf(loss), _pf(grad)), "andr_mlp_loss_grad")
        return float(loss[0]), grad

    def set_quadrature_model(self, params, confidence: float, near_ratio: float, far_table):
        params = np.ascontiguousarray(params, dtype=np.float32)
        far = np.ascontiguousarray(far_table, dtype=np.int32)
        if far.size != FAR_BINS:
            raise ValueError(f"far_table needs {FAR_BINS} entries")
        self._check(self.lib.andr_set_quadrature_model(self.h, _pf(params), confidence, near_ratio, _pi(far)),
                    "andr_set_quadrature_model")

    def quad_far_calibration(self, mesh, near_ratio: float, rel_tol: float, itri: int = 5):
        mv = _MeshView(mesh, None, itri)
        far = np.zeros(FAR_BINS, dtype=np.int32)
        self._check(self.lib.andr_quad_far_calibration(self.h, C.byref(mv.s), near_ratio, rel_tol, _pi(far)),
                    "andr_quad_far_calibration")
        return far


# ---- Fortran-style LSQR (lsqr_precon_ / lsqr_solve_), 1-based CSR --------------------------------
def lsqr_precon(m: int, n: int, ia, ja, a):
    """Calls lsqr_precon_ exactly as the Fortran host would: a is scaled in place, q returned."""
    lib = load_library()
    q = np.zeros(n)
    mm, nn_, nnz = C.c_int32(m), C.c_int32(n), C.c_int64(a.size)
    lib.lsqr_precon_(C.byref(mm), C.byref(nn_), C.byref(nnz), _p(q), ia.ctypes.data_as(_L), _pi(ja), _p(a))
    return q


def lsqr_forget():
    """Releases the seam's cached device copy of the last matrix (andr_lsqr_forget)."""
    load_library().andr_lsqr_forget()


def lsqr_solve(m: int, n: int, ia, ja, a, q, b, x, maxit: int = 0, eps: float = 1e-15):
    lib = load_library()
    mm, nn_, nnz = C.c_int32(m), C.c_int32(n), C.c_int64(a.size)
    mi, ep, nit, ierr = C.c_int32(maxit), C.c_double(eps), C.c_int32(0), C.c_int32(0)
    lib.lsqr_solve_(C.byref(mm), C.byref(nn_), C.byref(nnz), C.byref(mi), C.byref(ep), C.byref(nit), C.byref(ierr),
                    _p(q), ia.ctypes.data_as(_L), _pi(ja), _p(a), _p(b), _p(x))
    return {"nit": nit.value, "ierr": ierr.value, "maxit": mi.value}
