"""Surface meshes for the BEM hot path: generators for the benchmark bodies and readers for the
mesh formats the reference accepts.

Element connectivity is 0-based int32 (nelem, 3) for triangle meshes, or (nelem, 4) when the mesh
has 4-node quadrilaterals (a triangle then has -1 as its fourth entry); node coordinates are
float64 (nnodes, 3); `bcid` holds the 1-based wall id of every element, as `element(ie)%bcid` in
the reference (src/mod_mesh.f90:16-80).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np


@dataclasses.dataclass
class Mesh:
    xyz: np.ndarray          # (nnodes, 3) float64
    con: np.ndarray          # (nelem, 3) or (nelem, 4) int32, 0-based (-1: no fourth node)
    bcid: np.ndarray         # (nelem,) int32, walls numbered from 1
    name: str = ""
    walls: list | None = None  # wall names by wall id (walls[w - 1]), from the mesh file's physical names

    @property
    def nnodes(self) -> int:
        return int(self.xyz.shape[0])

    @property
    def nelem(self) -> int:
        return int(self.con.shape[0])

    @property
    def nwalls(self) -> int:
        return int(self.bcid.max()) if self.bcid.size else 0

    def contiguous(self) -> "Mesh":
        return Mesh(np.ascontiguousarray(self.xyz, dtype=np.float64),
                    np.ascontiguousarray(self.con, dtype=np.int32),
                    np.ascontiguousarray(self.bcid, dtype=np.int32), self.name, self.walls)


# --------------------------------------------------------------------------------------------
# geodesic sphere
# --------------------------------------------------------------------------------------------
def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v, f


_SPHERES: dict = {}


def geodesic_sphere(nu: int, radius: float = 1.0) -> Mesh:
    """Class-I geodesic sphere: every icosahedron face split into nu*nu triangles, vertices
    projected radially onto the sphere. F = 20 nu^2, V = 10 nu^2 + 2; outward normals
    (counter-clockwise seen from outside). nu = 2**s gives the "icosphere subdivision s" sizes
    (s = 3: 1280 triangles, s = 5: 20480)."""
    if nu not in _SPHERES:
        _SPHERES[nu] = _geodesic_unit(nu)
    s = _SPHERES[nu]
    return Mesh(s.xyz * radius, s.con.copy(), s.bcid.copy(), s.name)


def _geodesic_unit(nu: int) -> Mesh:
    v0, f0 = _icosahedron()
    key_to_id: dict = {}
    pts: list = []

    def vid(face, i, j):
        # point with barycentric (i, j, nu-i-j)/nu on face; shared edges/vertices deduplicated
        a, b, c = f0[face]
        k = nu - i - j
        # canonical key: sorted (vertex, weight) pairs of nonzero weights
        key = tuple(sorted(((int(a), i), (int(b), j), (int(c), k))))
        key = tuple((vv, ww) for vv, ww in key if ww > 0)
        if key in key_to_id:
            return key_to_id[key]
        p = (i * v0[a] + j * v0[b] + k * v0[c]) / nu
        key_to_id[key] = len(pts)
        pts.append(p)
        return key_to_id[key]

    tris = []
    for face in range(20):
        for i in range(nu):
            for j in range(nu - i):
                # upward triangle: (i,j), (i-? ) in barycentric lattice
                p0 = vid(face, i + 1, j)
                p1 = vid(face, i, j + 1)
                p2 = vid(face, i, j)
                tris.append((p0, p1, p2))
                if i + j < nu - 1:
                    q0 = vid(face, i + 1, j)
                    q1 = vid(face, i + 1, j + 1)
                    q2 = vid(face, i, j + 1)
                    tris.append((q0, q1, q2))
    xyz = np.array(pts, dtype=np.float64)
    xyz /= np.linalg.norm(xyz, axis=1, keepdims=True)
    con = np.array(tris, dtype=np.int32)
    # orient outward
    x1, x2, x3 = xyz[con[:, 0]], xyz[con[:, 1]], xyz[con[:, 2]]
    n = np.cross(x2 - x1, x3 - x1)
    flip = np.einsum("ij,ij->i", n, x1 + x2 + x3) < 0
    con[flip] = con[flip][:, [0, 2, 1]]
    return Mesh(xyz, con, np.ones(con.shape[0], dtype=np.int32), f"sphere_nu{nu}")


def icosphere(subdivision: int, radius: float = 1.0) -> Mesh:
    return geodesic_sphere(2 ** subdivision, radius)


def nu_for_elements(nelem: float) -> int:
    """Nearest geodesic frequency nu (5..20) to a target element count, F = 20 nu^2."""
    nu = int(round(math.sqrt(max(nelem, 1.0) / 20.0)))
    return max(5, min(20, nu))


# --------------------------------------------------------------------------------------------
# superellipsoids and fibres
# --------------------------------------------------------------------------------------------
def superellipsoid(axes, p: float, nu: int, stretch: bool = True) -> Mesh:
    """|x/a|^p + |y/b|^p + |z/c|^p = 1 meshed from a geodesic sphere. With `stretch` the sphere
    is first scaled affinely to the ellipsoid (a, b, c), then every node is moved along its ray
    from the origin onto the surface; the ray intersection is exact:
    t = (sum |d_i / a_i|^p)^(-1/p) for direction d."""
    a = np.asarray(axes, dtype=np.float64)
    s = geodesic_sphere(nu)
    d = s.xyz * a[None, :] if stretch else s.xyz.copy()
    t = np.sum(np.abs(d / a[None, :]) ** p, axis=1) ** (-1.0 / p)
    xyz = d * t[:, None]
    return Mesh(xyz, s.con.copy(), s.bcid.copy(), f"superellipsoid_p{p:.3f}_nu{nu}")


def random_superellipsoid(rng: np.random.Generator, nu: int) -> tuple[Mesh, dict]:
    axes = rng.uniform(0.5, 2.0, size=3)
    p = float(rng.uniform(2.0, 8.0))
    m = superellipsoid(axes, p, nu)
    return m, {"axes": axes.tolist(), "p": p, "nu": nu, "kind": "superellipsoid"}


def fibre(half_length: float, radius: float, p: float, nelem_target: int, axis: int = 0) -> Mesh:
    """Slender superellipsoid |s/L|^p + |y/r|^p + |z/r|^p = 1 (L = half_length, r = radius) with
    near-isotropic triangles: latitude rings spaced uniformly in arc length along the meridian,
    n_phi nodes per ring chosen so that the ring spacing matches the circumferential spacing, and
    triangle fans at the two tips. The long axis is `axis` (0, 1 or 2)."""
    L, r = float(half_length), float(radius)
    # meridian of the body in the (s, rho) half plane, sampled finely, then re-sampled by arc length
    t = np.linspace(0.0, np.pi, 4001)
    ct, st = np.cos(t), np.sin(t)
    # superellipse |s/L|^p + |rho/r|^p = 1 along the ray of angle t in the (s/L, rho/r) plane
    k = (np.abs(ct) ** p + np.abs(st) ** p) ** (-1.0 / p)
    s_m, rho_m = L * ct * k, r * st * k
    seg = np.hypot(np.diff(s_m), np.diff(rho_m))
    arc = np.concatenate([[0.0], np.cumsum(seg)])
    total = arc[-1]
    circ = 2.0 * np.pi * r
    # F ~ 2 n_phi n_ring with n_ring / n_phi ~ total / (circ / ... ) for isotropic cells
    ratio = total / circ
    n_phi = max(4, int(round(math.sqrt(nelem_target / (2.0 * ratio)))))
    n_ring = max(2, int(round(nelem_target / (2.0 * n_phi))))
    a = np.linspace(0.0, total, n_ring + 2)[1:-1]  # interior rings
    s_r = np.interp(a, arc, s_m)
    rho_r = np.interp(a, arc, rho_m)
    phi = 2.0 * np.pi * np.arange(n_phi) / n_phi
    pts = [np.array([L, 0.0, 0.0])]
    for i in range(n_ring):
        c, sn = np.cos(phi), np.sin(phi)
        # superellipse cross-section |y/R|^p + |z/R|^p = 1 scaled to the ring radius
        kk = (np.abs(c) ** p + np.abs(sn) ** p) ** (-1.0 / p)
        ring = np.stack([np.full(n_phi, s_r[i]), rho_r[i] * c * kk, rho_r[i] * sn * kk], 1)
        pts.extend(ring)
    pts.append(np.array([-L, 0.0, 0.0]))
    xyz = np.array(pts)
    tris = []
    top, bot = 0, len(pts) - 1

    def rid(i, j):
        return 1 + i * n_phi + (j % n_phi)

    for j in range(n_phi):
        tris.append((top, rid(0, j), rid(0, j + 1)))
    for i in range(n_ring - 1):
        for j in range(n_phi):  # each ring-to-ring quad split along a diagonal
            tris.append((rid(i, j), rid(i + 1, j), rid(i + 1, j + 1)))
            tris.append((rid(i, j), rid(i + 1, j + 1), rid(i, j + 1)))
    for j in range(n_phi):
        tris.append((bot, rid(n_ring - 1, j + 1), rid(n_ring - 1, j)))
    con = np.array(tris, dtype=np.int32)
    # the connectivity above is consistently oriented; a coarse ring (few nodes) makes the
    # polyhedron locally non-convex, so orient globally by the sign of the enclosed volume
    x1, x2, x3 = xyz[con[:, 0]], xyz[con[:, 1]], xyz[con[:, 2]]
    vol6 = np.einsum("ij,ij->", np.cross(x2 - x1, x3 - x1), x1)
    if vol6 < 0:
        con = con[:, [0, 2, 1]]
    perm = [0, 1, 2]
    perm[0], perm[axis] = perm[axis], perm[0]
    xyz = xyz[:, perm]
    if axis != 0:  # an odd permutation of the axes flips the orientation
        con = con[:, [0, 2, 1]]
    return Mesh(np.ascontiguousarray(xyz), con, np.ones(con.shape[0], dtype=np.int32),
                f"fibre_L{L:.3f}_r{r:.3f}_p{p:.2f}_F{con.shape[0]}")


def random_fibre(rng: np.random.Generator, nelem_target: int) -> tuple[Mesh, dict]:
    """Fibre-like body: slender superellipsoid of aspect ratio U[10, 50], exponent U[2, 8],
    cross-section radius 0.1 (half length 1 to 5), long axis drawn at random, meshed with
    near-isotropic triangles (see `fibre`)."""
    ar = float(rng.uniform(10.0, 50.0))
    p = float(rng.uniform(2.0, 8.0))
    axis = int(rng.integers(0, 3))
    m = fibre(ar * 0.1, 0.1, p, nelem_target, axis)
    return m, {"aspect": ar, "p": p, "axis": axis, "kind": "fibre", "nelem": m.nelem}


def batch_bodies(n_super: int = 512, n_fibre: int = 64, seed0: int = 0):
    """The batched workload: seeded superellipsoids (axes U[0.5,2], p U[2,8]) and fibres
    (aspect U[10,50]); element counts log-uniform in [500, 8000] rounded to the nearest
    geodesic size 20 nu^2. Body k uses numpy.random.default_rng(seed0 + k)."""
    bodies = []
    for k in range(n_super + n_fibre):
        rng = np.random.default_rng(seed0 + k)
        f = math.exp(rng.uniform(math.log(500.0), math.log(8000.0)))
        nu = nu_for_elements(f)
        if k < n_super:
            m, info = random_superellipsoid(rng, nu)
        else:
            m, info = random_fibre(rng, 20 * nu * nu)
        info["seed"] = seed0 + k
        bodies.append((m, info))
    return bodies


def body_in_sphere(inner: Mesh, outer_radius: float, outer_nu: int) -> Mesh:
    """The reference's bounded resistance set-up (run/mesh/sphere-in-sphere-*.msh): the body is
    wall 1, a far outer sphere is wall 2. Normals point out of the fluid: into the body on
    wall 1 and outward on wall 2."""
    o = geodesic_sphere(outer_nu, outer_radius)
    con_in = inner.con[:, [0, 2, 1]]  # flip: normals into the body
    xyz = np.vstack([inner.xyz, o.xyz])
    con = np.vstack([con_in, o.con + inner.nnodes]).astype(np.int32)
    bcid = np.concatenate([np.ones(inner.nelem, np.int32), np.full(o.nelem, 2, np.int32)])
    return Mesh(xyz, con, bcid, f"{inner.name}_in_sphere{outer_radius:g}")


# --------------------------------------------------------------------------------------------
# readers (surface triangles only)
# --------------------------------------------------------------------------------------------
def cubed_sphere(n: int, radius: float = 1.0) -> Mesh:
    """Sphere of 6 n^2 bilinear quadrilaterals: the faces of a cube divided n x n, nodes projected
    radially (equiangular spacing), counter-clockwise seen from outside (outward normals)."""
    t = np.tan(np.linspace(-np.pi / 4, np.pi / 4, n + 1))
    pts, quads, index = [], [], {}

    def node(p):
        key = tuple(np.round(p / np.linalg.norm(p), 12))
        if key not in index:
            index[key] = len(pts)
            pts.append(radius * p / np.linalg.norm(p))
        return index[key]

    axes = [(0, 1, 2), (1, 2, 0), (2, 0, 1)]
    for (a, b, c) in axes:
        for sign in (1.0, -1.0):
            for i in range(n):
                for j in range(n):
                    corner = []
                    for (di, dj) in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        p = np.zeros(3)
                        p[a] = sign
                        p[b] = t[i + di]
                        p[c] = t[j + dj]
                        corner.append(node(p))
                    quads.append(corner if sign > 0 else corner[::-1])
    xyz = np.array(pts)
    con = np.array(quads, dtype=np.int32)
    return Mesh(xyz, con, np.ones(len(con), np.int32), f"cubed_sphere{n}")


def split_some_quads(mesh: Mesh, every: int = 3) -> Mesh:
    """Mixed mesh: every `every`-th quadrilateral replaced by two triangles (same nodes)."""
    rows, bc = [], []
    for k, q in enumerate(mesh.con):
        if k % every == 0 and q[3] >= 0:
            rows += [[q[0], q[1], q[2], -1], [q[0], q[2], q[3], -1]]
            bc += [mesh.bcid[k], mesh.bcid[k]]
        else:
            rows.append(list(q))
            bc.append(mesh.bcid[k])
    return Mesh(mesh.xyz.copy(), np.array(rows, np.int32), np.array(bc, np.int32), mesh.name + "_mixed")


def read_gmsh(path: str) -> Mesh:
    """gmsh ASCII v2 or v4 surface mesh (src/lib_meshRead.f90:244-629): triangles (type 2) and
    4-node quadrilaterals (type 3) are kept, the physical tag becomes the wall id (the used tags in
    ascending order numbered 1, 2, ...: the tag itself when the tags are 1..n, as the reference
    assumes), and the names of $PhysicalNames become Mesh.walls. Nodes are renumbered densely in file
    order."""
    with open(path, "r") as fh:
        lines = fh.read().split("\n")
    i = 0
    version = 2.0
    nodes: dict = {}
    order: list = []
    tris: list = []
    tags: list = []
    phys_of_entity: dict = {}
    phys_names: dict = {}
    while i < len(lines):
        ln = lines[i].strip()
        if ln == "$PhysicalNames":
            n = int(lines[i + 1])
            for k in range(n):
                t = lines[i + 2 + k].split(None, 2)
                phys_names[int(t[1])] = t[2].strip().strip('"')
            i += 2 + n
        elif ln == "$MeshFormat":
            version = float(lines[i + 1].split()[0])
            i += 2
        elif ln == "$Entities" and version >= 4:
            hdr = [int(x) for x in lines[i + 1].split()]
            npts, ncur, nsurf, nvol = hdr[:4]
            j = i + 2 + npts + ncur
            for k in range(nsurf):
                t = lines[j + k].split()
                tag = int(t[0])
                nphys = int(t[7])
                phys_of_entity[tag] = int(t[8]) if nphys > 0 else tag
            i = j + nsurf + nvol
        elif ln == "$Nodes":
            if version < 4:
                n = int(lines[i + 1])
                for k in range(n):
                    t = lines[i + 2 + k].split()
                    nodes[int(t[0])] = (float(t[1]), float(t[2]), float(t[3]))
                    order.append(int(t[0]))
                i += 2 + n
            else:
                hdr = [int(x) for x in lines[i + 1].split()]
                nblocks = hdr[0]
                j = i + 2
                for _ in range(nblocks):
                    b = [int(x) for x in lines[j].split()]
                    nb = b[3]
                    ids = [int(lines[j + 1 + k]) for k in range(nb)]
                    for k in range(nb):
                        t = lines[j + 1 + nb + k].split()
                        nodes[ids[k]] = (float(t[0]), float(t[1]), float(t[2]))
                        order.append(ids[k])
                    j += 1 + 2 * nb
                i = j
        elif ln == "$Elements":
            if version < 4:
                n = int(lines[i + 1])
                for k in range(n):
                    t = [int(x) for x in lines[i + 2 + k].split()]
                    if t[1] in (2, 3):
                        ntag = t[2]
                        tags.append(t[3] if ntag > 0 else 1)
                        tris.append(t[3 + ntag:3 + ntag + (3 if t[1] == 2 else 4)])
                i += 2 + n
            else:
                hdr = [int(x) for x in lines[i + 1].split()]
                nblocks = hdr[0]
                j = i + 2
                for _ in range(nblocks):
                    b = [int(x) for x in lines[j].split()]
                    edim, etag, etype, nb = b
                    for k in range(nb):
                        t = [int(x) for x in lines[j + 1 + k].split()]
                        if etype in (2, 3):
                            tris.append(t[1:4] if etype == 2 else t[1:5])
                            tags.append(phys_of_entity.get(etag, etag))
                    j += 1 + nb
                i = j
        else:
            i += 1
    used = sorted({v for t in tris for v in t}, key=lambda v: order.index(v) if len(order) < 50000 else v)
    remap = {v: k for k, v in enumerate(used)}
    xyz = np.array([nodes[v] for v in used], dtype=np.float64)
    width = 4 if any(len(t) == 4 for t in tris) else 3
    con = np.array([[remap[v] for v in t] + [-1] * (width - len(t)) for t in tris], dtype=np.int32)
    utags = sorted(set(tags))
    wall = {t: k + 1 for k, t in enumerate(utags)}
    bcid = np.array([wall[t] for t in tags], dtype=np.int32)
    names = [phys_names.get(t, str(t)) for t in utags] if phys_names else None
    return Mesh(xyz, con, bcid, path.split("/")[-1], names)
