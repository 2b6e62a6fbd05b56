"""Host-side checks that need no GPU: the C-ABI library loads and exports every entry point the
header declares; mesh generators and the gmsh reader produce valid closed surfaces."""
import os
import re

import numpy as np
import pytest

from andromeda_amd import mesh as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "andromeda_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b((?:andr|lsqr)_\w+)\s*\(", src)))


def test_library_exports_header():
    import ctypes
    from andromeda_amd import api
    if not os.path.exists(api.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(api.LIB_PATH)
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(api.EXPORTED) <= set(names)


def test_missing_library_fails_loudly(tmp_path):
    from andromeda_amd import api
    with pytest.raises(api.AndromedaError):
        api._lib = None
        api.load_library(str(tmp_path / "nope.so"))
    api._lib = None


@pytest.mark.parametrize("nu", [1, 2, 3, 5])
def test_geodesic_sphere(nu):
    m = M.geodesic_sphere(nu)
    assert m.nelem == 20 * nu * nu and m.nnodes == 10 * nu * nu + 2
    x1, x2, x3 = (m.xyz[m.con[:, k]] for k in range(3))
    n = np.cross(x2 - x1, x3 - x1)
    assert np.abs(n.sum(0)).max() < 1e-12               # closed surface
    assert (np.einsum("ij,ij->i", n, x1 + x2 + x3) > 0).all()  # outward (convex body)
    assert np.allclose(np.linalg.norm(m.xyz, axis=1), 1.0)
    # every edge shared by exactly two triangles
    e = np.sort(np.concatenate([m.con[:, [0, 1]], m.con[:, [1, 2]], m.con[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    assert (cnt == 2).all()


def test_superellipsoid_on_surface():
    rng = np.random.default_rng(0)
    m, info = M.random_superellipsoid(rng, 6)
    a = np.array(info["axes"])
    f = np.sum(np.abs(m.xyz / a) ** info["p"], axis=1)
    assert np.abs(f - 1.0).max() < 1e-12


def test_batch_sizes():
    bodies = M.batch_bodies(8, 2)
    for mesh, info in bodies:
        assert 450 <= mesh.nelem <= 8100
        if info["kind"] == "superellipsoid":
            assert mesh.nelem == 20 * info["nu"] ** 2


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_fibre_mesh(axis):
    m = M.fibre(3.0, 0.1, 4.0, 2000, axis)
    x1, x2, x3 = (m.xyz[m.con[:, k]] for k in range(3))
    n = np.cross(x2 - x1, x3 - x1)
    assert np.abs(n.sum(0)).max() < 1e-12
    assert np.einsum("ij,ij->", n, x1) > 0  # outward: positive enclosed volume
    d = np.concatenate([m.con[:, [0, 1]], m.con[:, [1, 2]], m.con[:, [2, 0]]])
    assert len(set(map(tuple, d))) == len(d)  # consistently oriented 2-manifold
    area = 0.5 * np.linalg.norm(n, axis=1)
    longest = np.max([np.linalg.norm(x2 - x1, axis=1), np.linalg.norm(x3 - x2, axis=1),
                      np.linalg.norm(x1 - x3, axis=1)], axis=0)
    assert np.median(longest ** 2 / (2 * area)) < 3.0  # near-isotropic elements
    assert np.abs(m.xyz[:, axis]).max() > 2.9


def test_gmsh_reader_reference_mesh():
    m = M.read_gmsh(os.path.join(ROOT, "tests", "golden", "sphere-01.msh"))
    assert m.nelem == 836 and m.nnodes == 420
    x1, x2, x3 = (m.xyz[m.con[:, k]] for k in range(3))
    assert np.abs(np.cross(x2 - x1, x3 - x1).sum(0)).max() < 1e-10


def test_tensor_error_measure():
    from andromeda_amd.metrics import tensor_error
    R = np.diag([6.0, 6.0, 6.0, 8.0, 8.0, 8.0])
    X = R.copy()
    X[0, 4] = 1e-7  # coupling of a symmetric body: judged on sqrt(|K| |Omega|)
    assert abs(tensor_error(X, R) - 1e-7 / np.sqrt(np.linalg.norm(R[:3, :3]) * np.linalg.norm(R[3:, 3:]))) < 1e-15
    X = R * (1 + 1e-6)
    assert abs(tensor_error(X, R) - 1e-6) < 1e-12


def test_entry_error_measure():
    """Entry-wise relative error over the entries above 1e-3 of their block's norm: small entries
    (below the floor) are left out, the coupling blocks are floored on sqrt(|K| |Omega|)."""
    from andromeda_amd.metrics import entry_error
    R = np.diag([6.0, 6.0, 6.0, 8.0, 8.0, 8.0])
    R[0, 1] = R[1, 0] = 1e-4   # below 1e-3 |K|: not judged
    R[0, 4] = 0.5              # coupling entry above 1e-3 sqrt(|K| |Omega|)
    X = R.copy()
    X[0, 1] *= 2.0
    assert entry_error(X, R) == 0.0
    X[0, 4] = 0.5 * (1 + 3e-6)
    X[3, 3] = 8.0 * (1 - 1e-6)
    assert abs(entry_error(X, R) - 3e-6) < 1e-15


def test_mesh_validate_rejects_bad_inputs():
    """andr_mesh_validate (host only): every check an entry point makes before it reads a mesh --
    wall ids outside 1..nwalls, connectivity out of range, bad rule indices -- is refused with
    ANDR_ERR_ARG and a reason; a valid mesh passes."""
    from andromeda_amd import api
    if not os.path.exists(api.LIB_PATH):
        pytest.skip("library not built")
    m = M.geodesic_sphere(2)
    assert api.validate_mesh(m) == (0, "")
    bad = m.bcid.copy()
    bad[5] = 2
    rc, msg = api.validate_mesh(m, bcid=bad)
    assert rc == -1 and "bcid of element 5" in msg
    bad[5] = 0
    assert api.validate_mesh(m, bcid=bad)[0] == -1
    assert api.validate_mesh(m, bcid=bad * 0 + 2, nwalls=2)[0] == 0
    con = m.con.copy()
    con[3, 1] = m.nnodes
    rc, msg = api.validate_mesh(m, con=con)
    assert rc == -1 and "connectivity" in msg
    con[3, 1] = -1
    assert api.validate_mesh(m, con=con)[0] == -1
    assert api.validate_mesh(m, itri=6)[0] == -1
    assert api.validate_mesh(m, iquad=9)[0] == -1
