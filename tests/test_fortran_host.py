"""The Fortran host (ISO_C_BINDING module + implicit-interface LSQR calls) against the library
called from Python on the same mesh."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "andromeda_amd", "fortran", "resistance_host")
MESH = os.path.join(ROOT, "tests", "golden", "sphere-01.msh")


def _host():
    if not os.path.exists(HOST):
        pytest.skip("Fortran host not built (amdflang missing)")
    return HOST


def test_fortran_host_links_and_starts():
    r = subprocess.run([_host()], capture_output=True, text=True, timeout=60)
    assert "usage" in (r.stdout + r.stderr)


@pytest.mark.gpu
def test_fortran_host_matches_python(ctx):
    from andromeda_amd import api
    from andromeda_amd.mesh import read_gmsh
    r = subprocess.run([_host(), MESH, "fp64"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout
    m = re.search(r"lsqr: maxit (\d+) nit (\d+) ierr (\d+)\s+max\|q \+ n_x\|\s+(\S+)", out)
    assert m, out
    assert int(m.group(1)) == 10 * (2 * 836 + 50)  # maxit 0 written back as 10(m+n+50)
    assert int(m.group(3)) <= 3
    assert float(m.group(4)) < 0.05  # flux of u = x on a coarse sphere: q = -n_x
    FT = np.array([[float(v) for v in line.split(":")[1].split()]
                   for line in out.splitlines() if line.startswith("FT ")])
    assert FT.shape == (6, 6)
    mesh = read_gmsh(MESH)
    ref = ctx.resistance(mesh, precision=api.PREC_FP64, eps=1e-12)
    assert np.max(np.abs(FT - ref.FT)) <= 1e-9 * np.abs(ref.FT).max()
    # sphere of radius 0.5: drag 6 pi mu a U (coarse mesh: a few per cent)
    assert abs(FT[0, 0] / (6 * np.pi * 0.5) - 1) < 0.05
    # domain values through the Fortran interface: a uniform boundary flow is returned at the centre
    # (up to the reference's single-precision 8 pi), the same numbers as the Python binding's
    m = re.search(r"domain velocity:\s+(\S+)\s+(\S+)\s+(\S+)\s+closure error:\s+(\S+)", out)
    assert m, out
    vel = np.array([float(m.group(k)) for k in (1, 2, 3)])
    u = np.zeros((mesh.nnodes, 3))
    u[:, 0] = 1.0
    vp, _, ep = ctx.stokes_domain_values(mesh, mesh.xyz.mean(axis=0)[None, :], u, np.zeros((mesh.nelem, 3)))
    assert np.abs(vel - vp[0]).max() < 1e-7 and abs(abs(vel[0]) - 1.0) < 1e-6 and np.abs(vel[1:]).max() < 1e-6
    assert float(m.group(4)) < 1e-6 and ep[0] < 1e-6
