"""Jacobi-preconditioned BiCGSTAB, the Krylov alternative of the reference's LIS path for square
systems (src/mod_par.f90:288). CPU: the oracle's restatement against scipy's bicgstab with the same
preconditioner (same algorithm: equal iteration counts, equal iterates to rounding) and against a
direct solve. GPU: the device solver against the oracle and the direct solve, on random diagonally
dominant systems (an odd order, several right-hand sides) and on a Laplace BEM system."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import krylov_oracle as KO  # noqa: E402

from andromeda_amd import mesh as M  # noqa: E402


def _system(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) / np.sqrt(n)
    A += np.diag(2.0 + rng.uniform(0, 1, n))
    return A, rng.normal(size=(n, 3))


def test_oracle_matches_scipy_and_direct_solve():
    from scipy.sparse.linalg import LinearOperator, bicgstab
    A, B = _system(300, 1)
    b = B[:, 0]
    x, nit, st = KO.bicgstab_jacobi(A, b, tol=1e-10)
    assert st == 0
    assert np.linalg.norm(x - np.linalg.solve(A, b)) / np.linalg.norm(x) < 1e-8
    dinv = 1.0 / np.diag(A)
    M_ = LinearOperator(A.shape, matvec=lambda v: dinv * v)
    count = [0]
    xs, info = bicgstab(A, b, rtol=1e-10, atol=0.0, M=M_, maxiter=1000,
                        callback=lambda xk: count.__setitem__(0, count[0] + 1))
    assert info == 0 and abs(count[0] - nit) <= 1
    assert np.linalg.norm(xs - x) / np.linalg.norm(x) < 1e-9


def test_oracle_stopping_rules():
    A, B = _system(120, 2)
    b = B[:, 0]
    x, nit, st = KO.bicgstab_jacobi(A, b, tol=1e-14, maxit=3)
    assert (nit, st) == (3, 1)
    x, nit, st = KO.bicgstab_jacobi(A, np.zeros(120))
    assert (nit, st) == (0, 0) and not x.any()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [301, 1000])
def test_gpu_bicgstab_random(ctx, n):
    A, B = _system(n, n)
    X, st = ctx.bicgstab_dense(A, B, tol=1e-10)
    Xd = np.linalg.solve(A, B)
    assert st["status"] == [0, 0, 0]
    assert np.linalg.norm(X - Xd) / np.linalg.norm(Xd) < 1e-8
    for k in range(3):
        _, nit, _ = KO.bicgstab_jacobi(A, B[:, k], tol=1e-10)
        assert abs(st["nit"][k] - nit) <= 2


@pytest.mark.gpu
def test_gpu_bicgstab_laplace_system(ctx, orc):
    m = M.icosphere(3)
    Ho, Go, _ = orc.assemble_laplace(m)
    sysM, rhsM = orc.laplace_dirichlet_system(m, Ho, Go)
    b = rhsM @ (-m.xyz[:, 0])
    X, st = ctx.bicgstab_dense(sysM, b, tol=1e-10)
    xo, nit, sto = KO.bicgstab_jacobi(sysM, b, tol=1e-10)
    assert st["status"][0] == 0 and sto == 0 and abs(st["nit"][0] - nit) <= 2
    xd = np.linalg.solve(sysM, b)
    assert np.linalg.norm(X[:, 0] - xd) / np.linalg.norm(xd) < 1e-8
