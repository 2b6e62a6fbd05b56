"""Test infrastructure only (the oracle): a numpy restatement of the Krylov alternative of the
reference's LIS path for square systems, parLISslvSet "-i bicgstab -p jacobi -tol 1.0e-8
-maxiter 1000" (src/mod_par.f90:288). LIS (an external library, not in the reference's sources) is
not available here; its BiCGSTAB is the published preconditioned algorithm (van der Vorst 1992;
Barrett et al., Templates, 1994, fig. 2.10), restated below with M = diag(A) and LIS's default
stopping rule ||r||_2 <= tol ||b||_2. Pinned in the tests against scipy.sparse.linalg.bicgstab with
the same preconditioner and against numpy.linalg.solve."""
import numpy as np


def bicgstab_jacobi(A, b, x0=None, tol=1e-8, maxit=1000):
    """Returns (x, nit, status): status 0 converged, 1 iteration limit, 2 breakdown."""
    A = np.asarray(A, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.diag(A).copy()
    dinv = np.where(d != 0.0, 1.0 / np.where(d != 0.0, d, 1.0), 1.0)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
    r = b - A @ x
    bnorm = np.linalg.norm(b)
    if np.linalg.norm(r) <= tol * bnorm:
        return x, 0, 0
    rh = r.copy()
    rho = alpha = omega = 1.0
    p = np.zeros_like(b)
    v = np.zeros_like(b)
    for it in range(1, maxit + 1):
        rho_new = rh @ r
        if rho_new == 0.0:
            return x, it - 1, 2
        beta = 0.0 if it == 1 else (rho_new / rho) * (alpha / omega)
        rho = rho_new
        p = r + beta * (p - omega * v)
        y = dinv * p
        v = A @ y
        rv = rh @ v
        if rv == 0.0:
            return x, it, 2
        alpha = rho / rv
        x = x + alpha * y
        s = r - alpha * v
        z = dinv * s
        t = A @ z
        tt = t @ t
        omega = (t @ s) / tt if tt != 0.0 else 0.0
        x = x + omega * z
        r = s - omega * t
        if np.linalg.norm(r) <= tol * bnorm:
            return x, it, 0
        if omega == 0.0:
            return x, it, 2
    return x, maxit, 1
