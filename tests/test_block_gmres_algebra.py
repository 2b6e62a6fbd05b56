"""CPU check of the block GMRES algebra that the device kernels implement (andr_gmres.hip:
kb_binit, kb_bsmall, kb_bsolve), written out with the kernels' storage and indexing: Cholesky
factors with dropped columns, the Householder band update over a sliding window of L + 1 rows, the
rotated right-hand sides, the packed triangle and the back substitution. The solution is compared
with a direct solve; dead columns (a zero right-hand side, the padding to 8) must stay zero and a
right-hand side that repeats another must still be solved. The GPU tests check the kernels
themselves against the plain FP64 path."""
import math

import numpy as np

NB = 8


def chol_drop(G, nrm2, live, L, drop):
    R = np.zeros((8, 8))
    Ri = np.zeros((8, 8))
    brk = 0
    for c in range(L):
        d = G[c, c] - sum(R[t, c] ** 2 for t in range(c))
        if not live[c] or not (d > drop * nrm2[c]):
            brk |= int(live[c] and nrm2[c] > 0)  # a column without a basis vector is no breakdown
            continue
        R[c, c] = math.sqrt(d)
        for c2 in range(c + 1, L):
            R[c, c2] = (G[c, c2] - sum(R[t, c] * R[t, c2] for t in range(c))) / R[c, c]
    for k in range(L):
        if R[k, k] == 0:
            continue
        Ri[k, k] = 1 / R[k, k]
        for a in range(k - 1, -1, -1):
            if R[a, a] != 0:
                Ri[a, k] = -sum(R[a, t] * Ri[t, k] for t in range(a + 1, k + 1)) / R[a, a]
    return brk, R, Ri


def block_gmres(A, B, tol, m):
    """One cycle of block GMRES as the kernels run it; returns X (n x 8) and the rounds used."""
    n, L = A.shape[0], B.shape[1]
    R = np.zeros((n, NB))
    R[:, :L] = B
    bn = np.linalg.norm(R, axis=0)
    live = [k < L and bn[k] > 0 for k in range(NB)]
    Q = np.zeros((m + 1, n, NB))
    C1 = np.zeros((m + 1, 8, 8))
    C2 = np.zeros((m + 2, 8, 8))
    Rp = np.zeros((L * m) * (L * m + 1) // 2)
    hv = np.zeros(8 * L * m)
    gB = np.zeros((L * (m + 1), NB))
    G = R.T @ R
    _, R0, Ri = chol_drop(G, np.diag(G), live, L, 1e-24)  # kb_binit
    gB[:L, :L] = R0[:L, :L]
    Q[0] = R @ Ri
    used = None
    for j in range(m):
        W = A @ Q[j]  # the Krylov product
        for b in range(j + 1):
            C1[b] = Q[b].T @ W
        W = W - sum(Q[b] @ C1[b] for b in range(j + 1))
        for b in range(j + 1):
            C2[b] = Q[b].T @ W
        C2[j + 1] = W.T @ W
        # kb_bsmall
        G2 = C2[j + 1] - sum(C2[b][:L].T @ C2[b][:L] for b in range(j + 1))
        nrm = np.array([C2[j + 1][c, c] + sum((C1[b][:L, c] ** 2).sum() for b in range(j + 1)) for c in range(8)])
        brk, Rs, Ri = chol_drop(G2, nrm, live, L, 1e-20)
        LJ, rows = L * j, L * (j + 2)

        def hval(r, c):
            if c >= L or r >= rows:
                return 0.0
            if r < L * (j + 1):
                b, l = divmod(r, L)
                return C1[b][l, c] + C2[b][l, c]
            cc = r - L * (j + 1)
            return Rs[cc, c] if cc <= c else 0.0

        win = [[hval(i, c) if i <= L else 0.0 for i in range(8)] for c in range(8)]
        sv = np.zeros((8, 8))
        for q in range(LJ + L):
            if q < LJ:
                tau, v = hv[8 * q], [1.0] + [hv[8 * q + i] if i <= L else 0.0 for i in range(1, 8)]
            else:
                cq = q - LJ
                w = win[cq]
                x0, s2 = w[0], sum(w[i] ** 2 for i in range(1, L + 1))
                alpha, t, sc = x0, 0.0, 0.0
                if s2 > 0:
                    nx = math.sqrt(x0 * x0 + s2)
                    alpha = -nx if x0 > 0 else nx
                    t, sc = (alpha - x0) / alpha, 1 / (x0 - alpha)
                Rp[q * (q + 1) // 2 + q] = alpha
                sv[cq, 0] = hv[8 * q] = t
                for i in range(1, L + 1):
                    sv[cq, i] = hv[8 * q + i] = w[i] * sc
                tau, v = t, [1.0] + [sv[cq, i] if i <= L else 0.0 for i in range(1, 8)]
            for c in range(L):
                qc = LJ + c
                if q >= qc:
                    continue
                w = win[c]
                s = sum(v[i] * w[i] for i in range(8))
                w = [w[i] - tau * s * v[i] for i in range(8)]
                Rp[qc * (qc + 1) // 2 + q] = w[0]
                win[c] = [w[i + 1] if i < L else (hval(q + 1 + L, c) if i == L else 0.0) for i in range(8)]
        est = {}
        for k in range(NB):
            if not live[k]:
                continue
            gw = [gB[LJ + i, k] if i <= L else 0.0 for i in range(8)]
            for t in range(L):
                v = [1.0] + [sv[t, i] if i <= L else 0.0 for i in range(1, 8)]
                s = sum(v[i] * gw[i] for i in range(8))
                gw = [gw[i] - sv[t, 0] * s * v[i] for i in range(8)]
                gB[LJ + t, k] = gw[0]
                r = LJ + t + 1 + L
                gw = [gw[i + 1] if i < L else ((gB[r, k] if r < rows else 0.0) if i == L else 0.0) for i in range(8)]
            for i in range(L):
                gB[L * (j + 1) + i, k] = gw[i]
            est[k] = math.sqrt(sum(gw[i] ** 2 for i in range(L)))
        Q[j + 1] = (W - sum(Q[b] @ C2[b] for b in range(j + 1))) @ Ri
        if brk or all(est[k] <= 0.5 * tol * bn[k] for k in est):
            used = j + 1
            break
    kk = used or m
    # kb_bsolve
    N = L * kk
    g = gB.copy()
    for q in range(N - 1, -1, -1):
        off = q * (q + 1) // 2
        d = Rp[off + q]
        g[q] = g[q] / d if d != 0 else 0.0
        g[:q] -= np.outer(Rp[off:off + q], g[q])
    X = np.zeros((n, NB))
    for b in range(kk):
        Y = np.zeros((8, 8))
        Y[:L] = g[L * b:L * b + L]
        X += Q[b] @ Y
    return X, kk


def _system(n, seed):
    rng = np.random.default_rng(seed)
    A = np.eye(n) + rng.standard_normal((n, n)) / (2.5 * math.sqrt(n))  # spectrum in a disc around 1
    return A, rng.standard_normal((n, 6))


def test_block_gmres_solves_six_columns():
    A, B = _system(150, 1)
    X, kk = block_gmres(A, B, 1e-10, 60)
    ref = np.linalg.solve(A, B)
    assert np.abs(X[:, 6:]).max() == 0.0  # padding columns stay exact zeros
    assert np.linalg.norm(X[:, :6] - ref) / np.linalg.norm(ref) < 1e-9
    # one Krylov space for the six columns: fewer rounds than one column alone needs steps
    x1, k1 = block_gmres(A, B[:, :1], 1e-10, 60)
    assert np.linalg.norm(x1[:, 0] - ref[:, 0]) / np.linalg.norm(ref[:, 0]) < 1e-9
    assert kk < k1


def test_block_gmres_dead_and_repeated_columns():
    A, B = _system(120, 2)
    B[:, 2] = 0.0  # converged (dead) column
    B[:, 4] = B[:, 1]  # dependent column: dropped from the first basis block, still solved
    X, _ = block_gmres(A, B, 1e-10, 60)
    ref = np.linalg.solve(A, B)
    assert np.abs(X[:, 2]).max() == 0.0
    for k in (0, 1, 3, 4, 5):
        assert np.linalg.norm(X[:, k] - ref[:, k]) / np.linalg.norm(ref[:, k]) < 1e-9
