"""Error measure of a resistance tensor against a reference one.

FT is the 6x6 array returned by the resistance drivers (row k = force and torque of rigid-body
problem k). The translation block K (force/velocity) scales as a, the rotation block Omega
(torque/angular velocity) as a^3 and the two coupling blocks as a^2, so each block is compared on
its own scale: ||dK||/||K||, ||dOmega||/||Omega|| and, because the coupling vanishes for bodies
with a centre of symmetry, ||dC||/sqrt(||K|| ||Omega||) for both coupling blocks (Frobenius norms).
`entry_error` is the entry-wise measure beside it: the largest |dR_ij| / |R_ij| over the entries of
each block whose magnitude exceeds 1e-3 of that block's norm (the coupling blocks against
1e-3 sqrt(||K|| ||Omega||)).
"""
import numpy as np


def tensor_error(FT, FT_ref) -> float:
    FT, R = np.asarray(FT, dtype=np.float64), np.asarray(FT_ref, dtype=np.float64)
    nK = np.linalg.norm(R[:3, :3])
    nO = np.linalg.norm(R[3:, 3:])
    eK = np.linalg.norm(FT[:3, :3] - R[:3, :3]) / nK
    eO = np.linalg.norm(FT[3:, 3:] - R[3:, 3:]) / nO
    sc = np.sqrt(nK * nO)
    eC = max(np.linalg.norm(FT[:3, 3:] - R[:3, 3:]), np.linalg.norm(FT[3:, :3] - R[3:, :3])) / sc
    return float(max(eK, eO, eC))


def entry_error(FT, FT_ref, rel_floor: float = 1e-3) -> float:
    FT, R = np.asarray(FT, dtype=np.float64), np.asarray(FT_ref, dtype=np.float64)
    nK = np.linalg.norm(R[:3, :3])
    nO = np.linalg.norm(R[3:, 3:])
    sc = np.sqrt(nK * nO)
    worst = 0.0
    for rows, cols, scale in ((slice(0, 3), slice(0, 3), nK), (slice(3, 6), slice(3, 6), nO),
                              (slice(0, 3), slice(3, 6), sc), (slice(3, 6), slice(0, 3), sc)):
        r, f = R[rows, cols], FT[rows, cols]
        big = np.abs(r) > rel_floor * scale
        if np.any(big):
            worst = max(worst, float(np.max(np.abs(f[big] - r[big]) / np.abs(r[big]))))
    return worst
