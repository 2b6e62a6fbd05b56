"""Error measure of a resistance tensor against a reference one.

FT is the 6x6 array returned by the resistance drivers (row k = force and torque of rigid-body
problem k). The translation block K (force/velocity) scales as a, the rotation block Omega
(torque/angular velocity) as a^3 and the two coupling blocks as a^2, so each block is compared on
its own scale: ||dK||/||K||, ||dOmega||/||Omega|| and, because the coupling vanishes for bodies
with a centre of symmetry, ||dC||/sqrt(||K|| ||Omega||) for both coupling blocks (Frobenius norms).
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
