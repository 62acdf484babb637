"""Frame-correct conserved quantities for drift checks."""
import numpy as np

from solaris_amd import scenarios as S


def barycentric_invariants(y, m):
    """Energy and angular momentum in the barycentric frame computed from
    astrocentric states (the reference's Integrals use the astrocentric
    frame, where they are not conserved)."""
    msk = m > 0
    M = m.sum()
    rb = (m[:, None] * y[:, :3]).sum(0) / M
    vb = (m[:, None] * y[:, 3:]).sum(0) / M
    R = y[msk, :3] - rb
    Vv = y[msk, 3:] - vb
    mm = m[msk]
    ke = 0.5 * (mm * (Vv ** 2).sum(1)).sum()
    pe = 0.0
    for i in range(len(mm)):
        d = np.linalg.norm(R[i + 1:] - R[i], axis=1)
        pe -= S.GAUSS2 * (mm[i] * mm[i + 1:] / d).sum()
    L = (mm[:, None] * np.cross(R, Vv)).sum(0)
    return ke + pe, L


def drifts(y0, y, m):
    E0, L0 = barycentric_invariants(y0, m)
    E, L = barycentric_invariants(y, m)
    return abs(E / E0 - 1.0), float(np.linalg.norm(L - L0) / np.linalg.norm(L0))
