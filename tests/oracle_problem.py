"""Explicit dense form of the contact problem, built from the oracle's A / A^T.

Used only to check solver outputs against conditions the mathematics fixes
(KKT, exhaustive active-set optimum) on small problems.
"""
import numpy as np


def dense_problem(orc, scene, dt, friction=False):
    """Returns (B, invM, c, U) with f(l) = 1/2 l^T Q l + c^T l, Q = dt^2 B diag(invM) B^T."""
    nc = len(orc.contacts()["dij"])
    na = scene.n_active
    A = np.zeros((3 * nc, 6 * na))
    for k in range(6 * na):
        e = np.zeros(6 * na)
        e[k] = 1.0
        A[:, k] = orc.amat(e)
    ct = orc.contacts()
    dim = scene.dim
    off = scene.n_inactive
    invM = np.zeros(6 * na)
    U = np.zeros(6 * na)
    for i in range(na):
        invM[3 * i:3 * i + dim] = 1.0 / scene.m[off + i]
        U[3 * i:3 * i + dim] = scene.v[off + i, :dim]
        if dim == 2:
            invM[3 * na + 3 * i + 2] = 1.0 / scene.J[off + i, 2]
            U[3 * na + 3 * i + 2] = scene.omega[off + i, 2]
        else:
            invM[3 * na + 3 * i:3 * na + 3 * i + 3] = 1.0 / scene.J[off + i]
            U[3 * na + 3 * i:3 * na + 3 * i + 3] = scene.omega[off + i]
    Cg = np.zeros(3 * nc)
    for c in range(nc):
        Cg[3 * c:3 * c + dim] = ct["dij"][c] * ct["nij"][c, :dim]
    Cg += dt * (A @ U)
    if friction:
        return A, invM, Cg, U
    N = np.zeros((nc, 3 * nc))
    for c in range(nc):
        N[c, 3 * c:3 * c + dim] = ct["nij"][c, :dim]
    return N @ A, invM, N @ Cg, U


def exhaustive_nofriction(B, invM, c, dt):
    """Minimiser of 1/2 l^T Q l + c^T l over l >= 0 by active-set enumeration (tiny nc).

    Returns M^-1 B^T l (unique even when l is not)."""
    import itertools
    Q = dt * dt * (B * invM) @ B.T
    n = len(c)
    best = None
    for k in range(n + 1):
        for S in itertools.combinations(range(n), k):
            S = list(S)
            lam = np.zeros(n)
            if S:
                sol, *_ = np.linalg.lstsq(Q[np.ix_(S, S)], -c[S], rcond=None)
                if np.any(sol < -1e-12):
                    continue
                lam[S] = sol
            g = Q @ lam + c
            if np.all(g >= -1e-9 * max(1.0, np.abs(c).max())):
                f = 0.5 * lam @ Q @ lam + c @ lam
                if best is None or f < best[0] - 1e-15:
                    best = (f, lam)
    assert best is not None
    return best[1], invM * (B.T @ best[1])
