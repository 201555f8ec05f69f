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


def sparse_problem(orc, scene, dt, friction=False):
    """Sparse B (rows: nc or 3nc), invM, c and U, probing the oracle's A column by column."""
    import scipy.sparse as sp
    ct = orc.contacts()
    nc = len(ct["dij"])
    na = scene.n_active
    dim = scene.dim
    rows, cols, vals = [], [], []
    e = np.zeros(6 * na)
    for k in range(6 * na):
        e[k] = 1.0
        col = orc.amat(e)
        e[k] = 0.0
        nz = np.nonzero(col)[0]
        rows.append(nz)
        cols.append(np.full(nz.size, k))
        vals.append(col[nz])
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(3 * nc, 6 * na))
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
        return A, invM, Cg, U, ct["nij"]
    N = sp.csr_matrix((ct["nij"][:, :dim].ravel(), (np.repeat(np.arange(nc), dim),
                                                   (3 * np.arange(nc)[:, None] + np.arange(dim)).ravel())),
                      shape=(nc, 3 * nc))
    return N @ A, invM, N @ Cg, U, ct["nij"]


def cone_projection(lam, nij, mu, dim):
    """Coulomb-cone projection of every contact (lagrange_multiplier.hpp:243-268), vectorised."""
    l = lam.reshape(-1, 3).copy()
    n = nij[:, :3].copy()
    if dim == 2:
        n[:, 2] = 0.0
    ln = np.einsum("cd,cd->c", l[:, :dim], n[:, :dim])
    lt = l[:, :dim] - ln[:, None] * n[:, :dim]
    norm = np.linalg.norm(lt, axis=1)
    out = l.copy()
    cut = norm > mu * ln
    zero = cut & (ln <= -mu * norm)
    proj = cut & ~zero
    out[zero, :dim] = 0.0
    nn = (mu * mu) * (norm[proj] + ln[proj] / mu) / (mu * mu + 1)
    out[proj, :dim] = (nn / norm[proj])[:, None] * lt[proj] + (nn / mu)[:, None] * n[proj, :dim]
    return out.ravel()


def exact_solution(B, invM, c, dt, U, friction=False, nij=None, mu=0.3, dim=3, tol=1e-13, max_iter=400000):
    """High-accuracy minimiser of 1/2 l^T Q l + c^T l over the feasible set, Q = dt^2 B M^-1 B^T.

    Accelerated projected gradient with the fixed step 1/L (L = largest eigenvalue
    of Q by power iteration) and gradient restart; stops when the fixed-point
    residual |l - Proj(l - g/L)| < tol.  Returns u = U + dt M^-1 B^T l.
    """
    Bt = B.T.tocsr()
    Qv = lambda x: dt * dt * (B @ (invM * (Bt @ x)))
    rng = np.random.default_rng(0)
    x = rng.random(B.shape[0])
    for _ in range(200):
        x = Qv(x)
        x /= np.linalg.norm(x)
    L = float(x @ Qv(x)) * 1.01
    proj = (lambda z: cone_projection(z, nij, mu, dim)) if friction else (lambda z: np.maximum(z, 0.0))
    lam = np.zeros(B.shape[0])
    y = lam.copy()
    th = 1.0
    for _ in range(max_iter):
        g = Qv(y) + c
        ln = proj(y - g / L)
        if np.linalg.norm(ln - proj(ln - (Qv(ln) + c) / L)) < tol:
            lam = ln
            break
        th1 = 0.5 * (th * np.sqrt(4 + th * th) - th * th)
        beta = th * (1 - th) / (th * th + th1)
        if g @ (ln - lam) > 0:
            y = ln.copy()
            th1 = 1.0
        else:
            y = ln + beta * (ln - lam)
        lam = ln
        th = th1
    return lam, U + dt * invM * (Bt @ lam), L
