"""Coulomb-cone helpers of the GPU tests' KKT checks."""
import numpy as np


def cone_projection(v, n, mu):
    """Projection of 3-vectors v on the Coulomb cones {|x_t| <= mu x_n} of normals n."""
    vn = (v * n).sum(1)
    vt = v - vn[:, None] * n
    nt = np.linalg.norm(vt, axis=1)
    out = v.copy()
    zero = mu * nt <= -vn
    bd = ~(zero | (nt <= mu * vn))
    a = (vn[bd] + mu * nt[bd]) / (1 + mu * mu)
    out[bd] = a[:, None] * (n[bd] + mu * vt[bd] / nt[bd][:, None])
    out[zero] = 0.0
    return out
