"""Writes tests/golden/solve_cases.npz: exact optima of the first-step contact problem of
several seeded scenes, computed independently of the reference's APGD.

For each case: the scene is rebuilt from its seed (tests/solve_cases.py), the oracle
computes the contacts and the a-priori velocities, the sparse B is probed from the
oracle's A (which is pinned by test_oracle_geometry.py), and exact_solution() solves
the dual problem to a fixed-point residual of 1e-13.  Also stored: the oracle's own
adaptive-APGD error against that optimum, and the largest eigenvalue L of Q.

Run: python tests/golden/make_solve_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from tests.oracle_problem import exact_solution, sparse_problem  # noqa: E402
from tests.solve_cases import CASES, oracle_kw  # noqa: E402


def main():
    out = {}
    for name, case in CASES.items():
        sc, kw = case["scene"](), oracle_kw(case)
        o = oracle.Oracle(sc, **kw)
        nc = o.compute_contacts()
        o.set_a_priori_velocity(case["dt"])
        st = o.state()
        sc2 = sc.copy()
        sc2.v, sc2.omega = st["v"], st["omega"]
        res = o.solve(case["dt"])
        fr = case["problem"] == 1
        B, invM, c, U, nij = sparse_problem(o, sc2, case["dt"], friction=fr)
        lam, u_ex, L = exact_solution(B, invM, c, case["dt"], U, friction=fr, nij=nij, mu=case["mu"], dim=sc.dim)
        na = sc.n_active
        u_or = np.concatenate([res["u"].ravel(), res["w"].ravel()])
        e_or = float(np.abs(u_or - u_ex).max())
        out[name + "/u_exact"] = u_ex
        out[name + "/lam_exact"] = lam
        out[name + "/e_oracle"] = np.array(e_or)
        out[name + "/L"] = np.array(L)
        out[name + "/nc"] = np.array(nc)
        print(f"{name}: nc={nc} na={na} L={L:.4e} e_oracle={e_or:.3e}", flush=True)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "solve_cases.npz"), **out)


if __name__ == "__main__":
    main()
