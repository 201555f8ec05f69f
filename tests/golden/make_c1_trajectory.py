"""Writes tests/golden/c1_trajectory.npz: the oracle's 200-step trajectory of the 1,024-disk
workload (32 x 32 jittered lattice, radii U[0.008, 0.012], one plane, gravity; NoFriction, adaptive
APGD tol 1e-6, dt 0.005, dmax 0.025), positions every 40 steps, and the same run with the line
search's acceptance test perturbed by +-4e-16 (objective_perturbation), whose spread is the
trajectory's own sensitivity to rounding.  About 10 minutes of CPU; run from the repository root:

    python tests/golden/make_c1_trajectory.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from scopi_amd import scenes  # noqa: E402


def main():
    oracle.build()
    cfg = scenes.CONFIG1
    sc = scenes.disks_2d()
    out = {}
    for tag, pert in (("ref", 0.0), ("plus", 4e-16), ("minus", -4e-16)):
        o = oracle.Oracle(sc, vap=oracle.VAP_FPD, dmax=cfg["dmax"], kd_tree_radius=cfg["kd_tree_radius"],
                          tolerance=cfg["tolerance"], objective_perturbation=pert)
        pos = []
        for k in range(cfg["steps"]):
            o.run(cfg["dt"], k + 1, k)
            if (k + 1) % 40 == 0:
                pos.append(o.state()["pos"][:, :2].copy())
        out[tag] = np.stack(pos)
        print(tag, "done", flush=True)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "c1_trajectory.npz"), steps=np.arange(40, 201, 40),
                        **out)


if __name__ == "__main__":
    main()
