"""Launch-shape autotuner for the projected-gradient kernels.

The B^T gather kernel assigns a group of G lanes to one particle and the
per-contact kernels run a grid-stride loop over a fixed number of blocks; the
best G and block count depend on the contact count and on the layout of B.
`tune(scene, prm, dt)` solves the same contact problem once per candidate shape
(the solve starts from lambda = 0 every time, so every candidate performs the
same iterations) and keeps the fastest on the context.  It has no counterpart
in the reference, whose solver runs on the CPU.
"""
from __future__ import annotations

import time

from .hip import TUNE_BT_GROUP, TUNE_REDUCE_BLOCKS, HipScene, Params

BT_GROUPS = (1, 2, 4, 8, 16)
REDUCE_BLOCKS = (256, 512, 1024, 2048)


def _time_solve(hs: HipScene, prm: Params, dt: float, repeats: int) -> float:
    best = float("inf")
    for _ in range(repeats):
        t0 = time.perf_counter()
        hs.solve(prm, dt)
        best = min(best, time.perf_counter() - t0)
    return best


def tune(hs: HipScene, prm: Params, dt: float, repeats: int = 2, groups=BT_GROUPS, blocks=REDUCE_BLOCKS):
    """Sweep the launch shapes on the contacts currently held by `hs`.

    The caller has run compute_contacts and set_a_priori_velocity.  Returns
    {"bt_group", "reduce_blocks", "seconds", "table"} and leaves the best shape
    set on `hs`.
    """
    hs.solve(prm, dt)  # first solve allocates the solver's buffers
    table = {}
    hs.set_tuning(TUNE_REDUCE_BLOCKS, blocks[-1])
    for g in groups:
        hs.set_tuning(TUNE_BT_GROUP, g)
        table[(g, blocks[-1])] = _time_solve(hs, prm, dt, repeats)
    g_best = min((k for k in table if k[1] == blocks[-1]), key=table.get)[0]
    hs.set_tuning(TUNE_BT_GROUP, g_best)
    for b in blocks[:-1]:
        hs.set_tuning(TUNE_REDUCE_BLOCKS, b)
        table[(g_best, b)] = _time_solve(hs, prm, dt, repeats)
    g, b = min(table, key=table.get)
    hs.set_tuning(TUNE_BT_GROUP, g)
    hs.set_tuning(TUNE_REDUCE_BLOCKS, b)
    return {"bt_group": g, "reduce_blocks": b, "seconds": table[(g, b)],
            "table": {f"{k[0]}x{k[1]}": v for k, v in table.items()}}
