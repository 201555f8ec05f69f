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

from .hip import TUNE_BT_GROUP, TUNE_GRAPH_ROUNDS, TUNE_REDUCE_BLOCKS, HipScene, Params

BT_GROUPS = (1, 2, 4, 8, 16)
REDUCE_BLOCKS = (256, 512, 1024, 2048)
GRAPH_ROUNDS = (15, 63)


def _time_solve(hs: HipScene, prm: Params, dt: float, repeats: int) -> float:
    best = float("inf")
    for _ in range(repeats):
        t0 = time.perf_counter()
        hs.solve(prm, dt)
        best = min(best, time.perf_counter() - t0)
    return best


def tune(hs: HipScene, prm: Params, dt: float, repeats: int = 2, groups=BT_GROUPS, blocks=REDUCE_BLOCKS,
         graph_rounds=GRAPH_ROUNDS):
    """Sweep the launch shapes on the contacts currently held by `hs`, then whether a hipGraph of
    rounds (bitwise the eager loop) is faster on them.

    The caller has run compute_contacts and set_a_priori_velocity.  Returns
    {"bt_group", "reduce_blocks", "graph_rounds", "seconds", "table"} and leaves the best shape
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
    best_r, best_t = 0, table[(g, b)]
    graphs = {}
    for r in graph_rounds:
        hs.set_tuning(TUNE_GRAPH_ROUNDS, r)
        graphs[r] = _time_solve(hs, prm, dt, repeats)
        if graphs[r] < best_t:
            best_r, best_t = r, graphs[r]
    hs.set_tuning(TUNE_GRAPH_ROUNDS, best_r)
    table_out = {f"{k[0]}x{k[1]}": v for k, v in table.items()}
    table_out.update({f"graph{r}": v for r, v in graphs.items()})
    return {"bt_group": g, "reduce_blocks": b, "graph_rounds": best_r, "seconds": best_t, "table": table_out}


# Batch schedules: loop iterations per lane of the per-contact grid, lanes per particle of the
# B^T gather, and which scenes the per-scene work-queue kernel takes (at most `max_contacts`
# contacts; 0: none) with its workgroup size and grid.
ROW_ITERS = (2, 4, 8)
QUEUE_SHAPES = ((20000, 512, 64), (40000, 1024, 32), (40000, 1024, 64), (80000, 1024, 32))
BATCH_BT_GROUPS = (4, 8, 16)
BATCH_DEFAULT = {"row_iters": 2, "bt_group": 8, "queue_max_contacts": 0, "queue_threads": 256, "queue_blocks": 1024}


def apply_batch_schedule(b, sched):
    """Set a schedule (keys of BATCH_DEFAULT) on a HipBatch; the scenes' arithmetic then depends on it
    (block partitions fix the summation order), the same for a scene alone and in any mix."""
    from .hip import (TUNE_QUEUE_BLOCKS, TUNE_QUEUE_MAX_CONTACTS, TUNE_QUEUE_THREADS, TUNE_ROW_ITERS)
    b.set_tuning(TUNE_ROW_ITERS, sched["row_iters"])
    b.set_tuning(TUNE_BT_GROUP, sched["bt_group"])
    b.set_tuning(TUNE_QUEUE_MAX_CONTACTS, sched["queue_max_contacts"])
    b.set_tuning(TUNE_QUEUE_THREADS, sched["queue_threads"])
    b.set_tuning(TUNE_QUEUE_BLOCKS, sched["queue_blocks"])


def _time_batch(make_batch, sched, steps):
    import torch
    b = make_batch()
    apply_batch_schedule(b, sched)
    b.step(1)  # packing, first contacts, allocations
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.step(steps)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    b.close()
    return dt


def tune_batch(make_batch, steps=1, log=None):
    """Choose a batch schedule by measurement.

    make_batch() returns a new HipBatch holding the workload's scenes (the same scenes on every
    call).  Each candidate schedule gets a fresh batch, one untimed step, then `steps` timed
    steps.  Coordinate search: the iterations per lane with the queue off, then the work-queue shapes,
    then the B^T lane groups, each stage on the best schedule so far.  Returns {"schedule",
    "seconds_per_step", "table"} (table: one entry per candidate timed).
    """
    table = []

    def run(sched):
        t = _time_batch(make_batch, sched, steps)
        table.append(dict(sched, seconds_per_step=t))
        if log:
            log(f"[tune] {sched} -> {1e3 * t:.1f} ms per batched step")
        return t

    best = dict(BATCH_DEFAULT)
    best_t = run(best)
    for r in ROW_ITERS:
        if r == best["row_iters"]:
            continue
        cand = dict(best, row_iters=r)
        t = run(cand)
        if t < best_t:
            best, best_t = cand, t
    base = dict(best)
    for qmax, qthreads, qblocks in QUEUE_SHAPES:
        cand = dict(base, queue_max_contacts=qmax, queue_threads=qthreads, queue_blocks=qblocks)
        t = run(cand)
        if t < best_t:
            best, best_t = cand, t
    base = dict(best)
    for g in BATCH_BT_GROUPS:
        if g == base["bt_group"]:
            continue
        cand = dict(base, bt_group=g)
        t = run(cand)
        if t < best_t:
            best, best_t = cand, t
    return {"schedule": best, "seconds_per_step": best_t, "table": table}
