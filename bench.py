"""Benchmark: time-steps/s of the contact step at 131,072 spheres (APGD tol 1e-6), one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|fp32_refine] [--no-cpu]

Workload: 131,072 spheres, radii U[0.009, 0.011], on a
jittered 64x64x32 lattice (spacing 0.025, jitter +-0.003) in a box of 5 planes,
gravity (0,0,-9.81) through vap_fpd, m = 1, J = 0.1, dt = 1e-3, dmax = 0.02,
kd_tree_radius = 0.042^2, NoFriction APGD (dynamic descent, alpha 0.05) to
||lambda_{k+1} - lambda_k|| < 1e-6.  A step is one full pass of the device time
loop (obstacles, contact search, vap, B assembly, APGD, velocity update,
integration).  W untimed steps, then K timed steps from the state they left.

Prints one JSON line (rank 0) with the metric, a `roofline` object for the
dominant solver kernel (algorithmic bytes per launch / average launch time from
HIP events on the library's stream) and a `cpu_baseline` object (the CPU
restatement of the reference, timed on a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(nc, na, s, friction=False):
    """Algorithmic (unique) HBM bytes per launch of each projected-gradient kernel.

    s = value bytes (8 FP64, 4 FP32).  NoFriction: 9 coefficients per contact
    (n, a_i, a_j), int32 particle indices, one row per contact.
      k_bt    : lambda, 9 coefficients and the j-list entry per contact, 4 range
                ints + M^-1 (6) + w (6) per particle;
      k_row   : i, j, 9 coefficients, lambda+ read, Q lambda+ write per contact,
                w (6) per particle;
      k_trial : lambda_k, lambda_{k-1}, Q lambda_k, Q lambda_{k-1}, c read and
                lambda+ write per row.
    """
    if friction:
        rows = 3 * nc
        bt = nc * (3 * s + 18 * s + 4) + na * (16 + 12 * s)
        row = nc * (8 + 18 * s + 6 * s) + na * 6 * s
        trial = rows * 6 * s + nc * (4 * s)
    else:
        bt = nc * (s + 9 * s + 4) + na * (16 + 12 * s)
        row = nc * (8 + 9 * s + 2 * s) + na * 6 * s
        trial = nc * 6 * s
    return dict(bt=bt, row=row, trial=trial)


def cpu_baseline(sc, prm_kw, per_step_counts, budget_s=20.0):
    """The CPU restatement (oracle/, single thread) on a bounded sample of the workload.

    Sample: one full contact search on the 131k-sphere state plus as many APGD
    iterations (reference semantics: 3 Q products per iteration) as fit in the
    budget, on the real first-step problem.  Scaled to steps/s with the per-step
    iteration and line-search counts measured on the GPU (the reference costs
    3 Q products per iteration and 2 per rejected line-search trial).
    """
    import oracle
    t0 = time.perf_counter()
    o = oracle.Oracle(sc, **prm_kw)
    nc = o.compute_contacts()
    t_contacts = time.perf_counter() - t0
    o.set_a_priori_velocity(1e-3)
    # time APGD iterations on the real problem with a capped iteration count
    k = 2
    while True:
        o.set_params(**dict(prm_kw, max_ite=k))
        t1 = time.perf_counter()
        o.solve(1e-3)
        dt_solve = time.perf_counter() - t1
        st = o.stats()
        if dt_solve > budget_s / 4 or k >= 4096 or st["converged"]:
            break
        k *= 2
    t_q = dt_solve / max(1, st["q_applies"])
    it, bt = per_step_counts
    q_per_step = 3 * it + 2 * bt
    t_step = t_contacts + q_per_step * t_q
    return dict(value=1.0 / t_step, unit="time-steps/s", cores=1, kind="port",
                sample=(f"CPU restatement, 1 thread: full contact search on the 131k-sphere state "
                        f"({nc} contacts, {t_contacts:.2f} s) + {st['iterations']} APGD iterations "
                        f"({st['q_applies']} Q products, {dt_solve:.2f} s); scaled to one step with the "
                        f"GPU-measured {it:.1f} iterations + {bt:.1f} line-search rejections per step "
                        f"({q_per_step:.0f} Q products at {t_q * 1e3:.1f} ms)"),
                seconds_per_step=t_step, q_product_ms=t_q * 1e3, contact_search_s=t_contacts)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", choices=["fp64", "fp32_refine"], default="fp64")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--ny", type=int, default=64)
    ap.add_argument("--nz", type=int, default=32)
    ap.add_argument("--tune", action="store_true", help="sweep the solver's launch shapes before timing")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the contact step of one scene is not sharded across GPUs "
              "in this build (see DESIGN.md, 'More than one GPU').", file=sys.stderr)
        sys.exit(2)

    import torch
    from scopi_amd import hip, scenes

    log("[bench] building the 131k-sphere scene")
    sc = scenes.spheres_3d(args.nx, args.ny, args.nz)
    precision = hip.FP64 if args.precision == "fp64" else hip.FP32_REFINE
    prm = hip.params(vap=hip.VAP_FPD, dmax=0.02, kd_tree_radius=0.042 ** 2, tolerance=1e-6, precision=precision)
    dt = 1e-3
    torch.cuda.init()
    g = hip.HipScene(sc, capacity_contacts=20 * sc.n)
    log(f"[bench] warmup {args.warmup} steps")
    if args.warmup:
        g.run(prm, dt, args.warmup)
    if args.tune:
        from scopi_amd import tune
        g.compute_contacts(prm)
        g.set_a_priori_velocity(prm, dt)
        tuned = tune.tune(g, prm, dt)
        log(f"[bench] tuned launch shapes: {tuned}")
    g.set_kernel_timing(True)
    torch.cuda.synchronize()
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    info = g.run(prm, dt, args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kt = g.kernel_timing()
    steps_per_s = args.steps / elapsed
    nc = info["n_contacts"]
    na = sc.n_active
    it_per_step = info["iterations"] / args.steps
    bt_per_step = info["backtracks"] / args.steps
    s = 8 if precision == hip.FP64 else 4
    kb = kernel_bytes(nc, na, 8)
    per = {}
    for name, ms_key, n_key in (("bt", "bt_ms", "bt_launches"), ("row", "row_ms", "row_launches"),
                                ("trial", "trial_ms", "trial_launches")):
        n = kt[n_key]
        per[name] = dict(avg_us=(kt[ms_key] / n * 1e3) if n else None, launches=n, total_ms=kt[ms_key])
    dom = max(per, key=lambda k: per[k]["total_ms"])
    kname = {"bt": "k_bt", "row": "k_row", "trial": "k_trial"}[dom]
    avg_s = per[dom]["avg_us"] * 1e-6 if per[dom]["avg_us"] else None
    achieved = (kb[dom] / avg_s / 1e9) if avg_s else None
    log(f"[bench] {steps_per_s:.3f} steps/s, nc={nc}, iterations/step={it_per_step:.1f}, "
        f"backtracks/step={bt_per_step:.1f}, kernels={per}")
    result = {
        "metric": "time-steps/s at 131k spheres, APGD tol 1e-6",
        "value": steps_per_s,
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision == hip.FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{na} spheres ({args.nx}x{args.ny}x{args.nz} jittered lattice) + 5 planes, "
                               "NoFriction APGD tol 1e-6, dt 1e-3, dmax 0.02, kd_tree_radius 0.042^2",
                   "n_spheres": na, "n_contacts_last_step": nc, "precision": args.precision},
        "contact_iterations_per_s": nc * info["iterations"] / elapsed,
        "apgd_iterations_per_step": it_per_step,
        "line_search_rejections_per_step": bt_per_step,
        "kernels": per,
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": None,
                     "algorithmic_bytes_per_launch": kb[dom]},
    }
    if precision != hip.FP64:
        result["roofline"]["note"] = "bytes counted at FP64 values; FP32-phase launches move half of it"
    if not args.no_cpu:
        log("[bench] CPU baseline (bounded sample)")
        prm_kw = dict(vap=1, dmax=0.02, kd_tree_radius=0.042 ** 2, tolerance=1e-6)
        cb = cpu_baseline(sc, prm_kw, (it_per_step, bt_per_step))
        result["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
    print(json.dumps(result), flush=True)
    g.close()


if __name__ == "__main__":
    main()
