"""Benchmark: time-steps/s of the contact step at 131,072 spheres (APGD tol 1e-6), one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|fp32_refine] [--no-cpu]
                [--kernel-steps N] [--calibrate-copy MIB] [--tune]

Workload: 131,072 spheres, radii U[0.009, 0.011], on a
jittered 64x64x32 lattice (spacing 0.025, jitter +-0.003) in a box of 5 planes,
gravity (0,0,-9.81) through vap_fpd, m = 1, J = 0.1, dt = 1e-3, dmax = 0.02,
kd_tree_radius = 0.042^2, NoFriction APGD (dynamic descent, alpha 0.05) to
||lambda_{k+1} - lambda_k|| < 1e-6.  A step is one full pass of the device time
loop (obstacles, contact search, vap, B assembly, APGD, velocity update,
integration).  W untimed steps, then K timed steps from the state they left.

Prints one JSON line (rank 0) with the metric, a `roofline` object for the
dominant solver kernel and a `cpu_baseline` object (the CPU restatement of the
reference, timed on a bounded sample of the same workload, at 1 thread and at every
available host thread).  The K timed steps run uninstrumented; right after them,
--kernel-steps more steps run with HIP events on the library's stream around every
solver kernel (events serialise the host loop, so that pass only yields the
per-kernel average launch times): roofline.achieved = algorithmic bytes per launch
/ that average; roofline.traffic is the committed counter summary's figure for the
default FP64 workload (named in roofline.traffic_source), not a counter read in
this run.  The line also carries `other_precision` (the same workload with FP32
APGD and FP64 refinement, W + K steps from the initial state, and the ratio of the
two step rates), `warm_start` (the same workload with each solve started from the
last step's multipliers instead of the reference's lambda = 0) and `served` (a
short served run: Poisson arrivals of mixed sphere /
superellipsoid scenes, p50/p99 step latency, against the same requests stepped one
at a time), the other two parts of the headline metric.
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


# (rows per contact NR, B coefficients per contact NC, per transposed entry NT, per particle NW)
LAYOUTS = {"rotation_free": (1, 3, 3, 3), "rigid": (1, 9, 6, 6), "cone": (3, 21, 9, 6)}


def kernel_bytes(nc, nc_obs, na, s, layout="rotation_free"):
    """Algorithmic (unique) HBM bytes of one launch of each projected-gradient kernel.

    s = value bytes (8 FP64, 4 FP32); entries of the particle-major copy of B:
    2 per particle pair, 1 per obstacle contact.
      k_trial : lambda_k, lambda_{k-1}, Q lambda_{k-1}, c read and lambda+, Q lambda_k
                written per row; per contact i, j (int32) and NC coefficients (+ mu for
                the cone); per particle w_k (NW), which the row product gathers;
      k_bt    : per entry its contact index (int32) and NT coefficients; lambda+ once
                per row; per particle its entry offset (int32), M^-1 and w+ (NW each).
    """
    NR, NC, NT, NW = LAYOUTS[layout]
    rows = NR * nc
    entries = 2 * nc - nc_obs
    trial = rows * 6 * s + nc * (8 + NC * s) + (nc * s if layout == "cone" else 0) + na * NW * s
    bt = entries * (4 + NT * s) + rows * s + na * (4 + 2 * NW * s)
    return dict(bt=bt, trial=trial)


def cpu_threads():
    """Host threads this process may use: its CPU affinity, capped by OMP_NUM_THREADS when set."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(env))) if env.isdigit() else n


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(sc, prm_kw, per_step_counts, budget_s=20.0):
    """The CPU restatement (oracle/, built -O3 -march=x86-64-v3 -fopenmp) on a bounded sample of the workload.

    The reference runs its contact search, vap, update and move under OpenMP when built with it
    (contact_kdtree.hpp:220, vap_fpd.hpp:74, solver.hpp:446, 535) and its APGD serially.  Sample: one
    full contact search of the 131k-sphere state at 1 thread and at every available host thread,
    plus as many APGD iterations (reference semantics: 3 Q products per iteration) as fit in the
    budget on the real first-step problem (serial in both builds).  Scaled to steps/s with the
    per-step iteration and line-search counts measured on the GPU (the reference costs 3 Q
    products per iteration and 2 per rejected line-search trial).
    """
    import oracle
    nthreads = cpu_threads()
    t_contacts = {}
    for t in sorted({1, nthreads}):
        oracle.set_threads(t)
        t0 = time.perf_counter()
        o = oracle.Oracle(sc, **prm_kw)
        nc = o.compute_contacts()
        t_contacts[t] = time.perf_counter() - t0
    oracle.set_threads(1)
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
    step = {t: t_contacts[t] + q_per_step * t_q for t in t_contacts}

    def sample(t):
        return (f"CPU restatement of the reference, {t} thread(s) (OpenMP at the reference's pragma sites; APGD "
                f"serial): full contact search on the 131k-sphere state ({nc} contacts, {t_contacts[t]:.2f} s) + "
                f"{st['iterations']} APGD iterations ({st['q_applies']} Q products, {dt_solve:.2f} s); scaled to one "
                f"step with the GPU-measured {it:.1f} iterations + {bt:.1f} line-search rejections per step "
                f"({q_per_step:.0f} Q products at {t_q * 1e3:.1f} ms)")
    return dict(value=1.0 / step[nthreads], unit="time-steps/s", cores=nthreads, kind="port", sample=sample(nthreads),
                cpu_model=cpu_model(), build="g++ -O3 -march=x86-64-v3 -ffp-contract=off -fopenmp",
                single_thread={"value": 1.0 / step[1], "cores": 1, "sample": sample(1)},
                seconds_per_step=step[nthreads], q_product_ms=t_q * 1e3, contact_search_s=t_contacts)


TRAFFIC_PROFILE = "profiles/apgd_spheres131k_fp64_hbm_traffic.md"


def measured_traffic(kernel, precision, default_scene):
    """HBM bytes per launch of `kernel` from the committed counter summary of the default 131k-sphere
    FP64 workload (two rocprofv3 --pmc runs, FETCH_SIZE and WRITE_SIZE, calibrated on a plain copy),
    or None for another workload or precision (counters are not collected inside this run)."""
    import re
    path = os.path.join(ROOT, TRAFFIC_PROFILE)
    if precision != "fp64" or not default_scene or not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"\|\s*`" + re.escape(kernel) + r"<[^`]*`\s*\|(?:[^|]*\|){5}\s*([0-9.]+)\s*\|", line)
        if m:
            return float(m.group(1)) * 1e6
    return None


def copy_calibration(mb):
    """A plain device copy of `mb` MiB (read + write) so that a counter run (rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE) sees a kernel of known traffic beside the solver kernels."""
    import torch
    n = mb * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    del x, y


def bench_superellipsoids(args):
    """32,768 superellipsoids (32^3 jittered lattice, spacing 0.035, semi-axes U[0.008, 0.015],
    squareness U[0.4, 1.6], random orientations, seeded velocities U[-0.5, 0.5]), Friction mu 0.3,
    APGD tol 1e-6, dt 1e-3, dmax 0.02, kd_tree_radius (2 sqrt(3) 0.015 + 0.02)^2."""
    import torch
    from scopi_amd import hip, scenes
    cfg = scenes.CONFIG3
    log("[bench] building the 32k-superellipsoid scene")
    sc = scenes.superellipsoids_3d()
    precision = hip.FP64 if args.precision == "fp64" else hip.FP32_REFINE
    prm = hip.params(problem=hip.FRICTION, mu=cfg["mu"], vap=hip.VAP_FPD, dmax=cfg["dmax"],
                     kd_tree_radius=cfg["kd_tree_radius"], tolerance=cfg["tolerance"], precision=precision)
    dt = cfg["dt"]
    torch.cuda.init()
    g = hip.HipScene(sc, capacity_contacts=16 * sc.n)
    if args.row_iters is not None:
        g.set_tuning(hip.TUNE_ROW_ITERS, args.row_iters)
    if args.graph_rounds:
        g.set_tuning(hip.TUNE_GRAPH_ROUNDS, args.graph_rounds)
    log(f"[bench] warmup {args.warmup} steps")
    if args.warmup:
        g.run(prm, dt, args.warmup)
    torch.cuda.synchronize()
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    info = g.run(prm, dt, args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    result = {
        "metric": "time-steps/s at 32,768 superellipsoids, Friction mu 0.3, APGD tol 1e-6",
        "value": args.steps / elapsed, "unit": "time-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if precision == hip.FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "32,768 superellipsoids (32^3 jittered lattice, spacing 0.035), Friction mu 0.3, "
                               "APGD tol 1e-6, dt 1e-3, dmax 0.02", "precision": args.precision,
                   "graph_rounds": args.graph_rounds},
        "n_contacts_last_step": info["n_contacts"],
        "apgd_iterations_per_step": info["iterations"] / args.steps,
        "line_search_rejections_per_step": info["backtracks"] / args.steps,
    }
    print(json.dumps(result), flush=True)
    g.close()


def bench_disks(args):
    """The reference's own CPU-runnable case: 1,024 disks (2D) on a jittered 32x32 lattice above a plane,
    gravity through vap_fpd, NoFriction APGD tol 1e-6, dt 0.005, dmax 0.025, kd_tree_radius (0.024 +
    0.025)^2, 200 steps.  The GPU runs all 200 steps (after --warmup steps of a separate context), its
    rounds replayed as graphs of 63 rounds (or as --graph-rounds says); beside it, unless --no-cpu, the CPU restatement of the
    reference runs the first 25 of those steps on one host thread (about 0.8 s per step, so a bounded
    sample), and the GPU's time for the same first 25 steps is reported beside it."""
    import torch
    from scopi_amd import hip, scenes
    cfg = scenes.CONFIG1
    sc = scenes.disks_2d()
    prm = hip.params(vap=hip.VAP_FPD, dmax=cfg["dmax"], kd_tree_radius=cfg["kd_tree_radius"],
                     tolerance=cfg["tolerance"])
    torch.cuda.init()
    if args.warmup:
        w = hip.HipScene(sc)
        w.run(prm, cfg["dt"], args.warmup)
        w.close()
    g = hip.HipScene(sc)
    graph_rounds = 63 if args.graph_rounds is None else args.graph_rounds
    if graph_rounds:
        g.set_tuning(hip.TUNE_GRAPH_ROUNDS, graph_rounds)
    torch.cuda.synchronize()
    log(f"[bench] 1,024 disks: {cfg['steps']} steps on the GPU")
    sample = 25
    t0 = time.perf_counter()
    g.run(prm, cfg["dt"], sample)
    torch.cuda.synchronize()
    t_sample = time.perf_counter() - t0
    info = g.run(prm, cfg["dt"], cfg["steps"], sample)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    g.close()
    result = {
        "metric": "time-steps/s at 1,024 disks (2D), APGD tol 1e-6, 200 steps",
        "value": cfg["steps"] / elapsed, "unit": "time-steps/s", "n_gpus": 1, "steps": cfg["steps"],
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / cfg["steps"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "1,024 disks (32x32 jittered lattice, spacing 0.025) above a plane, NoFriction APGD "
                               "tol 1e-6, dt 0.005, dmax 0.025, kd_tree_radius (0.024+0.025)^2, 200 steps",
                   "graph_rounds": graph_rounds},
        "first_steps": {"steps": sample, "steps_per_s": sample / t_sample},
    }
    if not args.no_cpu:
        import oracle
        log(f"[bench] the CPU restatement: the first {sample} steps, one thread")
        oracle.set_threads(1)
        o = oracle.Oracle(sc, vap=1, dmax=cfg["dmax"], kd_tree_radius=cfg["kd_tree_radius"],
                          tolerance=cfg["tolerance"])
        t1 = time.perf_counter()
        o.run(cfg["dt"], sample)
        tc = time.perf_counter() - t1
        result["cpu_baseline"] = {"value": sample / tc, "unit": "time-steps/s", "cores": 1, "kind": "port",
                                  "sample": f"the first {sample} of the 200 steps (compare first_steps)",
                                  "cpu_model": cpu_model(),
                                  "apgd_iterations_per_step": o.stats()["total_iterations"] / sample}
    print(json.dumps(result), flush=True)


def bench_batched(args):
    """Independent sphere scenes (default 512, log-uniform 256..32,768 particles, each the 131k-sphere
    recipe scaled to its count: box of 5 planes, gravity, jittered lattice), NoFriction APGD tol 1e-6,
    dt 1e-3, dmax 0.02, kd_tree_radius 0.042^2, all scenes advanced together by one batch.  The
    one-at-a-time line steps every scene alone (its own context) for one step after the same warm-up."""
    import torch
    from scopi_amd import hip, scenes
    cfg = scenes.CONFIG4
    precision = hip.FP64 if args.precision == "fp64" else hip.FP32_REFINE
    nmin = args.nmin if args.nmin else cfg["nmin"]
    nmax = args.nmax if args.nmax else cfg["nmax"]
    log(f"[bench] building {args.scenes} scenes")
    scs = scenes.batched_sphere_scenes(args.scenes, nmin, nmax)
    prm = hip.params(vap=hip.VAP_FPD, dmax=cfg["dmax"], kd_tree_radius=cfg["kd_tree_radius"],
                     tolerance=cfg["tolerance"], precision=precision)
    dt = cfg["dt"]
    torch.cuda.init()
    tuned = None
    if not args.no_tune:
        from scopi_amd import tune

        def make_batch():
            nb = hip.HipBatch(prm, dt)
            for sc in scs:
                nb.add(sc)
            return nb
        log("[bench] autotuning the batch schedule")
        tuned = tune.tune_batch(make_batch, steps=2, log=log)
        log(f"[bench] schedule: {tuned['schedule']}")
    b = hip.HipBatch(prm, dt)
    if tuned:
        tune.apply_batch_schedule(b, tuned["schedule"])
    if args.queue_max_contacts is not None:
        b.set_tuning(hip.TUNE_QUEUE_MAX_CONTACTS, args.queue_max_contacts)
    if args.queue_threads is not None:
        b.set_tuning(hip.TUNE_QUEUE_THREADS, args.queue_threads)
    if args.queue_blocks is not None:
        b.set_tuning(hip.TUNE_QUEUE_BLOCKS, args.queue_blocks)
    if args.row_iters is not None:
        b.set_tuning(hip.TUNE_ROW_ITERS, args.row_iters)
    if args.bt_group is not None:
        b.set_tuning(hip.TUNE_BT_GROUP, args.bt_group)
    ids = [b.add(sc) for sc in scs]
    log(f"[bench] warmup {args.warmup} steps")
    b.step(args.warmup)
    torch.cuda.synchronize()
    log(f"[bench] timing {args.steps} batched steps")
    t0 = time.perf_counter()
    b.step(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    last = [b.info(sid) for sid in ids]  # the last step's solve of every scene
    contacts = sum(i["n_contacts"] for i in last)
    citer = sum(i["n_contacts"] * i["iterations"] for i in last)
    size = b.size()
    b.close()
    one = None
    if args.one_at_a_time:
        log("[bench] one-at-a-time: every scene alone, one step after the warm-up")
        t_one = 0.0
        for sc in scs:
            g = hip.HipScene(sc)
            if args.warmup:
                g.run(prm, dt, args.warmup)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            g.run(prm, dt, 1)
            torch.cuda.synchronize()
            t_one += time.perf_counter() - t1
            g.close()
        one = {"scene_steps_per_s": len(scs) / t_one, "ms_per_round_of_all_scenes": 1e3 * t_one,
               "sample": f"all {len(scs)} scenes, one step each after {args.warmup} warm-up steps"}
    value = len(scs) * args.steps / elapsed
    result = {
        "metric": f"scene time-steps/s, {len(scs)} independent sphere scenes (log-uniform {nmin}.."
                  f"{nmax} particles), APGD tol 1e-6",
        "value": value, "unit": "scene-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if precision == hip.FP64 else "f32+f64", "data": "synthetic",
        "config": {"workload": f"{len(scs)} independent sphere scenes, log-uniform {nmin}..{nmax} "
                               "particles each (131k-sphere recipe scaled), 5 planes each, NoFriction, dt 1e-3, "
                               "dmax 0.02, kd_tree_radius 0.042^2", "precision": args.precision,
                   "particles": size["particles"], "queue_max_contacts": args.queue_max_contacts,
                   "queue_threads": args.queue_threads},
        "contacts_last_step": contacts,
        "apgd_iterations_last_step": {"mean": sum(i["iterations"] for i in last) / len(last),
                                      "max": max(i["iterations"] for i in last)},
        "contact_iterations_per_s": citer * args.steps / elapsed,
        "all_converged": all(i["converged"] == 1 for i in last),
        "one_at_a_time": one,
        "speedup_vs_one_at_a_time": (value / one["scene_steps_per_s"]) if one else None,
        "schedule": tuned["schedule"] if tuned else None,
        "autotuner": ({"candidates": tuned["table"], "timed_steps_per_candidate": 2} if tuned else None),
    }
    print(json.dumps(result), flush=True)


def served_leg(requests, rate, request_steps, sample_steps, hbm_budget_gib, one_at_a_time, warm_start=True,
               schedules=None):
    """Served scenes: Poisson arrivals (`rate` per s) of `requests` scene requests, 80% sphere scenes (the
    131k-sphere recipe scaled, NoFriction) and 20% superellipsoid scenes (the 32k recipe scaled, Friction
    mu 0.3), log-uniform 512..65,536 particles, `request_steps` steps each, admitted by estimated HBM
    footprint into one resident batch per kind (scopi_amd.serve).  The one-at-a-time figure steps every
    request's scene alone (its own context) for `sample_steps` timed steps after one warm-up step and
    scales each to its step count.  With warm_start (the default), every scene's solves, served and
    alone, start from its own last step's multipliers, kept on the device."""
    import torch
    from scopi_amd import hip, serve
    kinds, footprint, prm = serve.hip_kinds(warm_start=warm_start, schedules=schedules)
    reqs = serve.poisson_requests(requests, rate, [("spheres", 0.8), ("superellipsoids", 0.2)], (512, 65536),
                                  request_steps, seed=2025)
    log(f"[bench] serving {len(reqs)} requests ({sum(r.kind == 'superellipsoids' for r in reqs)} superellipsoid "
        f"scenes), {request_steps} steps each, rate {rate}/s")
    done = lambda r: log(f"[bench] request {r.rid} ({r.kind}, {r.n} particles) done at {r.finished:.2f} s")
    srv = serve.Server(kinds, serve.served_scene, footprint, hbm_budget=int(hbm_budget_gib * (1 << 30)),
                       on_finish=done)
    st = srv.run(reqs)
    log(f"[bench] served: {json.dumps(st)}")
    one = None
    if one_at_a_time:
        log("[bench] one-at-a-time sample")
        t_total = 0.0
        for r in reqs:
            g = hip.HipScene(serve.served_scene(r))
            g.run(prm[r.kind], 1e-3, 1)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            k = min(sample_steps, r.steps)
            g.run(prm[r.kind], 1e-3, k)
            torch.cuda.synchronize()
            t_total += (time.perf_counter() - t1) / k * r.steps
            g.close()
        one = {"scene_steps_per_s": st["scene_steps"] / t_total, "seconds_for_all_requests": t_total,
               "sample": f"each request alone: {sample_steps} timed steps after 1 warm-up step, scaled to its "
                         f"{request_steps} steps"}
    return {
        "scene_steps_per_s": st["scene_steps_per_s"], "scene_steps": st["scene_steps"],
        "step_latency_ms": st["step_latency_ms"], "wait_ms": st["wait_ms"], "completion_s": st["completion_s"],
        "peak_resident_scenes": st["peak_resident_scenes"], "warm_start": bool(warm_start),
        "workload": f"{len(reqs)} requests, Poisson rate {rate}/s, {request_steps} steps each, sizes log-uniform "
                    f"512..65,536, spheres NoFriction / superellipsoids Friction mu 0.3, HBM budget {hbm_budget_gib} GiB",
        "one_at_a_time": one,
        "speedup_vs_one_at_a_time": (st["scene_steps_per_s"] / one["scene_steps_per_s"]) if one else None,
    }


def bench_served(args):
    """The served form alone as the JSON line (see served_leg)."""
    import torch
    torch.cuda.init()
    schedules = None
    if args.served_schedule:
        schedules = {k: json.loads(v) for k, v in (item.split("=", 1) for item in args.served_schedule)}
    sv = served_leg(args.requests, args.rate, args.request_steps, args.sample_steps, args.hbm_budget_gib,
                    args.one_at_a_time, warm_start=not args.served_cold, schedules=schedules)
    result = {
        "metric": "served scene time-steps/s (Poisson arrivals, 80% sphere / 20% superellipsoid scenes, "
                  "512..65,536 particles) with p50/p99 step latency",
        "value": sv["scene_steps_per_s"], "unit": "scene-steps/s", "n_gpus": 1, "steps": sv["scene_steps"],
        "warmup": 0, "ms_per_step": sv["step_latency_ms"]["p50"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic", "config": {"workload": sv["workload"]},
        "step_latency_ms": sv["step_latency_ms"], "wait_ms": sv["wait_ms"], "completion_s": sv["completion_s"],
        "peak_resident_scenes": sv["peak_resident_scenes"], "one_at_a_time": sv["one_at_a_time"],
        "speedup_vs_one_at_a_time": sv["speedup_vs_one_at_a_time"],
    }
    print(json.dumps(result), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", choices=["fp64", "fp32_refine"], default="fp64")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32-leg", action="store_true",
                    help="spheres: skip the FP32 + FP64-refinement pass at the same size (the ratio to FP64)")
    ap.add_argument("--no-served-leg", action="store_true", help="spheres: skip the short served leg")
    ap.add_argument("--no-warm-leg", action="store_true", help="spheres: skip the warm-start pass")
    ap.add_argument("--served-requests", type=int, default=12, help="spheres: requests of the served leg")
    ap.add_argument("--served-request-steps", type=int, default=40, help="spheres: steps per served request")
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--ny", type=int, default=64)
    ap.add_argument("--nz", type=int, default=32)
    ap.add_argument("--tune", action="store_true", help="sweep the solver's launch shapes before timing")
    ap.add_argument("--graph-rounds", type=int, default=None,
                    help="replay the projected-gradient rounds as a hipGraph of this many rounds (odd; 0: eager); "
                         "default 63 for the disks workload (each launch holds 8 to 16 blocks there: 88 against 83 "
                         "steps/s eager), eager for the others (131k spheres, superellipsoids: no gain)")
    ap.add_argument("--kernel-steps", type=int, default=3,
                    help="steps of the instrumented pass (HIP events around every solver kernel)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="run one plain device copy of MIB MiB first (traffic calibration for counter runs)")
    ap.add_argument("--workload", choices=["spheres", "superellipsoids", "batched", "served", "disks"],
                    default="spheres",
                    help="spheres: the headline 131k-sphere step; disks: 1,024 disks, 200 steps; "
                         "superellipsoids: 32,768 superellipsoids, "
                         "Friction mu 0.3; batched: independent sphere scenes advanced together (secondary lines)")
    ap.add_argument("--scenes", type=int, default=512, help="batched: number of scenes")
    ap.add_argument("--one-at-a-time", action="store_true", help="batched / served: also step every scene alone")
    ap.add_argument("--queue-max-contacts", type=int, default=None,
                    help="batched: scenes of at most this many contacts go to the per-scene work-queue kernel")
    ap.add_argument("--queue-threads", type=int, default=None, help="batched: work-queue workgroup size")
    ap.add_argument("--queue-blocks", type=int, default=None, help="batched: work-queue workgroups")
    ap.add_argument("--row-iters", type=int, default=None,
                    help="spheres / batched / superellipsoids: loop iterations per lane of k_trial")
    ap.add_argument("--bt-group", type=int, default=None, help="spheres / batched: lanes per particle of k_bt")
    ap.add_argument("--reduce-blocks", type=int, default=None, help="spheres: block cap of the per-contact grid")
    ap.add_argument("--nmin", type=int, default=0, help="batched: smallest scene (default 256 particles)")
    ap.add_argument("--nmax", type=int, default=0, help="batched: largest scene (default 32,768 particles)")
    ap.add_argument("--no-tune", action="store_true",
                    help="batched: skip the autotuner (scopi_amd.tune.tune_batch) and run the library's defaults")
    ap.add_argument("--served-cold", action="store_true",
                    help="served: start every solve from lambda = 0 instead of the scene's last multipliers")
    ap.add_argument("--requests", type=int, default=24, help="served: number of requests")
    ap.add_argument("--served-schedule", action="append", default=[], metavar="KIND=JSON",
                    help="served: batch schedule of one request kind, e.g. spheres='{\"row_iters\": 4, ...}' "
                         "(keys of scopi_amd.tune.BATCH_DEFAULT)")
    ap.add_argument("--rate", type=float, default=4.0, help="served: mean arrivals per second")
    ap.add_argument("--request-steps", type=int, default=100, help="served: steps asked by each request")
    ap.add_argument("--sample-steps", type=int, default=2, help="served: timed steps per request alone")
    ap.add_argument("--hbm-budget-gib", type=float, default=96.0, help="served: device memory for resident scenes")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the contact step is not sharded across GPUs.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from scopi_amd import hip, scenes

    if args.workload == "superellipsoids":
        return bench_superellipsoids(args)
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "served":
        return bench_served(args)
    if args.workload == "disks":
        return bench_disks(args)
    log("[bench] building the 131k-sphere scene")
    sc = scenes.spheres_3d(args.nx, args.ny, args.nz)
    default_scene = (args.nx, args.ny, args.nz) == (64, 64, 32)
    precision = hip.FP64 if args.precision == "fp64" else hip.FP32_REFINE
    prm = hip.params(vap=hip.VAP_FPD, dmax=0.02, kd_tree_radius=0.042 ** 2, tolerance=1e-6, precision=precision)
    dt = 1e-3
    torch.cuda.init()
    if args.calibrate_copy:
        copy_calibration(args.calibrate_copy)
    g = hip.HipScene(sc, capacity_contacts=20 * sc.n)
    if args.graph_rounds:
        g.set_tuning(hip.TUNE_GRAPH_ROUNDS, args.graph_rounds)
    if args.bt_group is not None:
        g.set_tuning(hip.TUNE_BT_GROUP, args.bt_group)
    if args.reduce_blocks is not None:
        g.set_tuning(hip.TUNE_REDUCE_BLOCKS, args.reduce_blocks)
    if args.row_iters is not None:
        g.set_tuning(hip.TUNE_ROW_ITERS, args.row_iters)
    log(f"[bench] warmup {args.warmup} steps")
    if args.warmup:
        g.run(prm, dt, args.warmup)
    if args.tune:
        from scopi_amd import tune
        g.compute_contacts(prm)
        g.set_a_priori_velocity(prm, dt)
        tuned = tune.tune(g, prm, dt)
        log(f"[bench] tuned launch shapes: {tuned}")
    torch.cuda.synchronize()
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    info = g.run(prm, dt, args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    steps_per_s = args.steps / elapsed
    nc = info["n_contacts"]
    na = sc.n_active
    it_per_step = info["iterations"] / args.steps
    bt_per_step = info["backtracks"] / args.steps
    it32 = info["iterations_fp32"] / max(1, info["iterations"])

    # Instrumented pass: HIP events on the library's stream around every solver kernel
    # (they serialise the host loop, so this pass is not the throughput figure).
    per, kt_steps_per_s = {}, None
    if args.kernel_steps > 0:
        log(f"[bench] instrumented pass, {args.kernel_steps} steps")
        g.set_kernel_timing(True)
        t1 = time.perf_counter()
        info_k = g.run(prm, dt, args.kernel_steps)
        torch.cuda.synchronize()
        kt_steps_per_s = args.kernel_steps / (time.perf_counter() - t1)
        kt = g.kernel_timing()
        g.set_kernel_timing(False)
        nc = info_k["n_contacts"]
        it32 = info_k["iterations_fp32"] / max(1, info_k["iterations"])
        for name, ms_key, n_key in (("bt", "bt_ms", "bt_launches"), ("trial", "trial_ms", "trial_launches")):
            n = kt[n_key]
            per[name] = dict(avg_us=(kt[ms_key] / n * 1e3) if n else None, launches=n, total_ms=kt[ms_key])
    ct = g.contacts() if nc else dict(ij=np.zeros((0, 2), np.int64))
    nc_obs = int(np.count_nonzero(ct["ij"][:, 0] < sc.n_inactive))
    # FP32 phase launches move s = 4 bytes per value, FP64 ones s = 8 (weighted by iterations)
    kb64 = kernel_bytes(nc, nc_obs, na, 8)
    kb32 = kernel_bytes(nc, nc_obs, na, 4)
    kb = {k: it32 * kb32[k] + (1 - it32) * kb64[k] for k in kb64}
    roofline = None
    if per:
        dom = max(per, key=lambda k: per[k]["total_ms"])
        avg_s = per[dom]["avg_us"] * 1e-6 if per[dom]["avg_us"] else None
        achieved = (kb[dom] / avg_s / 1e9) if avg_s else None
        for k in per:
            a = per[k]["avg_us"]
            per[k]["algorithmic_bytes"] = kb[k]
            per[k]["achieved_gbs"] = kb[k] / (a * 1e-6) / 1e9 if a else None
        roofline = {"bound": "hbm", "kernel": {"bt": "k_bt", "trial": "k_trial"}[dom],
                    "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                    "traffic": measured_traffic({"bt": "k_bt", "trial": "k_trial"}[dom], args.precision,
                                                default_scene),
                    "traffic_source": (TRAFFIC_PROFILE + " (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; not measured in "
                                       "this run)") if measured_traffic({"bt": "k_bt", "trial": "k_trial"}[dom],
                                                                        args.precision, default_scene) else None,
                    "algorithmic_bytes_per_launch": kb[dom], "avg_launch_us": per[dom]["avg_us"]}
    log(f"[bench] {steps_per_s:.3f} steps/s, nc={nc}, iterations/step={it_per_step:.1f}, "
        f"backtracks/step={bt_per_step:.1f}, kernels={per}")
    # the other precision on the same workload: FP32 APGD with FP64 refinement against the FP64 step
    other = None
    if not args.no_fp32_leg:
        oprec = hip.FP32_REFINE if precision == hip.FP64 else hip.FP64
        oname = "fp32_refine" if oprec == hip.FP32_REFINE else "fp64"
        log(f"[bench] {oname} pass: warmup {args.warmup}, timing {args.steps} steps")
        oprm = hip.params(vap=hip.VAP_FPD, dmax=0.02, kd_tree_radius=0.042 ** 2, tolerance=1e-6, precision=oprec)
        g2 = hip.HipScene(sc, capacity_contacts=20 * sc.n)
        if args.warmup:
            g2.run(oprm, dt, args.warmup)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        info2 = g2.run(oprm, dt, args.steps)
        torch.cuda.synchronize()
        el2 = time.perf_counter() - t2
        g2.close()
        sps2 = args.steps / el2
        fp32_rate, fp64_rate = (sps2, steps_per_s) if oprec == hip.FP32_REFINE else (steps_per_s, sps2)
        other = {"precision": oname, "steps_per_s": sps2, "ms_per_step": 1e3 * el2 / args.steps,
                 "apgd_iterations_per_step": info2["iterations"] / args.steps,
                 "fp32_iterations_per_step": info2["iterations_fp32"] / args.steps,
                 "residual_last_solve": info2["residual"], "tolerance": 1e-6,
                 "fp32_refine_over_fp64": fp32_rate / fp64_rate}
        log(f"[bench] {oname}: {sps2:.3f} steps/s, FP32+FP64 / FP64 = {fp32_rate / fp64_rate:.3f}")
    # warm start (lambda of the last step as the first iterate; not the reference's lambda = 0)
    warm = None
    if not args.no_warm_leg:
        log(f"[bench] warm-start pass ({args.precision}): warmup {args.warmup}, timing {args.steps} steps")
        wprm = hip.params(vap=hip.VAP_FPD, dmax=0.02, kd_tree_radius=0.042 ** 2, tolerance=1e-6, precision=precision,
                          warm_start=1)
        g3 = hip.HipScene(sc, capacity_contacts=20 * sc.n)
        if args.warmup:
            g3.run(wprm, dt, args.warmup)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        info3 = g3.run(wprm, dt, args.steps)
        torch.cuda.synchronize()
        el3 = time.perf_counter() - t3
        g3.close()
        warm = {"precision": args.precision, "steps_per_s": args.steps / el3, "ms_per_step": 1e3 * el3 / args.steps,
                "apgd_iterations_per_step": info3["iterations"] / args.steps,
                "line_search_rejections_per_step": info3["backtracks"] / args.steps,
                "over_cold_start": (args.steps / el3) / steps_per_s}
        log(f"[bench] warm start: {args.steps / el3:.3f} steps/s, {info3['iterations'] / args.steps:.1f} iterations/step")
    served = None
    if not args.no_served_leg:
        served = served_leg(args.served_requests, args.rate, args.served_request_steps, args.sample_steps,
                            args.hbm_budget_gib, True, warm_start=not args.served_cold)
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
                   "n_spheres": na, "n_contacts_last_step": nc, "precision": args.precision,
                   "graph_rounds": args.graph_rounds},
        "contact_iterations_per_s": nc * info["iterations"] / elapsed,
        "apgd_iterations_per_step": it_per_step,
        "line_search_rejections_per_step": bt_per_step,
        "kernels": per,
        "instrumented_steps_per_s": kt_steps_per_s,
        "roofline": roofline,
        "other_precision": other,
        "warm_start": warm,
        "served": served,
    }
    if not args.no_cpu:
        log("[bench] CPU baseline (bounded sample)")
        prm_kw = dict(vap=1, dmax=0.02, kd_tree_radius=0.042 ** 2, tolerance=1e-6)
        cb = cpu_baseline(sc, prm_kw, (it_per_step, bt_per_step))
        result["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample", "cpu_model", "build",
                                                     "single_thread")}
    print(json.dumps(result), flush=True)
    g.close()


if __name__ == "__main__":
    main()
