"""Benchmark of the MPH-WC step on one MI355X.

Workload: the 3D two-fluid dam break at ~1M fluid particles (128x64x128 fluid particles,
type 0 below / type 1 above mid-height, in a 384x64x192 box with 3-layer walls on the x and
z faces, y periodic; 1,272,064 particles), FP64.  A step is one full iteration of the
reference's loop body (source/main.cpp:386-431) for every particle, rebuilds included.

`value` is timed over --steps steps from step --from-step (1000: the developed flow, where
the lists are rebuilt every few steps as they are for most of a run); the first steps of
the flow (from rest, lists rebuilt every 20-50 steps) are reported as `start_of_flow`.

Prints ONE JSON line: steps/s (value), pair-interactions/s, the roofline of the dominant
kernel from HIP events on the solver's stream, and the reference's own OpenMP build timed
on a bounded sample of the same deck on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic FP64 flops per interacting pair (ordered, r <= max radius), counting
# add/sub/mul/div/sqrt as 1 and fma as 2, min-image free: sweep A / sweep B
FLOPS_PER_PAIR = {False: (27, 49), True: (43, 117)}   # keyed by "surface tension present"
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
ERROR_FIELDS = ("position", "velocity", "force", "density_a", "vol_strain_p", "pressure_p")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_deck(name: str):
    from mphexplicit_amd import decks
    if name == "dambreak3d_1m":
        return decks.dam_break_1m()
    if name == "dambreak3d":
        return decks.dam_break_3d()
    if name == "droplet2d":
        return decks.droplet_2d()
    raise SystemExit("unknown workload %s" % name)


def cpu_baseline(deck, steps: int, cores: int):
    """Reference build (oracle/_ref) on the host: load + 1 warm step untimed, then `steps`."""
    from oracle.oracle import Reference, reference_available
    if not reference_available(deck.dim):
        return None
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    t0 = time.perf_counter()
    ref = Reference(deck, nthreads=cores)      # deck files go to a temporary directory
    ref.step(1)
    t1 = time.perf_counter()
    rb = ref.step(steps)
    t2 = time.perf_counter()
    log("cpu baseline: load+warm %.1fs, %d steps %.1fs, rebuilds %d" % (t1 - t0, steps, t2 - t1, rb))
    return {"value": steps / (t2 - t1), "unit": "steps/s", "cores": cores, "kind": "reference",
            "sample": "steps 2-%d of the same deck (%d particles), reference source/main.cpp built with "
                      "g++ -O3 -fopenmp, %d OpenMP threads, %d rebuilds in the sample"
                      % (steps + 1, deck.n, cores, rb)}


def measured_traffic(workload: str, cluster_size: int, kernel: str):
    """HBM bytes per launch of a sweep from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE runs
    (profiles/hbm_traffic_<workload>.json), if they were taken on the same configuration"""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_%s.json" % workload)
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("workload") != workload or t.get("cluster_size") != cluster_size:
        return None
    k = t.get("kernels", {}).get(kernel)
    return float(k["bytes"]) if k else None


def timed_steps(s, args, n):
    """n steps timed by the wall clock around the stream (synchronised on both sides)"""
    stepper = s.step_graph if args.graph else s.step
    s.sync()
    t0 = time.perf_counter()
    stepper(n)
    s.sync()
    return n / (time.perf_counter() - t0)


def timed_run(s, args):
    """warmup, `start_of_flow` steps, advance to --from-step, the timed window (value), then a
    HIP-event window of the same developed flow"""
    stepper = s.step_graph if args.graph else s.step
    stepper(args.warmup)
    start = {"from_step": args.warmup, "steps": args.steps, "value": timed_steps(s, args, args.steps),
             "unit": "steps/s"}
    cur = s.stats()["steps"]
    if cur < args.from_step:
        stepper(args.from_step - cur)
    rb0 = s.stats()["rebuilds"]
    from_step = s.stats()["steps"]
    sps = timed_steps(s, args, args.steps)
    rebuilds = s.stats()["rebuilds"] - rb0
    k2 = max(10, args.steps // 4)
    tm = s.step_timed(k2)
    return sps, tm, k2, {"start_of_flow": start, "from_step": from_step, "rebuilds": rebuilds}


def field_errors(a, b):
    """per field: normwise max|a-b|/max|b| and relative L2 |a-b|/|b|"""
    import numpy as np
    out = {}
    for k in ERROR_FIELDS:
        x, y = a[k].astype(np.float64), b[k].astype(np.float64)
        d = x - y
        out[k] = {"normwise": float(np.max(np.abs(d)) / max(np.max(np.abs(y)), 1e-300)),
                  "rel_l2": float(np.linalg.norm(d) / max(np.linalg.norm(y), 1e-300))}
    return out


def batch_leg(args, Solver):
    """64 independent uneven cases (2D/3D droplets and bridges, 2e3..6e4 particles each) in
    one context: every kernel launch covers all cases, steps are replayed HIP graphs"""
    from mphexplicit_amd import decks
    cases = decks.batch_cases(count=64)
    n = [d.n for d in cases]
    t0 = time.perf_counter()
    b = Solver.from_decks(cases)
    create = time.perf_counter() - t0
    b.step_graph(args.warmup)
    b.sync()
    t0 = time.perf_counter()
    b.step_graph(args.steps)
    b.sync()
    el = time.perf_counter() - t0
    k2 = max(10, args.steps // 4)
    tm = b.step_timed(k2)
    st = b.stats()
    b.close()
    sps = args.steps / el
    return {"cases": len(cases), "particles": int(sum(n)), "min_case": int(min(n)), "max_case": int(max(n)),
            "value": sps, "unit": "steps/s (all cases advance one step)", "ms_per_step": 1e3 / sps,
            "particle_steps_per_s": sps * sum(n), "pair_interactions_per_s": tm["pairs"] / k2 * sps,
            "create_s": create, "cluster_size": st["cluster_size"], "graph": True,
            "phase_ms": {"move": tm["move_ms"] / k2, "rebuild": tm["rebuild_ms"] / k2,
                         "sweep_a": tm["sweep_a_ms"] / k2, "sweep_b": tm["sweep_b_ms"] / k2}}


SUR_SCALES = dict(u_scale=0.1, a_scale=10.0, dens_scale=50.0)


def pretrain_decks():
    """The other trajectories the network is trained on before the measured deck's own: the
    2D square droplet of 80 x 80 and the 3D dam break of 64 x 32 x 64 (plain FP64 runs)."""
    from mphexplicit_amd import decks
    return [("droplet2d_80", decks.droplet_2d()), ("dambreak3d_64", decks.dam_break_3d())]


def pretrain(args):
    """Weights trained on each pre-training deck's plain trajectory in turn (from step
    --sur-warm, --sur-samples samples x --sur-iters Adam steps each; the Adam moments start
    afresh on each deck), as the measured deck's training then continues from them."""
    from mphexplicit_amd.solver import Solver, Surrogate
    w, info = None, []
    for name, d in pretrain_decks():
        t0 = time.perf_counter()
        s = Solver.from_deck(d)
        s.step(args.sur_warm)
        sur = Surrogate(s, train_stride=args.sur_stride, lr=1e-3, seed=5, radius_ratio=args.sur_radius,
                        **SUR_SCALES)
        if w is not None:
            sur.set_weights(w)
        losses = sur.train(args.sur_samples, args.sur_iters)
        w = sur.weights()
        sur.close()
        s.close()
        info.append({"deck": name, "particles": d.n, "loss_first": float(losses[0]),
                     "loss_last": float(losses[-1]), "s": time.perf_counter() - t0})
        log("pre-training on %s: loss %.3g -> %.3g" % (name, losses[0], losses[-1]))
    return w, info


def surrogate_leg(args, deck):
    """Learned 2 dt steps on the same deck: train on the deck's own plain trajectory (from a
    developed flow; with pre-training, continuing from the weights trained on the 2D droplet
    and the 64 x 32 x 64 dam break), then time plain steps, bare surrogate steps and judged
    blocks (K surrogate steps, one-step judge against two plain steps, rollback above the
    bound) in the same stretch of the flow."""
    from mphexplicit_amd.solver import Solver, Surrogate
    w0, pre = pretrain(args) if args.sur_pretrain else (None, [])
    s = Solver.from_deck(deck)
    s.step(args.sur_warm)
    sur = Surrogate(s, train_stride=args.sur_stride, lr=1e-3, seed=5, radius_ratio=args.sur_radius, **SUR_SCALES)
    if w0 is not None:
        sur.set_weights(w0)
    s.sync()
    t0 = time.perf_counter()
    losses = sur.train(args.sur_samples, args.sur_iters)
    t_train = time.perf_counter() - t0
    plain_sps = timed_steps(s, args, 2 * args.sur_steps)
    sur.step(2)
    s.sync()
    t0 = time.perf_counter()
    sur.step(args.sur_steps)
    s.sync()
    t_bare = (time.perf_counter() - t0) / args.sur_steps
    t0 = time.perf_counter()
    st = sur.advance(args.sur_blocks, args.sur_k, args.sur_tol)
    t_judged = time.perf_counter() - t0
    dt = deck.data.dt
    plain_sim = dt * plain_sps
    out = {"train_s": t_train, "train_samples": args.sur_samples, "train_iters": args.sur_iters,
           "train_stride": args.sur_stride, "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
           "ms_per_surrogate_step": 1e3 * t_bare, "ms_per_two_plain_steps": 2e3 / plain_sps,
           "bare_sim_seconds_per_s": 2.0 * dt / t_bare,
           "judged": st, "judged_sim_seconds_per_s": st["sim_time"] / t_judged,
           "plain_sim_seconds_per_s": plain_sim, "tol": args.sur_tol, "k": args.sur_k,
           "radius_ratio": args.sur_radius or deck.data.radius_ratio_p, "pretrained_on": pre}
    out["judged_vs_plain"] = out["judged_sim_seconds_per_s"] / plain_sim
    out["bare_vs_plain"] = out["bare_sim_seconds_per_s"] / plain_sim
    sur.close()
    s.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="dambreak3d_1m")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--graph", action="store_true", help="replay a captured HIP graph per step")
    ap.add_argument("--from-step", type=int, default=1000,
                    help="step from which the timed window starts (the developed flow)")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-case batch leg")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned-step leg")
    ap.add_argument("--sur-warm", type=int, default=100)
    ap.add_argument("--sur-samples", type=int, default=120)
    ap.add_argument("--sur-iters", type=int, default=8)
    ap.add_argument("--sur-stride", type=int, default=32)
    ap.add_argument("--sur-steps", type=int, default=10)
    ap.add_argument("--sur-blocks", type=int, default=8)
    ap.add_argument("--sur-k", type=int, default=10)
    ap.add_argument("--sur-tol", type=float, default=0.01)
    ap.add_argument("--sur-pretrain", action="store_true",
                    help="train on the 2D droplet (80x80) and 64x32x64 dam break trajectories first "
                         "(measured worse on the 1M deck: more blocks rolled back)")
    ap.add_argument("--sur-radius", type=float, default=0.0,
                    help="message radius of the surrogate / l0 (0: the largest kernel radius)")
    ap.add_argument("--late-steps", type=int, default=200,
                    help="steps of a longer developed-flow window after the timed one; 0: skip")
    ap.add_argument("--error-steps", type=int, default=500,
                    help="steps after which the mixed leg's fields are compared with FP64")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from mphexplicit_amd.solver import MPH_FP64, MPH_MIXED, Solver

    deck = make_deck(args.workload)
    prec = MPH_FP64 if args.precision == "fp64" else MPH_MIXED
    log("workload %s: %d particles" % (deck.name, deck.n))
    t0 = time.perf_counter()
    s = Solver.from_deck(deck, prec)
    log("create %.2fs" % (time.perf_counter() - t0))
    steps_per_s, tm, k2, win = timed_run(s, args)
    st = s.stats()
    surface = bool(any(t > 0 for t in deck.data.surface_tension))
    fa, fb = FLOPS_PER_PAIR[surface]
    pairs_per_step = tm["pairs"] / k2
    phases = {"move": tm["move_ms"] / k2, "rebuild": tm["rebuild_ms"] / k2,
              "sweep_a": tm["sweep_a_ms"] / k2, "sweep_b": tm["sweep_b_ms"] / k2}
    dom = max(("sweep_a", "sweep_b"), key=lambda k: phases[k])
    fl = (fa if dom == "sweep_a" else fb) * pairs_per_step
    achieved = fl / (phases[dom] * 1e-3) / 1e12
    roofline = {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFLOPS,
                "traffic": measured_traffic(deck.name, st["cluster_size"], "k_" + dom) if prec == MPH_FP64 else None,
                "kernel": "k_" + dom,
                "note": "FP64 VALU kernel; MI355X's fp64 vector and matrix peaks are both 78.6 TFLOP/s; "
                        "achieved = %d algorithmic flops/pair x %.4g pairs per launch / mean launch time; "
                        "traffic = HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE (scaled by the "
                        "counters of a device copy of known size), profiles/hbm_traffic_%s.json" % (fa if dom == "sweep_a" else fb,
                                                                      pairs_per_step, deck.name)}
    mixed = None
    if prec == MPH_FP64 and not args.no_mixed:
        # the same deck with FP32 pair math: timing, and the error of every field against the
        # FP64 run after the same number of steps
        m = Solver.from_deck(deck, MPH_MIXED)
        m_sps, m_tm, _, m_win = timed_run(m, args)
        # the error after the same number of steps from the same state: a fresh FP64 and mixed
        # pair run --error-steps steps
        e64, e32 = Solver.from_deck(deck, MPH_FP64), Solver.from_deck(deck, MPH_MIXED)
        e64.step(args.error_steps)
        e32.step(args.error_steps)
        n_err = args.error_steps
        errs = field_errors(e32.fields(names=ERROR_FIELDS), e64.fields(names=ERROR_FIELDS))
        e64.close()
        e32.close()
        m_ph = {"sweep_a": m_tm["sweep_a_ms"] / k2, "sweep_b": m_tm["sweep_b_ms"] / k2}
        m_dom = max(m_ph, key=lambda k: m_ph[k])
        m_ach = (fa if m_dom == "sweep_a" else fb) * (m_tm["pairs"] / k2) / (m_ph[m_dom] * 1e-3) / 1e12
        mixed = {"value": m_sps, "unit": "steps/s", "ms_per_step": 1e3 / m_sps,
                 "speedup_vs_fp64": m_sps / steps_per_s,
                 "phase_ms": {"move": m_tm["move_ms"] / k2, "rebuild": m_tm["rebuild_ms"] / k2, **m_ph},
                 "roofline": {"bound": "valu", "achieved": m_ach, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                              "frac": m_ach / FP32_PEAK_TFLOPS, "kernel": "k_" + m_dom + "_mixed"},
                 "error_vs_fp64_after_steps": n_err, "error_vs_fp64": errs}
        m.close()
    late = None
    if args.late_steps > 0:
        # a longer window of the same developed flow (the timed window is --steps long)
        cur = s.stats()["steps"]
        rb0 = s.stats()["rebuilds"]
        late = {"from_step": cur, "steps": args.late_steps, "value": timed_steps(s, args, args.late_steps),
                "unit": "steps/s", "rebuilds": s.stats()["rebuilds"] - rb0}
    batch = None if args.no_batch else batch_leg(args, Solver)
    surrogate = None
    if prec == MPH_FP64 and not args.no_surrogate:
        surrogate = surrogate_leg(args, deck)
    cpu = None
    if not args.no_cpu_baseline:
        cores = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
        try:
            cpu = cpu_baseline(deck, args.cpu_steps, cores)
        except Exception as e:  # the baseline is reported, never required
            log("cpu baseline failed: %s" % e)
    out = {
        "metric": "3D 1M-particle steps/s & pair-interactions/s; surrogate vs plain GPU at tol",
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 / steps_per_s,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == MPH_FP64 else "f64 state / f32 pair math",
        "data": "synthetic (seeded lattice deck, numpy PCG64)",
        "config": {"workload": deck.name, "particles": deck.n, "fluid": int((deck.type < 2).sum()),
                   "precision": args.precision, "dt": deck.data.dt, "graph": bool(args.graph),
                   "list_width": st["list_width"],
                   "cluster_size": st["cluster_size"], "list_grid": st["list_grid"]},
        "timed_window": {"from_step": win["from_step"], "steps": args.steps, "rebuilds": win["rebuilds"],
                         "note": "value: steps/s of the --steps steps from step --from-step (developed flow)"},
        "start_of_flow": win["start_of_flow"],
        "pair_interactions_per_s": pairs_per_step * steps_per_s,
        "pairs_per_step": pairs_per_step,
        "slots_per_pair": st["eval_slots"] / pairs_per_step if pairs_per_step else None,
        "simulated_seconds_per_s": deck.data.dt * steps_per_s,
        "rebuilds_in_profiled_window": tm["rebuilds"],
        "phase_ms": phases,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "developed_flow": late,
        "mixed": mixed,
        "batch64": batch,
        "surrogate": surrogate,
    }
    if cpu:
        # the CPU sample is the start of the flow (steps 2-21, few rebuilds); the GPU's
        # start-of-flow window is the like-for-like comparison
        out["speedup_vs_cpu"] = steps_per_s / cpu["value"]
        out["speedup_vs_cpu_same_window"] = win["start_of_flow"]["value"] / cpu["value"]
    print(json.dumps(out), flush=True)
    s.close()


if __name__ == "__main__":
    main()
