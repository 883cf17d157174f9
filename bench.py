#!/usr/bin/env python3
"""Benchmark of the MI355X IFGF forward map.

Workload: sound-soft kite x = cos t + 0.65 cos 2t - 0.65, y = 1.5 sin t at
k = 200, 6,400 equal-parameter patches x 10 Fejer nodes (N = 64,000),
combined-field operator (eta = k), 5 x 5 Chebyshev nodes per cone, 10 tree
levels, all FP64. One step = one forward-map application (the operator GMRES
applies once per iteration) on a seeded complex density already resident in
HBM. The near-field kernel values and the folded leaf operator depend only on
the geometry and are evaluated once at plan creation (counted in
plan_setup_s); the on-the-fly variant that re-evaluates them in every
application is timed too (on_the_fly_matvecs_per_s). After the timed steps, GMRES(50) solves to 1e-8 for the plane wave
d = (cos pi/4, sin pi/4) give the time to solve (the median of three solves
after a first one whose one-time costs are reported apart). The mixed-precision
operator (FP32 cone work) and its FGMRES solve with FP64 refinement are timed
beside it, and so are 64-column applications through the batched kernels and
their 64-direction lockstep solve, sixteen random star obstacles applied one
after another and as one captured group launch, and the field evaluator on a
256 x 256 target grid.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stage_work(st, formulation, P, Ps, Pang, stored):
    """Algorithmic work of each stage of one application: (bound, amount, unit).

    Stored operators stream 16 bytes per near pair / leaf evaluation from HBM;
    kernel generation costs 40 (single layer) or 60 (double layer / combined)
    FLOP per pair plus 8 per complex multiply-add; a cone interpolation costs
    4P FLOP plus 20 for its centre factor; a cone transform 4P(Ps+Pang)."""
    gen = 40 if formulation == 0 else 60
    w = {}
    if stored:
        w["near"] = ("hbm", 16.0 * st["near_stored"], "B")
        w["leaf"] = ("hbm", 16.0 * st["leaf"], "B")
    else:
        w["near"] = ("mfma", (gen + 8.0) * st["near"], "FLOP")
        w["leaf"] = ("mfma", (gen + 8.0) * st["leaf"] + 4.0 * P * (Ps + Pang) * st["leaf_cones"], "FLOP")
    w["propagation"] = ("mfma", (4.0 * P + 20) * st["prop"] + 4.0 * P * (Ps + Pang) * st["prop_cones"], "FLOP")
    w["interpolation"] = ("mfma", (4.0 * P + 20) * st["interp"], "FLOP")
    return w


def timed_solve(fn, reps=5, settle=3):
    """(result of the last call, cold seconds, warm seconds): the first call
    carries the one-time costs (first launches of the solver kernels, the
    workspaces the plan then keeps); the next settle calls are untimed (right
    after a plan's setup the first few solves run slower while the fresh
    workspaces settle: 9.7, 9.6, 8.9 ms, then 8.5 ms, for the mixed FGMRES on
    the 64k kite); the next reps calls are warm. Warm seconds: {"min",
    "median", "max"} over those calls."""
    import torch
    ts = []
    for q in range(reps + 1 + settle):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        if q == 0 or q > settle:
            ts.append(time.perf_counter() - t0)
    warm = sorted(ts[1:])
    return out, ts[0], {"min": round(warm[0], 5), "median": round(warm[len(warm) // 2], 5),
                        "max": round(warm[-1], 5)}


def cpu_baseline(kw, x, name):
    """The oracle (CPU restatement, the reference's Amos Hankel) on one
    forward-map application of the same workload, oracle setup excluded: once
    on all the host threads OpenMP is given (OMP_NUM_THREADS; the count is
    reported) and once on one core. Returns (record, the oracle's output):
    the output is the full-size parity check of the GPU operator."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = O.set_threads(0)
    t0 = time.time()
    om = O.ForwardMap(kw["curve"][0], kw["curve"][1], kw["patch_lims"], formulation=O.COMBINED,
                      k=kw["k"], eta=kw["k"], delta=kw["delta"], p=kw["p"], nlevels=kw["nlevels"])
    log(f"cpu baseline: oracle setup {time.time() - t0:.1f} s on {threads} threads")
    t0 = time.time()
    y = om.apply(x)
    dt = time.time() - t0
    log(f"cpu baseline: one oracle application {dt:.2f} s on {threads} threads")
    O.set_threads(1)
    t0 = time.time()
    y1 = om.apply(x)
    dt1 = time.time() - t0
    O.set_threads(threads)
    log(f"cpu baseline: one oracle application {dt1:.2f} s on 1 thread (same bits: {bool((y1 == y).all())})")
    rec = {"value": 1.0 / dt, "unit": "matvecs/s", "cores": threads, "kind": "port",
           "sample": f"1 forward-map application (combined field, 5x5) of the {name} by the oracle "
                     f"(OpenMP over targets and cones, {threads} threads) with the reference's Amos Hankel; "
                     "oracle setup excluded",
           "one_core": {"value": 1.0 / dt1, "cores": 1}}
    return rec, y


def star_directions_block(I, configs, np, log, reps=1):
    """The headline many-right-hand-side comparison on the 256,000-point star at
    k = 500 (C3 geometry, C4 directions): 64 plane waves 2 pi j / 64 solved to
    the FP64 true residual 1e-8 (a) by 64 single FP64 GMRES(50) solves, (b) by
    the FP64 lockstep GMRES over batched applications and (c) by the
    mixed-precision refinement on a mixed plan; wall time of each (warm), the
    ratios, and the scattered field at 64 points on r = 3 of (b) and (c)
    against (a)'s FP64 solutions."""
    import torch
    kw, _ = configs.star_k500(nrhs_max=64)
    t0 = time.time()
    f64 = I.ForwardMap(**kw, autotune=True, device_setup=True)
    setup64 = time.time() - t0
    x, y, *_ = f64.points()
    th = 2 * np.pi * np.arange(64) / 64
    B = np.stack([-configs.plane_wave(x, y, kw["k"], (np.cos(t_), np.sin(t_))) for t_ in th], axis=1)
    tt = 2 * np.pi * np.arange(64) / 64
    tx, ty = 3.0 * np.cos(tt), 3.0 * np.sin(tt)
    f64.solve(B[:, 0], restart=50, tol=1e-8, maxit=2000)  # first call (kernel loads, workspaces)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    singles, its1 = [], []
    for j in range(64):
        d_, it_, rr_, _ = f64.solve(B[:, j], restart=50, tol=1e-8, maxit=2000)
        singles.append(d_)
        its1.append(it_)
    t_single = time.perf_counter() - t0
    log(f"64 star directions: 64 single FP64 solves {t_single:.3f} s (iterations {min(its1)}..{max(its1)})")
    far = [f64.propagate_unacc(tx, ty, d_) for d_ in singles]
    (X64, its, rr), _, t_lock = timed_solve(lambda: f64.solve_multi(B, restart=50, tol=1e-8, maxit=2000), reps=reps,
                                            settle=1)
    e_lock = max(float(np.linalg.norm(f64.propagate_unacc(tx, ty, X64[:, j]) - far[j]) / np.linalg.norm(far[j]))
                 for j in range(64))
    log(f"64 star directions: FP64 lockstep {t_lock} s, max true relres {rr.max():.2e}, far field vs single {e_lock:.1e}")
    del f64
    t0 = time.time()
    fmx = I.ForwardMap(**dict(kw, precision="mixed"), autotune=True, device_setup=True)
    setupm = time.time() - t0
    (Xm, outer, itm, rrm), _, t_ref = timed_solve(
        lambda: fmx.solve_multi_refined(B, restart=50, tol=1e-8, maxit=2000), reps=reps, settle=1)
    e_mix = max(float(np.linalg.norm(fmx.propagate_unacc(tx, ty, Xm[:, j]) - far[j]) / np.linalg.norm(far[j]))
                for j in range(64))
    log(f"64 star directions: mixed refinement {t_ref} s, {outer} steps, max true relres {rrm.max():.2e}, "
        f"far field vs FP64 {e_mix:.1e}")
    del fmx
    return {"workload": "star r=1+0.3cos5t, k=500, N=256000, 64 plane waves 2 pi j/64, CFIE 5x5, tol 1e-8",
            "plan_setup_s": {"fp64": round(setup64, 3), "mixed": round(setupm, 3)},
            "fp64_single_solves_s": round(t_single, 4), "single_iterations": [min(its1), max(its1)],
            "fp64_lockstep_s": t_lock["median"], "fp64_lockstep_true_relres_max": float(rr.max()),
            "fp64_lockstep_far_field_rel_diff_max": e_lock,
            "mixed_refined_s": t_ref["median"], "mixed_refinement_steps": outer,
            "mixed_true_relres_max": float(rrm.max()), "mixed_far_field_rel_diff_vs_fp64_max": e_mix,
            "speedup_mixed_vs_fp64_single": round(t_single / t_ref["median"], 3),
            "speedup_mixed_vs_fp64_lockstep": round(t_lock["median"] / t_ref["median"], 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-on-the-fly", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-group", action="store_true")
    ap.add_argument("--no-field", action="store_true")
    ap.add_argument("--no-star-directions", action="store_true",
                    help="skip the 64-direction comparison on the 256k star")
    ap.add_argument("--workload", choices=("kite", "star"), default="kite",
                    help="kite: the 64k kite at k=200 (default); star: the 256k star r=1+0.3cos5t at k=500")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 20 device copies of 256 MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda:0")
        b_ = torch.empty_like(a)
        for _ in range(20):
            b_.copy_(a)
        torch.cuda.synchronize()
        log("copy calibration: 20 copies of 268435456 bytes")
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the forward map "
              "runs on a single GPU (no sharding across GPUs)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import ifgf2d_amd as I
    from ifgf2d_amd import configs

    kw, d = configs.kite_k200() if args.workload == "kite" else configs.star_k500()
    t0 = time.time()
    # every setup step on the GPU (sort, near-singular search, correction, tree)
    fm = I.ForwardMap(**kw, autotune=True, device_setup=True)
    t_setup = time.time() - t0
    kw["delta"] = fm.delta
    st = fm.stats()
    log(f"plan: N={fm.N} setup {t_setup:.2f} s stats {st}")

    # seeded splitmix64 + Box-Muller density (configs.seeded_density; the oracle's
    # C++ draws the same values)
    x_host = configs.seeded_density(fm.N, seed=0)
    dev = torch.device("cuda:0")
    xs = torch.from_numpy(x_host).to(dev)
    ys = torch.empty_like(xs)
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        fm.apply_device(xs.data_ptr(), ys.data_ptr())
    torch.cuda.synchronize()

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fm.apply_device(xs.data_ptr(), ys.data_ptr())
    torch.cuda.synchronize()
    t_run = time.perf_counter() - t0
    ms_per_step = 1e3 * t_run / args.steps
    value = args.steps / t_run
    log(f"{args.steps} applications in {t_run:.3f} s -> {value:.1f} matvecs/s")
    # per-stage times from a separate pass with events around every stage
    # (recording them costs several microseconds per stage, so they stay out
    # of the timed loop above)
    nprof = max(1, min(args.steps, 200))
    fm.profile(True)
    for _ in range(nprof):
        fm.apply_device(xs.data_ptr(), ys.data_ptr())
    stages, count = fm.stage_times()
    fm.profile(False)
    stage_avg = {k_: v / max(count, 1) for k_, v in stages.items()}
    log(f"stages (ms/apply over {count} profiled applications) "
        + ", ".join(f"{k_} {v:.4f}" for k_, v in stage_avg.items()))

    # dominant stage and its roofline (algorithmic bytes or FLOP per application)
    P = 25
    ext = fm.extra_stats()
    work = stage_work(dict(st, **ext), fm.formulation, P, 5, 5, ext["stored"])
    nlev = kw["nlevels"]
    stored = ext["stored"]
    if stored:
        # one push launch per level carries propagation and interpolation;
        # per evaluation it reads its 32 B (s, theta, factor) record and
        # writes its 16 B value; per parent-cone node it reads a 16 B
        # child-slot index and the children's pushed values (16 B per
        # propagation evaluation); per parent cone a 4 B row of the inverse
        # node factors, stored once per (level, radial interval, s node), each
        # 16 B factor counted once
        flop = work["propagation"][1] + work["interpolation"][1]
        nev = st["prop"] + st["interp"]
        pcones = st["cones"] - ext["leaf_cones"]
        byts = (48.0 * nev + 16.0 * 25 * pcones + 16.0 * st["prop"] + 4.0 * pcones
                + 16.0 * ext["ifac_entries"])
        work["propagation"] = ("mixed", (flop, byts), "FLOP+B")
        work.pop("interpolation")
    dom = max(work, key=lambda s: stage_avg[s])
    bound, amount, unit = work[dom]
    secs = stage_avg[dom] * 1e-3
    launches = {"near": 1, "leaf": 1, "propagation": (nlev - 2) if stored else (nlev - 3), "interpolation": 1}
    other = None
    if bound == "mixed":
        flop, byts = amount
        f_flop = flop / secs / 1e12 / PEAK_FP64_TFLOPS
        f_byte = byts / secs / 1e9 / PEAK_HBM_GBS
        other = {"flop_per_application": flop, "bytes_per_application": byts,
                 "fp64_frac": round(f_flop, 4), "hbm_frac": round(f_byte, 4)}
        bound, amount, unit = ("hbm", byts, "B") if f_byte >= f_flop else ("mfma", flop, "FLOP")
    if bound == "hbm":
        achieved, peak, runit = amount / secs / 1e9, PEAK_HBM_GBS, "GB/s"
    else:
        achieved, peak, runit = amount / secs / 1e12, PEAK_FP64_TFLOPS, "TFLOP/s"
    kname = {"near": "k_near_items" if stored else "k_near",
             "leaf": "k_leaf_items" if stored else "k_leaf",
             "propagation": "k_level_push" if stored else "k_prop",
             "interpolation": "k_interp"}[dom]
    roofline = {"bound": bound, "kernel": kname, "achieved": round(achieved, 3), "peak": peak,
                "unit": runit, "frac": round(achieved / peak, 4), "traffic": None,
                "algorithmic_per_application": amount, "algorithmic_unit": unit,
                "launches_per_application": launches[dom],
                "algorithmic_per_launch": amount / launches[dom],
                "avg_launch_ms": stage_avg[dom] / launches[dom]}
    if other:
        roofline["both_resources"] = other
    # the on-the-fly variant of the same operator (kernels re-evaluated per application)
    otf = None
    if not args.no_on_the_fly:
        fo = I.ForwardMap(**dict(kw, on_the_fly=True))
        for _ in range(3):
            fo.apply_device(xs.data_ptr(), ys.data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nrep = max(1, args.steps // 5)
        for _ in range(nrep):
            fo.apply_device(xs.data_ptr(), ys.data_ptr())
        torch.cuda.synchronize()
        otf = round(nrep / (time.perf_counter() - t0), 3)
        log(f"on-the-fly kernels: {otf} matvecs/s")
        del fo
    tag = "kite_k200" if args.workload == "kite" else "star_k500"
    traffic_file = os.path.join(ROOT, "profiles", f"traffic_{roofline['kernel']}_{tag}.json")
    if os.path.exists(traffic_file):
        roofline["traffic"] = json.load(open(traffic_file)).get("bytes_per_launch")

    # the timed matvec against the unaccelerated FP64 operator (direct sum plus the
    # same correction) at the first 2,000 patch-order targets
    y_acc = fm.compute_Ax_acc(x_host)
    tgt = np.arange(min(2000, fm.N))
    y_dir = fm.compute_Ax_unacc(x_host, tgt)
    check = {"targets": int(tgt.size),
             "rel_err_vs_direct": float(np.linalg.norm(y_acc[tgt] - y_dir) / np.linalg.norm(y_dir))}
    log(f"matvec vs direct FP64 sum at {tgt.size} targets: {check['rel_err_vs_direct']:.2e}")

    solve = None
    rhs = dens = None
    if not args.no_solve:
        xg, yg, *_ = fm.points()
        rhs = -configs.plane_wave(xg, yg, kw["k"], d)
        (dens, it, relres, hist), t_cold, t_solve = timed_solve(
            lambda: fm.solve(rhs, restart=50, tol=1e-8, maxit=2000))
        solve = {"time_to_solve_s": t_solve["median"], "warm_s": t_solve, "first_call_s": round(t_cold, 4),
                 "gmres_iterations": it, "true_relres": relres, "restart": 50, "tol": 1e-8,
                 "last_call_phases_s": {k_: round(v, 5) for k_, v in fm.solver_times().items()}}
        log(f"GMRES: {it} iterations, {t_solve} s warm, true relres {relres:.2e}, phases {fm.solver_times()}")

    # mixed precision: FP32 leaf operator, cone coefficients, propagation and
    # interpolation (FP64 near field, correction, accumulation), and the
    # FP64-outer FGMRES that recovers the 1e-8 FP64 true residual
    mixed = None
    if not args.no_mixed:
        fmx = I.ForwardMap(**dict(kw, precision="mixed"), autotune=True)
        for _ in range(args.warmup):
            fmx.apply_device(xs.data_ptr(), ys.data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fmx.apply_device(xs.data_ptr(), ys.data_ptr())
        torch.cuda.synchronize()
        mv = args.steps / (time.perf_counter() - t0)
        ym = fmx.compute_Ax_acc(x_host)
        y64 = fmx.compute_Ax_acc(x_host, fp64=True)
        mixed = {"matvecs_per_s": round(mv, 3),
                 "rel_diff_vs_fp64_operator": float(np.linalg.norm(ym - y64) / np.linalg.norm(y64))}
        log(f"mixed precision: {mv:.1f} matvecs/s, operator rel diff vs FP64 {mixed['rel_diff_vs_fp64_operator']:.2e}")
        if not args.no_solve:
            (dm, oit, iit, rr, _), t_fc, t_f = timed_solve(
                lambda: fmx.solve_fgmres(rhs, restart=20, tol=1e-8, inner_restart=60))
            mixed["fgmres"] = {"time_to_solve_s": t_f["median"], "warm_s": t_f, "first_call_s": round(t_fc, 4),
                               "outer_iterations": oit, "inner_iterations": iit, "true_relres": rr,
                               "density_rel_diff_vs_fp64_gmres": float(np.linalg.norm(dm - dens) / np.linalg.norm(dens)),
                               "last_call_phases_s": {k_: round(v, 5) for k_, v in fmx.solver_times().items()}}
            log(f"mixed FGMRES: {oit} outer + {iit} inner iterations, {t_f} s warm, true relres {rr:.2e}, "
                f"phases {fmx.solver_times()}")
        del fmx

    # several right-hand sides per application: 64 columns in blocks of 8
    # through the batched kernels, FP64 and mixed precision
    batched = None
    if not args.no_batched:
        batched = {"nrhs": 64}
        Xb = torch.from_numpy(configs.seeded_density(64 * fm.N, seed=1).reshape(64, fm.N)).to(dev)
        Yb = torch.empty_like(Xb)
        for prec in ("fp64", "mixed", "fp64_valu", "mixed_valu"):
            fb = I.ForwardMap(**dict(kw, precision=prec.split("_")[0]), nrhs_max=64, autotune=True,
                              mfma=not prec.endswith("valu"))
            fb.apply_device(Xb.data_ptr(), Yb.data_ptr(), nrhs=64)
            torch.cuda.synchronize()
            reps = max(2, args.steps // 100)
            t0 = time.perf_counter()
            for _ in range(reps):
                fb.apply_device(Xb.data_ptr(), Yb.data_ptr(), nrhs=64)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / reps
            batched[prec] = {"ms_per_application": round(dt * 1e3, 4), "column_matvecs_per_s": round(64 / dt, 1)}
            log(f"batched {prec}: 64 columns in {dt * 1e3:.3f} ms -> {64 / dt:.1f} column-matvecs/s")
            if prec == "fp64" and not args.no_solve:
                # 64 incidence directions 2 pi j / 64 solved together (lockstep GMRES(50))
                xg, yg, *_ = fb.points()
                th = 2 * np.pi * np.arange(64) / 64
                B = np.stack([-configs.plane_wave(xg, yg, kw["k"], (np.cos(t_), np.sin(t_))) for t_ in th], axis=1)
                (X64, its, rr), t_mc, t_m = timed_solve(lambda: fb.solve_multi(B, restart=50, tol=1e-8, maxit=2000),
                                                        reps=2)
                batched["solve_64_directions"] = {"time_to_solve_s": t_m["median"], "warm_s": t_m,
                                                  "first_call_s": round(t_mc, 4),
                                                  "iterations_min": int(its.min()),
                                                  "iterations_max": int(its.max()),
                                                  "true_relres_max": float(rr.max())}
                log(f"64 directions solved together: {t_m} s, iterations {its.min()}..{its.max()}, "
                    f"max true relres {rr.max():.2e}")
            if prec == "mixed" and not args.no_solve and "solve_64_directions" in batched:
                # the same 64 directions: lockstep mixed-precision GMRES corrections,
                # FP64 residuals (ifgf_refine_multi)
                (Xm, outer, its, rr), t_mc, t_m = timed_solve(
                    lambda: fb.solve_multi_refined(B, restart=50, tol=1e-8, maxit=2000), reps=2)
                dmax = max(float(np.linalg.norm(Xm[:, j] - X64[:, j]) / np.linalg.norm(X64[:, j])) for j in range(64))
                batched["solve_64_directions_mixed_refined"] = {
                    "time_to_solve_s": t_m["median"], "warm_s": t_m, "first_call_s": round(t_mc, 4),
                    "refinement_steps": outer,
                    "inner_iterations_min": int(its.min()), "inner_iterations_max": int(its.max()),
                    "true_relres_max": float(rr.max()), "density_rel_diff_vs_fp64_max": dmax}
                log(f"64 directions, mixed refinement: {t_m} s, {outer} steps, inner iterations "
                    f"{its.min()}..{its.max()}, max true relres {rr.max():.2e}, density diff {dmax:.1e}")
            del fb
        del Xb, Yb

    star_dirs = None
    if not args.no_star_directions and not args.no_solve:
        star_dirs = star_directions_block(I, configs, np, log)

    # sixteen seeded random star obstacles at k=100 (~10 nodes per wavelength):
    # their forward maps one after another against one captured group launch
    group = None
    if not args.no_group:
        maps = [I.ForwardMap(**kw_, autotune=True) for kw_ in configs.random_stars(16, k=100.0, seed=0)]
        gx = [torch.from_numpy(configs.seeded_density(m.N, seed=2 + j)).to(dev) for j, m in enumerate(maps)]
        gy = [torch.empty_like(v) for v in gx]
        torch.cuda.synchronize()
        reps = max(5, args.steps // 10)
        for m, a_, b_ in zip(maps, gx, gy):
            m.apply_device(a_.data_ptr(), b_.data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            for m, a_, b_ in zip(maps, gx, gy):
                m.apply_device(a_.data_ptr(), b_.data_ptr())
        torch.cuda.synchronize()
        t_seq = (time.perf_counter() - t0) / reps
        grp = I.ObstacleGroup(maps, [v.data_ptr() for v in gx], [v.data_ptr() for v in gy])
        grp.launch()
        grp.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            grp.launch()
        grp.synchronize()
        t_grp = (time.perf_counter() - t0) / reps
        gst = grp.stats()
        group = {"obstacles": len(maps), "N": [m.N for m in maps], "nodes_per_wavelength": 10,
                 "sequential_ms_per_matvec": round(t_seq * 1e3, 4), "group_ms_per_matvec": round(t_grp * 1e3, 4),
                 "speedup": round(t_seq / t_grp, 2), "group_launches": gst["launches"],
                 "group_workgroups": gst["workgroups"]}
        log(f"16 random stars: sequential {t_seq * 1e3:.3f} ms, one group launch {t_grp * 1e3:.3f} ms per matvec "
            f"({gst['launches']} fused kernel launches)")
        grp.close()
        # cone orders chosen per obstacle by the autotuner (relative error 1e-4
        # against the unaccelerated operator), fused per class of equal order
        t0 = time.time()
        orders = []
        for kw_ in configs.random_stars(16, k=100.0, seed=0):
            ps_, pa_, e_, ok_ = I.autotune_cone_order(kw_["curve"], kw_["patch_lims"], kw_["k"], kw_["nlevels"],
                                                      1e-4, p=kw_["p"])
            orders.append((ps_, pa_, e_))
        t_tune = time.time() - t0
        maps2 = [I.ForwardMap(**kw_, Ps=o_[0], Pang=o_[1], autotune=True)
                 for kw_, o_ in zip(configs.random_stars(16, k=100.0, seed=0), orders)]
        for m, a_, b_ in zip(maps2, gx, gy):
            m.apply_device(a_.data_ptr(), b_.data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            for m, a_, b_ in zip(maps2, gx, gy):
                m.apply_device(a_.data_ptr(), b_.data_ptr())
        torch.cuda.synchronize()
        t_seq2 = (time.perf_counter() - t0) / reps
        grp2 = I.ObstacleGroup(maps2, [v.data_ptr() for v in gx], [v.data_ptr() for v in gy])
        grp2.launch()
        grp2.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            grp2.launch()
        grp2.synchronize()
        t_grp2 = (time.perf_counter() - t0) / reps
        group["autotuned_cone_orders"] = {
            "tol": 1e-4, "orders": [f"{o_[0]}x{o_[1]}" for o_ in orders],
            "errors_max": max(o_[2] for o_ in orders), "autotune_s": round(t_tune, 3),
            "classes": grp2.stats()["classes"], "sequential_ms_per_matvec": round(t_seq2 * 1e3, 4),
            "group_ms_per_matvec": round(t_grp2 * 1e3, 4)}
        log(f"16 random stars, autotuned cone orders (1e-4): {sorted(set(group['autotuned_cone_orders']['orders']))}, "
            f"sequential {t_seq2 * 1e3:.3f} ms, group {t_grp2 * 1e3:.3f} ms")
        grp2.close()
        del maps2
        del maps, gx, gy

    # the scattered field on a 256 x 256 grid over the scatterer (65,536 targets)
    # through the IFGF field evaluator (5x5 and 9x9) against the GPU direct sum
    field = None
    if not args.no_field and args.workload == "kite":
        g = np.linspace(-2.0, 2.0, 256)
        tx, ty = (a_.ravel().copy() for a_ in np.meshgrid(g, g))
        u = torch.from_numpy(x_host).to(dev)
        sub = np.random.default_rng(7).choice(tx.size, 2000, replace=False)
        t0 = time.perf_counter()
        direct_all = fm.propagate_unacc(tx, ty, x_host)
        t_direct = time.perf_counter() - t0
        field = {"targets": int(tx.size), "grid": "256 x 256 on [-2, 2]^2",
                 "direct_sum_s": round(t_direct, 4)}
        for P_ in (5, 9):
            t0 = time.perf_counter()
            f = fm.field(tx, ty, Ps=P_, Pang=P_)
            t_create = time.perf_counter() - t0
            out_d = torch.empty(f.M, dtype=torch.complex128, device=dev)
            for _ in range(3):
                f.apply_device(u.data_ptr(), out_d.data_ptr())
            torch.cuda.synchronize()
            reps = max(10, args.steps // 20)
            t0 = time.perf_counter()
            for _ in range(reps):
                f.apply_device(u.data_ptr(), out_d.data_ptr())
            torch.cuda.synchronize()
            t_eval = (time.perf_counter() - t0) / reps
            vals = out_d.cpu().numpy()
            err = float(np.linalg.norm(vals[sub] - direct_all[sub]) / np.linalg.norm(direct_all[sub]))
            fs = f.stats()
            field[f"{P_}x{P_}"] = {"ms_per_evaluation": round(t_eval * 1e3, 4), "create_s": round(t_create, 3),
                                   "rel_err_vs_direct": err, "levels": fs["nlevels"], "cones": fs["cones"],
                                   "stored_bytes": fs["stored_bytes"],
                                   "speedup_vs_direct": round(t_direct / t_eval, 1)}
            log(f"field {P_}x{P_}: {t_eval * 1e3:.3f} ms per evaluation of 65,536 targets "
                f"(direct sum {t_direct:.3f} s), rel err {err:.2e}")
            f.close()
            del out_d

    cpu = parity = None
    if not args.no_cpu_baseline:
        cpu, y_oracle = cpu_baseline(
            kw, x_host, "64,000-point kite at k=200" if args.workload == "kite" else "256,000-point star at k=500")
        # the timed operator against the oracle at full size (same density)
        parity = {"N": fm.N, "bar": 1e-12,
                  "rel_diff": float(np.linalg.norm(y_acc - y_oracle) / np.linalg.norm(y_oracle))}
        log(f"GPU FP64 operator vs oracle at N={fm.N}: {parity['rel_diff']:.2e}")

    out = {
        "metric": "time-to-solve (rel. residual 1e-8) + IFGF matvecs/s vs plain FP64 on GPU",
        "value": round(value, 3), "unit": "matvecs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": ("kite k=200, 6400 patches x 10 Fejer nodes (N=64000), combined "
                                "field eta=k, 5x5 cone nodes, 10 levels, FP64") if args.workload == "kite" else
                               ("star r=1+0.3cos5t k=500, 25600 patches x 10 Fejer nodes (N=256000), combined "
                                "field eta=k, 5x5 cone nodes, 12 levels, FP64"),
                   "N": fm.N, "nrhs": 1, "nlevels": kw["nlevels"], "Ps": 5, "Pang": 5,
                   "parallelism": "single GPU"},
        "stages_ms": {k_: round(v, 5) for k_, v in stage_avg.items()},
        "work": {k_: st[k_] for k_ in ("near", "leaf", "prop", "interp", "cones", "corr")},
        "plan_setup_s": round(t_setup, 3),
        "stored_operator_bytes": ext["stored_bytes"],
        "autotuned_lanes": ext["lanes"],
        "on_the_fly_matvecs_per_s": otf,
        "matvec_check": check,
        "solve": solve,
        "mixed_precision": mixed,
        "batched_rhs": batched,
        "star_64_directions": star_dirs,
        "obstacle_group": group,
        "field": field,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "oracle_parity": parity,
    }
    if cpu:
        out["gpu_over_cpu"] = round(value / cpu["value"], 1)
        out["gpu_over_cpu_one_core"] = round(value / cpu["one_core"]["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
