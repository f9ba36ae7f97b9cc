#!/usr/bin/env python3
"""Benchmark of the MI355X quasi-static Newton step (LPM-CPP's SolverStatic path).

Workload: the FP64 drop-in configuration -- a 3D simple-cubic elastic cube of
40^3 = 64,000 particles (r = 0.1 mm, shells 6+12), E = 70 GPa, nu = 0.33, bottom
face fixed, top face pulled 0.05 mm in 20 quasi-static load steps.  One bench
step is one complete solve (20 load steps: tangent assembly, modified-Newton
iterations with FP64 forward passes and Jacobi-preconditioned CG, state update)
of a device-resident system; the particle topology (bond search, conns, CSR
pattern) is built before timing and reported as setup_ms (the process's first
handle) and setup_warm_ms (a second build, as a running server pays it).

Printed JSON: value = FP64 solves/s; the FP32-inner mixed-precision path on the
same workload, its error against the FP64 solution, p50/p99 solve latency, the
SpMV roofline (algorithmic bytes over the CG SpMV's average duration, from
start/stop events attached to every CG SpMV dispatch of one extra solve, since
the timed solves replay HIP graphs), assembly and forward-pass rates, and the
CPU baseline: the oracle (a multithreaded FP64 restatement of the reference) on
the first load steps, with the reference's CG and Newton settings, scaled to
solves/s.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lpm-cpp_amd"))

METRIC = "solves/sec + p99 latency at 1e-8 rel. err vs FP64; SpMV/assembly HBM GB/s"
HBM_PEAK_GBS = 8000.0
NEWTON_TOL = 1e-10
# FP64 drop-in: Jacobi-PCG stopped where the reference stops MKL's CG (||r||^2 <=
# 1e-12 ||r0||^2, solver.h:312), Newton tightened to 1e-10 so results meet 1e-8.
FP64_OPTS = dict(precond=1, rtol=1e-12, atol=0.0, precision=0)
# Mixed (the served path's scheme): FP32 stiffness and CG as an inexact inner
# solve (||r||^2 <= 3e-4 ||r0||^2) of the FP64 modified Newton, whose FP64
# residual refines to the same 1e-10, each load step starting from the last
# step's increment (solver option 6).  The inner tolerance is the fastest of
# 1e-5 ... 1e-2 on this workload (looser: more Newton iterations; tighter: more
# CG iterations per solve); the error stays near 2e-11 across that range.
MIXED_RTOL = 3e-4
MIXED_OPTS = dict(precond=1, rtol=MIXED_RTOL, atol=0.0, precision=1, predict=1.0)
# Its parts, each measured alone: the mixed path without the predictor, the
# FP64 path with the predictor, and the FP64 path at the mixed path's inexact
# inner tolerance (the gain of FP32 alone is mixed_no_predictor over it).
MIXED_NOPRED_OPTS = dict(precond=1, rtol=MIXED_RTOL, atol=0.0, precision=1)
FP64_PRED_OPTS = dict(precond=1, rtol=1e-12, atol=0.0, precision=0, predict=1.0)
FP64_INEXACT_OPTS = dict(precond=1, rtol=MIXED_RTOL, atol=0.0, precision=0)
# the reference's settings: MKL CG unpreconditioned, dpar[0] = dpar[1] = 1e-12
# (solver.h:311-313), Newton tol_iter = 1e-5 and 30 iterations (examples)
REF_OPTS = dict(precond=0, rtol=1e-12, atol=1e-12)


def solve_once(s, gids, prob, opts, P):
    s.restart()
    t0 = time.perf_counter()
    cg = newton = 0
    for step in prob.steps:
        P.apply_step(s, gids, step)
        ok, st = s.solve_step(prob.undamaged, prob.max_iter, NEWTON_TOL, **opts)
        if not ok:
            raise RuntimeError("a load step did not converge")
        cg += st[0]
        newton += st[2]
    s.synchronize()
    return time.perf_counter() - t0, cg, newton


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs), q))


def cpu_baseline(prob, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from lpmg import problems as P

    oracle.set_threads(threads)
    t0 = time.perf_counter()
    o, gids = P.build(oracle.Oracle, prob)
    o.forward()
    setup = time.perf_counter() - t0
    n_sample = 3
    t0 = time.perf_counter()
    cg, ok = 0, True
    for step in prob.steps[:n_sample]:
        P.apply_step(o, gids, step)
        ok_k, st = o.solve_step(prob.undamaged, prob.max_iter, 1e-5, **REF_OPTS)
        ok = ok and ok_k
        cg += int(st[0])
    t_step = (time.perf_counter() - t0) / n_sample
    return dict(value=1.0 / (t_step * len(prob.steps)), unit="solves/s", cores=threads, kind="port",
                sample=f"load steps 1-{n_sample} of {len(prob.steps)} of the same workload (assembly + "
                       f"Newton with unpreconditioned CG, rtol=atol=1e-12 on squared norms, tol_iter=1e-5), "
                       f"{cg} CG iterations, {t_step:.2f} s per load step, scaled to {len(prob.steps)} "
                       f"steps; setup {setup:.2f} s not included",
                step_s=t_step, converged=bool(ok))


# CT serving stream: Poisson arrivals (50 requests/s), particle counts
# log-uniform in [5k, 250k], a/W uniform in [0.2, 0.5], 30 load steps, critical
# bond stretch 1e-3.  Newton tolerance 1e-9: FP64 roundoff in the absolute
# particle coordinates puts a floor near 2e-10 of the reaction on the residual at
# 1e5 particles, which the reference's arithmetic shares.
SERVE_RATE = 50.0
SERVE_SEED = 1
SERVE_TOL = 1e-9
SERVE_FUSE_ROWS = 50000
SERVE_SLOTS = 16
SERVE_OPENING = 7.5e-6  # pin opening per load step, in units of W
# Served path: FP32 inner solves and the load-step predictor; the server runs
# the requests of at most SERVE_FUSE_ROWS particles in fused Jacobi-CG launches
# and the larger ones with the multigrid CG on their own streams (precond < 0:
# the server's choice).
# The CT stream keeps the inner tolerance at 1e-4: at 3e-4 its served
# displacements moved 2.0e-8 from the plain FP64 ones (Newton tolerance 1e-9,
# whose floor on these plates is about that), at 1e-4 1.6e-9.
SERVED_OPTS = dict(precond=-1, rtol=1e-4, atol=0.0, precision=1, predict=1.0)


def ct_stream(n_req, seed=SERVE_SEED):
    """The first n_req requests of the seeded stream (each request draws its
    gap, size and a/W in turn, so a shorter replay is a prefix of a longer one)."""
    rng = np.random.default_rng(seed)
    gaps, sizes, aw = [], [], []
    for _ in range(n_req):
        gaps.append(rng.exponential(1.0 / SERVE_RATE))
        sizes.append(int(np.exp(rng.uniform(np.log(5e3), np.log(2.5e5)))))
        aw.append(rng.uniform(0.2, 0.5))
    arrivals = np.cumsum(gaps) - gaps[0]
    return arrivals, np.array(sizes), np.array(aw)


def replay(probs, arrivals, opts, slots, fuse):
    """Submit each request at its arrival time; latency = submit -> finished."""
    from lpmg import solver_opts
    from lpmg.serve import Request, Server
    o = solver_opts(**opts)
    reqs = [Request(p, o, out_fields=("XYZ", "BDAMAGE"), sizes={"BDAMAGE": len(p.extra["bonds"][0])})
            for p in probs]
    with Server(max_active=slots, batching=fuse) as srv:
        t0 = time.perf_counter()
        tickets = []
        for r, t in zip(reqs, arrivals):
            dt = t - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            tickets.append(srv.submit(r))
        results = [None] * len(reqs)
        for k, tk in enumerate(tickets):
            while results[k] is None:
                results[k] = srv.wait(tk, timeout=20.0)
                if results[k] is None:  # report the load steps the running requests have finished
                    done = [int((q.step_stats[2::3] > 0).sum()) for q in reqs]
                    print(f"#   {time.perf_counter() - t0:.0f} s: load steps done per request {done}",
                          file=sys.stderr, flush=True)
        makespan = time.perf_counter() - t0
        info = srv.info()
    bad = [st.get("error") for rc, st in results if rc != 0]
    if bad:
        raise RuntimeError(f"served requests failed: {bad}")
    lat = [st["total_ms"] for _, st in results]
    return reqs, results, dict(value=len(reqs) / makespan, unit="solves/s", makespan_s=makespan,
                               p50_ms=percentile(lat, 50), p99_ms=percentile(lat, 99),
                               cg_iterations=float(sum(st["cg"] for _, st in results)),
                               peak_running=info["peak_running"], max_batch=info["max_batch"],
                               fused_chunks=info["chunks"], graphs=info["graphs"])


def serving_bench(n_req, plain=True):
    """The CT request stream replayed twice on the GPU: the plain FP64 path
    (one request at a time on its own handle, Jacobi-PCG at the reference's
    CG tolerance) and the served mixed path (16 slots, FP32 inner solves, fused
    Jacobi-CG launches for requests up to 50k particles, multigrid CG for the
    larger ones)."""
    from lpmg import problems as P
    arrivals, sizes, aw = ct_stream(n_req)
    probs = [P.ct_plate(int(n), float(a), steps=30, opening=SERVE_OPENING, crit=1e-3, tol=SERVE_TOL)
             for n, a in zip(sizes, aw)]
    print(f"# serving: {n_req} CT requests, particles {[len(p.xyz) for p in probs]}", file=sys.stderr, flush=True)
    rmx, res_mx, served = replay(probs, arrivals, SERVED_OPTS, SERVE_SLOTS, SERVE_FUSE_ROWS)
    print(f"# serving mixed: {served}", file=sys.stderr, flush=True)
    if not plain:
        return dict(particles=[len(p.xyz) for p in probs], served_mixed=served)
    r64, _, plain = replay(probs, arrivals, FP64_OPTS, 1, False)
    print(f"# serving plain FP64: {plain}", file=sys.stderr, flush=True)
    errs, flags_equal = [], 0
    for p, a, b in zip(probs, r64, rmx):
        u64 = a.out["XYZ"].reshape(-1, 3) - p.xyz
        umx = b.out["XYZ"].reshape(-1, 3) - p.xyz
        errs.append(float(np.linalg.norm(umx - u64) / np.linalg.norm(u64)))
        flags_equal += int(np.array_equal(a.out["BDAMAGE"] > 0.5, b.out["BDAMAGE"] > 0.5))
    return dict(workload=f"CT stream: first {n_req} requests (seed {SERVE_SEED}, Poisson {SERVE_RATE:g}/s, particles "
                         f"log-uniform 5k-250k, a/W U[0.2,0.5], 30 load steps of pin opening {SERVE_OPENING:g} W, "
                         f"critical bond stretch 1e-3, Newton tol {SERVE_TOL:g})",
                particles=[len(p.xyz) for p in probs], plain_fp64=plain, served_mixed=served,
                speedup_vs_plain_fp64=served["value"] / plain["value"],
                max_disp_rel_err_vs_fp64=max(errs), requests_with_identical_damage_flags=flags_equal)


def spmv_configs():
    """Back-to-back SpMV (50 launches each, FP64 then FP32) on the largest
    system of each configuration, assembled at rest: the hex plate 100x100, SC
    40^3, the BCC J2 bar 30x30x60 cells, a 250k-particle CT plate and the BCC
    polycrystal 80^3 cells.  GB/s count the algorithmic bytes (matrix values
    and block columns once, row pointers, x and y once); the FP32 matrices up to
    the CT plate fit the 256 MB Infinity Cache, the polycrystal's (1.6 GB)
    does not."""
    import lpmg
    from lpmg import problems as P
    out = []
    for name, mk in [("hex_plate_100x100", lambda: P.hex_plate()), ("sc_cube_40", lambda: P.sc_cube(n=40)),
                     ("bcc_j2_bar_30x30x60", lambda: P.bcc_j2_bar()),
                     ("ct_plate_250k", lambda: P.ct_plate(250000, 0.35)),
                     ("bcc_polycrystal_80", lambda: P.bcc_polycrystal(80, 200))]:
        prob = mk()
        s, _ = P.build(lpmg.System, prob)
        s.forward()
        s.assemble(precision=1)
        rec = dict(workload=name, particles=s.N, bsr_blocks=s.nnzb)
        for prec, which in (("fp64", 0), ("fp32", 1)):
            s.time_kernel(which, reps=5)
            ms, nbytes = s.time_kernel(which, reps=50)
            rec[prec] = dict(us=ms * 1e3, algorithmic_bytes=nbytes, GBps=nbytes / (ms * 1e-3) / 1e9,
                             frac_of_hbm_peak=nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS)
        print(f"# {name}: {rec}", file=sys.stderr, flush=True)
        out.append(rec)
        s.close()
    return dict(metric="back-to-back SpMV rate per configuration", results=out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed solves")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=40, help="particles per cube edge")
    ap.add_argument("--load-steps", type=int, default=20,
                    help="load steps per solve (fewer only to shorten profiler runs)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--serve-requests", type=int, default=0,
                    help="replay the first N requests of the CT serving stream through the plain FP64 "
                         "path and the served mixed path (minutes per large request on the plain path)")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the CT stream replay (--serve-requests N, at least 1)")
    ap.add_argument("--no-serve-plain", action="store_true",
                    help="with --serve-only: skip the plain FP64 replay (profiling the served path)")
    ap.add_argument("--spmv-configs", action="store_true",
                    help="only the back-to-back SpMV rates (FP64 and FP32) on every configuration's "
                         "largest system, one JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every solve of this workload runs on a single MI355X "
              "(multi-GPU sharding is described in DESIGN.md, not run).", file=sys.stderr)
        sys.exit(2)

    import lpmg
    from lpmg import problems as P

    if args.serve_only:
        print(json.dumps(serving_bench(max(1, args.serve_requests), plain=not args.no_serve_plain)))
        return
    if args.spmv_configs:
        print(json.dumps(spmv_configs()))
        return

    prob = P.sc_cube(n=args.n, steps=20, total=0.05, E=70e3, nu=0.33)
    prob.steps = prob.steps[:args.load_steps]
    # the first handle of the process also pays the HIP runtime's start and the
    # code object's load; a second build of the same system is the setup a
    # running server pays per request
    t0 = time.perf_counter()
    s, gids = P.build(lpmg.System, prob)
    s.synchronize()
    setup_ms = (time.perf_counter() - t0) * 1e3
    s.close()
    t0 = time.perf_counter()
    s, gids = P.build(lpmg.System, prob)
    s.synchronize()
    setup_warm_ms = (time.perf_counter() - t0) * 1e3
    print(f"# setup {setup_ms:.1f} ms (first handle), {setup_warm_ms:.1f} ms (next): {s.N} particles, "
          f"{s.nbond} bonds, {s.nnzb} blocks", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        solve_once(s, gids, prob, FP64_OPTS, P)
    times, cgs, newtons = [], [], []
    for k in range(args.steps):
        t, cg, nw = solve_once(s, gids, prob, FP64_OPTS, P)
        times.append(t)
        cgs.append(cg)
        newtons.append(nw)
        print(f"# fp64 solve {k}: {t * 1e3:.1f} ms, {cg:.0f} CG iterations, {nw:.0f} Newton", file=sys.stderr,
              flush=True)
    u64 = s.get("XYZ") - s.get("XYZ_INITIAL")
    f64 = s.get("PIN")
    # one more solve of the same workload with start/stop events attached to
    # every CG SpMV dispatch on the system's stream (the timed solves replay
    # graphs, where per-kernel events cannot be placed)
    s.profile(True)
    t_prof, _, _ = solve_once(s, gids, prob, FP64_OPTS, P)
    prof = s.profile_read()
    s.profile(False)
    spmv_ms = prof["spmv_ms"] / max(prof["spmv_launches"], 1)
    spmv_bytes = prof["spmv_bytes_fp64"]
    achieved = spmv_bytes / (spmv_ms * 1e-3) / 1e9

    mixed = None
    if not args.no_mixed:
        for _ in range(args.warmup):
            solve_once(s, gids, prob, MIXED_OPTS, P)
        mt, mcg = [], []
        for k in range(args.steps):
            t, cg, nw = solve_once(s, gids, prob, MIXED_OPTS, P)
            mt.append(t)
            mcg.append(cg)
            print(f"# mixed solve {k}: {t * 1e3:.1f} ms, {cg:.0f} CG iterations, {nw:.0f} Newton",
                  file=sys.stderr, flush=True)
        um = s.get("XYZ") - s.get("XYZ_INITIAL")
        fm = s.get("PIN")
        s.profile(True)
        solve_once(s, gids, prob, MIXED_OPTS, P)
        mprof = s.profile_read()
        s.profile(False)
        m_spmv_ms = mprof["spmv_ms"] / max(mprof["spmv_launches"], 1)

        def part(opts):
            runs = [solve_once(s, gids, prob, opts, P) for _ in range(max(1, args.steps))]
            u = s.get("XYZ") - s.get("XYZ_INITIAL")
            return dict(value=len(runs) / sum(t for t, _, _ in runs), unit="solves/s",
                        cg_iterations=float(np.mean([c for _, c, _ in runs])),
                        newton_iterations=float(np.mean([n for _, _, n in runs])),
                        disp_rel_err_vs_fp64=float(np.linalg.norm(u - u64) / np.linalg.norm(u64)))

        parts = dict(mixed_no_predictor=part(MIXED_NOPRED_OPTS), fp64_predictor=part(FP64_PRED_OPTS),
                     fp64_same_inner_tol=part(FP64_INEXACT_OPTS))
        for k, v in parts.items():
            print(f"# {k}: {1e3 / v['value']:.1f} ms, {v['cg_iterations']:.0f} CG iterations, "
                  f"{v['newton_iterations']:.0f} Newton", file=sys.stderr, flush=True)
        mixed = dict(value=len(mt) / sum(mt), unit="solves/s", p50_ms=percentile(mt, 50) * 1e3,
                     p99_ms=percentile(mt, 99) * 1e3, cg_iterations=float(np.mean(mcg)),
                     speedup_vs_fp64=(len(mt) / sum(mt)) / (len(times) / sum(times)),
                     disp_rel_err_vs_fp64=float(np.linalg.norm(um - u64) / np.linalg.norm(u64)),
                     force_rel_err_vs_fp64=float(np.linalg.norm(fm - f64) / np.linalg.norm(f64)),
                     spmv_fp32_ms=m_spmv_ms,
                     spmv_fp32_GBps=mprof["spmv_bytes_fp32"] / (m_spmv_ms * 1e-3) / 1e9,
                     inner_rtol_sq=MIXED_OPTS["rtol"], predictor=MIXED_OPTS["predict"],
                     parts=parts)

    asm_ms, asm_bytes = s.time_kernel(2, reps=10)
    fwd_ms, fwd_bytes = s.time_kernel(3, reps=20)
    copy_ms, copy_bytes = s.time_kernel(4, reps=10)
    sp64_ms, sp64_bytes = s.time_kernel(0, reps=50)
    sp32_ms, sp32_bytes = float("nan"), 0.0
    if not args.no_mixed:
        s.assemble(precision=1)  # the FP64-only solves above left no current FP32 copy
        sp32_ms, sp32_bytes = s.time_kernel(1, reps=50)
    s.close()
    serving = serving_bench(args.serve_requests) if args.serve_requests > 0 else None
    cpu = None if args.no_cpu else cpu_baseline(prob, args.cpu_threads)

    value = len(times) / sum(times)
    chunks = min(4, max(2, -(-int(s.max_nconn) // 16)))  # 16-block chunks per pass over a row
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sum(times) / len(times) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"SC {args.n}^3 elastic cube, FP64 drop-in path (BSR3 SpMV + Jacobi-PCG), "
                               f"20 load steps of 0.0025 mm, r=0.1 mm, E=70 GPa, nu=0.33",
                   "particles": s.N, "dof": s.dim * s.N, "bonds": s.nbond, "bsr_blocks": s.nnzb,
                   "load_steps": len(prob.steps), "newton_tol": NEWTON_TOL, "cg_rtol_sq": FP64_OPTS["rtol"]},
        "p50_ms": percentile(times, 50) * 1e3,
        "p99_ms": percentile(times, 99) * 1e3,
        "setup_ms": setup_ms,
        "setup_warm_ms": setup_warm_ms,
        "cg_iterations_per_solve": float(np.mean(cgs)),
        "newton_iterations_per_solve": float(np.mean(newtons)),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": f"cg_spmv_kernel<3,double,{chunks}>", "avg_ms": spmv_ms,
                     "launches": prof["spmv_launches"], "algorithmic_bytes": spmv_bytes,
                     "timing": "start/stop events attached to every CG SpMV dispatch (hipExtLaunchKernelGGL) "
                               f"of one extra solve of the workload ({t_prof * 1e3:.0f} ms with events)"},
        "assembly": {"ms": asm_ms, "algorithmic_bytes": asm_bytes, "GBps": asm_bytes / (asm_ms * 1e-3) / 1e9},
        "forward": {"ms": fwd_ms, "algorithmic_bytes": fwd_bytes, "GBps": fwd_bytes / (fwd_ms * 1e-3) / 1e9},
        "stream_copy": {"ms": copy_ms, "bytes": copy_bytes, "GBps": copy_bytes / (copy_ms * 1e-3) / 1e9},
        "spmv_back_to_back": {"fp64_ms": sp64_ms, "fp64_GBps": sp64_bytes / (sp64_ms * 1e-3) / 1e9,
                              "fp32_ms": sp32_ms, "fp32_GBps": sp32_bytes / (sp32_ms * 1e-3) / 1e9},
        "mixed": mixed,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    traffic_file = os.path.join(ROOT, "profiles", "spmv_fp64_sc40_traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            out["roofline"]["traffic"] = json.load(f).get("hbm_bytes_per_launch")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
