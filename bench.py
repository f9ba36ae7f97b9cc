#!/usr/bin/env python3
"""Benchmark of the MI355X quasi-static Newton step (LPM-CPP's SolverStatic path).

Headline (`value`): served mixed-precision throughput on the CT request stream --
seeded Poisson arrivals at 50 requests/s of 2D hexagonal compact-tension
specimens, particles log-uniform in 5k-250k, a/W uniform in 0.2-0.5, critical
bond stretch 1e-3, 30 load steps each.  One bench step is one request of that
stream: `--steps K` requests are replayed at their arrival times through the
serving runtime (8 worker slots, device-resident state, memory-gated
admission, fused Jacobi-CG launches for requests up to 10k particles, multigrid
CG for larger 2D ones, FP32 stiffness and inner solves inside the FP64 Newton
iteration, load-step predictor); `--warmup W` requests of a separate small
warm-up stream run first.  value = K / makespan; p50/p99 are per-request
latencies (submit -> finished).

Beside it, in the same JSON line:
  serving.plain_comparison  the first P requests (P = min(K, --plain-requests))
                            through the plain FP64 GPU path (one request at a
                            time on its own handle, FP64 stiffness and
                            Jacobi-PCG stopped where the reference stops MKL's
                            CG, the reference's load-step start) and through
                            the served path: speed-up, largest displacement
                            error against the FP64 result, identical damage flags
  serving.kernels           per-iteration times of the served solvers (FP64/FP32
                            Jacobi and multigrid CG, the build's own events over
                            fixed-length solves) and back-to-back SpMVs on the
                            stream's largest request and one near 20k particles
  polycrystal               the 1.0M-particle BCC polycrystal (80^3 cells, 200
                            grains, cubic C11/C12/C44): FP64 vs mixed, error,
                            and the SpMV on a matrix no cache holds
  roofline                  the CG SpMV of that polycrystal's FP64 solve, timed
                            live with start/stop events attached to every CG
                            SpMV dispatch of one load step
  sc_cube                   the SC 40^3 FP64 drop-in workload: solves/s, the
                            mixed path and its parts, its CG SpMV roofline,
                            assembly and forward-pass rates
  cpu_baseline              the oracle (the FP64 OpenMP restatement of the
                            reference) on the box's host cores: the stream's
                            smallest request among the first K in full, plus
                            the hexagonal plate 100x100 in full and the SC cube's
                            first load steps, each next to the GPU on the same
                            work
Sections after the headline check the remaining time (--time-budget) and are
reported as skipped rather than overrunning it.
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
T0 = time.perf_counter()


def log(msg):
    print(f"# [{time.perf_counter() - T0:6.1f} s] {msg}", file=sys.stderr, flush=True)


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs), q))


# ---------------------------------------------------------------- CT stream

# Newton tolerance 1e-9: FP64 roundoff in the absolute particle coordinates puts
# a floor near 2e-10 of the reaction on the residual at 1e5 particles, which the
# reference's arithmetic shares.
SERVE_RATE = 50.0
SERVE_SEED = 1
WARM_SEED = 2
SERVE_TOL = 1e-9
SERVE_FUSE_ROWS = 10000
SERVE_SLOTS = 8
SERVE_OPENING = 7.5e-6  # pin opening per load step, in units of W
# Served path: FP32 inner solves at ||r||^2 <= 1e-5 ||r0||^2 and the load-step
# predictor; precond < 0 lets the server pick fused Jacobi (<= 10k particles)
# or smoothed-aggregation multigrid (larger 2D requests).  At 3e-4 the served
# displacements moved 2.0e-8 from the plain FP64 ones (the Newton tolerance's
# floor), at 1e-4 1.6e-9.  With the multigrid CG, 1e-5 takes fewer Newton
# iterations per damage pass than 1e-4 (90 against 117 on the 205,645-particle
# request) and fewer CG iterations in all: the 20-request replay served 0.958
# solves/s against 0.839 (3e-5: 0.920; 1e-6 over-solves); with the aggregates
# split along the broken bonds 1.353 against 1.251 (3e-5: 1.360, 3e-6: 1.301,
# 1e-6: 1.259).  Slots and fusion limit: 8 and 10k at 1.606 solves/s against
# 12 slots 1.512, 16 slots 1.476, a 20k limit 1.574, a 5k limit 1.608.
SERVED_OPTS = dict(precond=-1, rtol=1e-5, atol=0.0, precision=1, predict=1.0)
# Plain FP64 GPU path (the drop-in): FP64 stiffness, Jacobi-PCG stopped at
# ||r||^2 <= 1e-12 ||r0||^2 (the reference's dpar[0], solver.h:312), the
# reference's load-step start.
PLAIN_OPTS = dict(precond=1, rtol=1e-12, atol=0.0, precision=0)


def ct_stream(n_req, seed=SERVE_SEED, lo=5e3, hi=2.5e5):
    """The first n_req requests of a seeded stream (each request draws its gap,
    size and a/W in turn, so a shorter replay is a prefix of a longer one)."""
    rng = np.random.default_rng(seed)
    gaps, sizes, aw = [], [], []
    for _ in range(n_req):
        gaps.append(rng.exponential(1.0 / SERVE_RATE))
        sizes.append(int(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        aw.append(rng.uniform(0.2, 0.5))
    arrivals = np.cumsum(gaps) - gaps[0]
    return arrivals, np.array(sizes), np.array(aw)


def ct_problems(sizes, aw):
    from lpmg import problems as P
    return [P.ct_plate(int(n), float(a), steps=30, opening=SERVE_OPENING, crit=1e-3, tol=SERVE_TOL)
            for n, a in zip(sizes, aw)]


def replay(probs, arrivals, opts, slots, fuse, label):
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
                if results[k] is None:  # progress: load steps each request has finished
                    done = [int((q.step_stats[2::3] > 0).sum()) for q in reqs]
                    log(f"{label}: load steps done per request {done}")
        makespan = time.perf_counter() - t0
        info = srv.info()
    bad = [st.get("error") for rc, st in results if rc != 0]
    if bad:
        raise RuntimeError(f"{label}: requests failed: {bad}")
    lat = [st["total_ms"] for _, st in results]
    stats = dict(value=len(reqs) / makespan, unit="solves/s", requests=len(reqs), makespan_s=makespan,
                 p50_ms=percentile(lat, 50), p99_ms=percentile(lat, 99),
                 cg_iterations=float(sum(st["cg"] for _, st in results)),
                 newton_iterations=float(sum(st["newton"] for _, st in results)),
                 peak_running=info["peak_running"], max_batch=info["max_batch"],
                 fused_chunks=info["chunks"], graphs=info["graphs"],
                 latency_ms=[round(x, 1) for x in lat])
    log(f"{label}: {len(reqs)} requests, {stats['value']:.4f} solves/s, makespan {makespan:.1f} s, "
        f"p50 {stats['p50_ms'] / 1e3:.1f} s, p99 {stats['p99_ms'] / 1e3:.1f} s")
    return reqs, results, stats


def compare(probs, ref_reqs, reqs):
    """Largest relative L2 displacement error against the reference results and
    the number of requests whose broken-bond sets are identical."""
    errs, same = [], 0
    for p, a, b in zip(probs, ref_reqs, reqs):
        ua = a.out["XYZ"].reshape(-1, 3) - p.xyz
        ub = b.out["XYZ"].reshape(-1, 3) - p.xyz
        errs.append(float(np.linalg.norm(ub - ua) / np.linalg.norm(ua)))
        same += int(np.array_equal(a.out["BDAMAGE"] > 0.5, b.out["BDAMAGE"] > 0.5))
    return errs, same


def serving_bench(K, W, P, plain=True):
    arrivals, sizes, aw = ct_stream(max(K, P))
    probs = ct_problems(sizes, aw)
    log(f"CT stream: {len(probs)} requests, particles {[len(p.xyz) for p in probs]}")
    if W > 0:  # warm-up: small requests of another seeded stream (code objects, graphs, pools)
        wa, ws, ww = ct_stream(W, seed=WARM_SEED, lo=5e3, hi=2e4)
        replay(ct_problems(ws, ww), wa, SERVED_OPTS, SERVE_SLOTS, SERVE_FUSE_ROWS, "warm-up")
    rs, res_s, served = replay(probs[:K], arrivals[:K], SERVED_OPTS, SERVE_SLOTS, SERVE_FUSE_ROWS,
                               f"served mixed, first {K}")
    out = dict(served_mixed=served, particles=[len(p.xyz) for p in probs[:K]])
    out["kernels"] = serving_kernels(probs[:K])
    if plain and P > 0:
        r64, _, plain_st = replay(probs[:P], arrivals[:P], PLAIN_OPTS, 1, False, f"plain FP64, first {P}")
        if P == K:
            rsP, servedP = rs, served
        else:
            rsP, _, servedP = replay(probs[:P], arrivals[:P], SERVED_OPTS, SERVE_SLOTS, SERVE_FUSE_ROWS,
                                     f"served mixed, first {P}")
        errs, same = compare(probs[:P], r64, rsP)
        out["plain_comparison"] = dict(
            requests=P, particles=[len(p.xyz) for p in probs[:P]], plain_fp64=plain_st, served_mixed=servedP,
            speedup_vs_plain_fp64=servedP["value"] / plain_st["value"], max_disp_rel_err_vs_fp64=max(errs),
            disp_rel_err_vs_fp64=errs, requests_with_identical_damage_flags=same)
    return out, probs


def serving_kernels(probs):
    """Per-iteration times of the served solvers, from the build's own events
    (lpmg_time_kernel kinds 5-8, 10, 11: fixed-length solves whose captured chunks are
    replayed as a served solve replays them), and back-to-back SpMVs, on the
    stream's largest request and on the request nearest 20k particles (a fused
    size), loaded by their first load step and assembled at it."""
    import lpmg
    from lpmg import problems as P
    sizes = [len(p.xyz) for p in probs]
    picks = sorted({int(np.argmax(sizes)), int(np.argmin([abs(n - 20000) for n in sizes]))})
    out = []
    for k in picks:
        prob = probs[k]
        s, gids = P.build(lpmg.System, prob)
        s.forward()
        P.apply_step(s, gids, prob.steps[0])
        s.forward()
        s.assemble(precision=1)
        s.update_rr()
        rec = dict(request=k, particles=s.N, bsr_blocks=s.nnzb)
        for name, which in (("spmv_fp64", 0), ("spmv_fp32", 1)):
            s.time_kernel(which, reps=5)
            ms, b = s.time_kernel(which, reps=40)
            rec[name] = dict(us=ms * 1e3, GBps=b / (ms * 1e-3) / 1e9)
        for name, which, reps in (("jacobi_cg_fp64", 5, 320), ("jacobi_cg_fp32", 6, 320),
                                  ("multigrid_cg_fp64", 7, 64), ("multigrid_cg_fp32", 8, 64),
                                  ("sa_multigrid_cg_fp64", 10, 64), ("sa_multigrid_cg_fp32", 11, 64)):
            s.time_kernel(which, reps=16)  # builds the multigrid hierarchy on first use
            ms, b = s.time_kernel(which, reps=reps)
            rec[name] = dict(us_per_iteration=ms * 1e3)
            if b > 0:
                rec[name]["GBps"] = b / (ms * 1e-3) / 1e9
        s.close()
        log(f"served solver kernels, request {k}: {rec}")
        out.append(rec)
    return out


# ---------------------------------------------------------------- polycrystal (1.0M particles)

POLY_TOL = 1e-10
# the server's choice of solver (smoothed-aggregation multigrid for a BCC request above the fusion
# limit); inner tolerance 1e-4 (1.39 s against 1.56 s at 1e-5: the same 34 Newton iterations)
POLY_MIXED = dict(precond=-1, rtol=1e-4, atol=0.0, precision=1, predict=1.0)


def serve_one(prob, opts):
    from lpmg import solver_opts
    from lpmg.serve import Request, Server
    rq = Request(prob, solver_opts(**opts), out_fields=("XYZ",))
    with Server(max_active=1, batching=SERVE_FUSE_ROWS) as srv:
        t0 = time.perf_counter()
        rc, st = srv.wait(srv.submit(rq))
        t = time.perf_counter() - t0
    if rc != 0:
        raise RuntimeError(f"polycrystal request failed: {st.get('error')}")
    return t, st, rq.out["XYZ"].reshape(-1, 3) - prob.xyz


def polycrystal_fp64_profile():
    """The polycrystal's FP64 path only, for profiler runs: build, the first
    load step with events on every CG SpMV, back-to-back SpMV, assembly and a
    plain copy of the matrix values (byte-counter calibration)."""
    import lpmg
    from lpmg import problems as P
    pc = P.bcc_polycrystal(80, 200)
    s, gids = P.build(lpmg.System, pc)
    s.forward()
    P.apply_step(s, gids, pc.steps[0])
    s.profile(True)
    ok, _ = s.solve_step(pc.undamaged, pc.max_iter, POLY_TOL, **PLAIN_OPTS)
    prof = s.profile_read()
    s.profile(False)
    sp64_ms, sp64_b = s.time_kernel(0, reps=20)
    asm_ms, asm_b = s.time_kernel(2, reps=3)
    copy_ms, copy_b = s.time_kernel(4, reps=5)
    rec = dict(converged=bool(ok), cg_spmv_us=prof["spmv_ms"] / max(prof["spmv_launches"], 1) * 1e3,
               cg_spmv_launches=prof["spmv_launches"], spmv_bytes=sp64_b, spmv_back_to_back_us=sp64_ms * 1e3,
               assembly_ms=asm_ms, assembly_bytes=asm_b, copy_us=copy_ms * 1e3, copy_bytes=copy_b,
               particles=s.N, bsr_blocks=s.nnzb)
    s.close()
    return rec


def polycrystal_bench():
    import lpmg
    from lpmg import problems as P
    t0 = time.perf_counter()
    pc = P.bcc_polycrystal(80, 200)
    pc.tol = POLY_TOL
    gen_s = time.perf_counter() - t0
    log(f"polycrystal: {len(pc.xyz)} particles generated in {gen_s:.1f} s")
    tm, stm, um = serve_one(pc, POLY_MIXED)
    log(f"polycrystal mixed: {tm:.2f} s, {stm['cg']:.0f} CG, {stm['newton']:.0f} Newton")
    t64, st64, u64 = serve_one(pc, PLAIN_OPTS)
    log(f"polycrystal FP64: {t64:.2f} s, {st64['cg']:.0f} CG, {st64['newton']:.0f} Newton")
    # the CG SpMV of the FP64 path, timed live over the first load step, and
    # back-to-back SpMV rates on the assembled matrix (2.96 GB FP64, 1.56 GB
    # FP32: neither fits the 256 MB Infinity Cache)
    s, gids = P.build(lpmg.System, pc)
    s.forward()
    P.apply_step(s, gids, pc.steps[0])
    s.profile(True)
    t_prof = time.perf_counter()
    ok, _ = s.solve_step(pc.undamaged, pc.max_iter, POLY_TOL, **PLAIN_OPTS)
    t_prof = time.perf_counter() - t_prof
    prof = s.profile_read()
    s.profile(False)
    if not ok:
        raise RuntimeError("polycrystal instrumented load step did not converge")
    s.assemble(precision=1)
    s.time_kernel(0, reps=3)
    sp64_ms, sp64_b = s.time_kernel(0, reps=20)
    sp32_ms, sp32_b = s.time_kernel(1, reps=20)
    asm_ms, asm_b = s.time_kernel(2, reps=3)
    s.time_kernel(9, reps=2)
    rd_ms, rd_b = s.time_kernel(9, reps=10)
    spmv_ms = prof["spmv_ms"] / max(prof["spmv_launches"], 1)
    rec = dict(
        workload=f"BCC polycrystal 80^3 cells, 200 Voronoi grains (seed 7), {len(pc.xyz)} particles, cubic "
                 f"C11/C12/C44 = 230/135/117 GPa, 10 tension steps of 1e-4, Newton tol {POLY_TOL:g}",
        particles=s.N, bsr_blocks=s.nnzb, generate_s=gen_s,
        fp64=dict(value=1.0 / t64, unit="solves/s", seconds=t64, cg_iterations=st64["cg"],
                  newton_iterations=st64["newton"], device_bytes=st64["bytes"]),
        mixed=dict(value=1.0 / tm, unit="solves/s", seconds=tm, cg_iterations=stm["cg"],
                   newton_iterations=stm["newton"], device_bytes=stm["bytes"], inner_rtol_sq=POLY_MIXED["rtol"],
                   predictor=POLY_MIXED["predict"],
                   solver="served: the server's choice, smoothed-aggregation multigrid CG (option 4) for a BCC "
                          "request above the fusion limit"),
        speedup_mixed_vs_fp64=t64 / tm,
        disp_rel_err_mixed_vs_fp64=float(np.linalg.norm(um - u64) / np.linalg.norm(u64)),
        spmv_back_to_back=dict(fp64_us=sp64_ms * 1e3, fp64_GBps=sp64_b / (sp64_ms * 1e-3) / 1e9,
                               fp64_frac_of_hbm_peak=sp64_b / (sp64_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                               fp32_us=sp32_ms * 1e3, fp32_GBps=sp32_b / (sp32_ms * 1e-3) / 1e9,
                               fp32_frac_of_hbm_peak=sp32_b / (sp32_ms * 1e-3) / 1e9 / HBM_PEAK_GBS),
        assembly=dict(ms=asm_ms, algorithmic_bytes=asm_b, GBps=asm_b / (asm_ms * 1e-3) / 1e9),
        # a plain streaming read of the FP64 values: what a read-dominated
        # kernel can attain on this part
        stream_read=dict(us=rd_ms * 1e3, bytes=rd_b, GBps=rd_b / (rd_ms * 1e-3) / 1e9,
                         spmv_back_to_back_over_read=(sp64_b / (sp64_ms * 1e-3)) / (rd_b / (rd_ms * 1e-3))))
    chunks = min(4, max(2, -(-int(s.max_nconn) // 16)))
    achieved = prof["spmv_bytes_fp64"] / (spmv_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": f"cg_spmv_kernel<3,double,{chunks}>",
            "avg_ms": spmv_ms, "launches": prof["spmv_launches"], "algorithmic_bytes": prof["spmv_bytes_fp64"],
            "workload": "1.0M-particle BCC polycrystal, FP64 Jacobi-PCG, first load step",
            "timing": "start/stop events attached to every CG SpMV dispatch (hipExtLaunchKernelGGL) of the first "
                      f"load step of the FP64 solve ({t_prof * 1e3:.0f} ms with events)",
            # the attainable rate: a plain streaming read of the same FP64 values
            "attainable_read_GBps": rec["stream_read"]["GBps"],
            "frac_of_attainable_read": achieved / rec["stream_read"]["GBps"]}
    s.close()
    log(f"polycrystal roofline: {spmv_ms * 1e3:.1f} us per CG SpMV, {achieved:.0f} GB/s")
    return rec, roof


# ---------------------------------------------------------------- SC 40^3 (the FP64 drop-in)

NEWTON_TOL = 1e-10
# FP64 drop-in: Jacobi-PCG stopped where the reference stops MKL's CG (||r||^2 <=
# 1e-12 ||r0||^2, solver.h:312), Newton tightened to 1e-10 so results meet 1e-8.
FP64_OPTS = dict(precond=1, rtol=1e-12, atol=0.0, precision=0)
# Mixed: FP32 stiffness and CG as an inexact inner solve (||r||^2 <= 3e-4
# ||r0||^2, the fastest of 1e-5 ... 1e-2 on this workload) of the FP64 modified
# Newton, each load step starting from the last step's increment.
MIXED_RTOL = 3e-4
MIXED_OPTS = dict(precond=1, rtol=MIXED_RTOL, atol=0.0, precision=1, predict=1.0)
# Its parts, each measured alone; fp64_predictor_same_inner_tol differs from the
# mixed path only in the precision of the stiffness and the inner solve.
PARTS = dict(mixed_no_predictor=dict(precond=1, rtol=MIXED_RTOL, atol=0.0, precision=1),
             fp64_predictor=dict(precond=1, rtol=1e-12, atol=0.0, precision=0, predict=1.0),
             fp64_same_inner_tol=dict(precond=1, rtol=MIXED_RTOL, atol=0.0, precision=0),
             fp64_predictor_same_inner_tol=dict(precond=1, rtol=MIXED_RTOL, atol=0.0, precision=0, predict=1.0))
# FP64 Jacobi-preconditioned BiCGSTAB at the drop-in's tolerances
BICG_OPTS = dict(precond=3, rtol=1e-12, atol=0.0, precision=0)
# the reference's settings: MKL CG unpreconditioned, dpar[0] = dpar[1] = 1e-12
# (solver.h:311-313), Newton tol_iter = 1e-5 and 30 iterations (examples)
REF_OPTS = dict(precond=0, rtol=1e-12, atol=1e-12)
REF_TOL = 1e-5


def solve_once(s, gids, prob, opts, tol=NEWTON_TOL, steps=None):
    from lpmg import problems as P
    s.restart()
    t0 = time.perf_counter()
    cg = newton = 0
    for step in (prob.steps if steps is None else prob.steps[:steps]):
        P.apply_step(s, gids, step)
        ok, st = s.solve_step(prob.undamaged, prob.max_iter, tol, **opts)
        if not ok:
            raise RuntimeError("a load step did not converge")
        cg += st[0]
        newton += st[2]
    s.synchronize()
    return time.perf_counter() - t0, cg, newton


def sc_fp64_profile(load_steps=2):
    """The SC 40^3 FP64 path only, for profiler runs: a few load steps, one of
    them with events on every CG SpMV, and the kernel timings."""
    import lpmg
    from lpmg import problems as P
    prob = P.sc_cube(n=40, steps=20, total=0.05, E=70e3, nu=0.33)
    s, gids = P.build(lpmg.System, prob)
    t, cg, nw = solve_once(s, gids, prob, FP64_OPTS, steps=load_steps)
    s.profile(True)
    solve_once(s, gids, prob, FP64_OPTS, steps=load_steps)
    prof = s.profile_read()
    s.profile(False)
    rec = dict(load_steps=load_steps, seconds=t, cg_iterations=cg,
               cg_spmv_us=prof["spmv_ms"] / max(prof["spmv_launches"], 1) * 1e3,
               cg_spmv_launches=prof["spmv_launches"])
    for name, which, reps in (("spmv", 0, 50), ("assembly", 2, 10), ("forward", 3, 20), ("copy", 4, 10)):
        ms, b = s.time_kernel(which, reps=reps)
        rec[name] = dict(us=ms * 1e3, bytes=b, GBps=b / (ms * 1e-3) / 1e9)
    s.close()
    return rec


def sc_bicgstab_profile(load_steps=2):
    """The SC 40^3 FP64 path with BiCGSTAB inner solves, a few load steps, for
    profiler runs."""
    import lpmg
    from lpmg import problems as P
    prob = P.sc_cube(n=40, steps=20, total=0.05, E=70e3, nu=0.33)
    s, gids = P.build(lpmg.System, prob)
    solve_once(s, gids, prob, BICG_OPTS, steps=load_steps)
    t, it, nw = solve_once(s, gids, prob, BICG_OPTS, steps=load_steps)
    s.close()
    return dict(load_steps=load_steps, seconds=t, bicgstab_iterations=it, newton_iterations=nw,
                us_per_iteration=t / max(it, 1) * 1e6)


def sc_bench(n_solves):
    import lpmg
    from lpmg import problems as P
    prob = P.sc_cube(n=40, steps=20, total=0.05, E=70e3, nu=0.33)
    t0 = time.perf_counter()
    s, gids = P.build(lpmg.System, prob)
    s.synchronize()
    setup_ms = (time.perf_counter() - t0) * 1e3
    solve_once(s, gids, prob, FP64_OPTS)
    times, cgs, newtons = [], [], []
    for _ in range(n_solves):
        t, cg, nw = solve_once(s, gids, prob, FP64_OPTS)
        times.append(t)
        cgs.append(cg)
        newtons.append(nw)
    log(f"SC 40^3 FP64: {np.mean(times) * 1e3:.1f} ms per solve, {np.mean(cgs):.0f} CG, {np.mean(newtons):.0f} Newton")
    u64 = s.get("XYZ") - s.get("XYZ_INITIAL")
    f64 = s.get("PIN")
    s.profile(True)
    t_prof, _, _ = solve_once(s, gids, prob, FP64_OPTS)
    prof = s.profile_read()
    s.profile(False)
    spmv_ms = prof["spmv_ms"] / max(prof["spmv_launches"], 1)
    achieved = prof["spmv_bytes_fp64"] / (spmv_ms * 1e-3) / 1e9
    solve_once(s, gids, prob, MIXED_OPTS)
    mt, mcg = [], []
    for _ in range(n_solves):
        t, cg, _ = solve_once(s, gids, prob, MIXED_OPTS)
        mt.append(t)
        mcg.append(cg)
    um = s.get("XYZ") - s.get("XYZ_INITIAL")
    fm = s.get("PIN")
    log(f"SC 40^3 mixed: {np.mean(mt) * 1e3:.1f} ms per solve")

    def part(opts):
        runs = [solve_once(s, gids, prob, opts) for _ in range(n_solves)]
        u = s.get("XYZ") - s.get("XYZ_INITIAL")
        return dict(value=len(runs) / sum(t for t, _, _ in runs), unit="solves/s",
                    cg_iterations=float(np.mean([c for _, c, _ in runs])),
                    newton_iterations=float(np.mean([n for _, _, n in runs])),
                    disp_rel_err_vs_fp64=float(np.linalg.norm(u - u64) / np.linalg.norm(u64)))

    parts = {k: part(v) for k, v in PARTS.items()}
    bicg = part(BICG_OPTS)
    # the reference's own settings (unpreconditioned CG, tol_iter 1e-5) on the
    # first three load steps: the GPU side of the CPU comparison
    t_ref, cg_ref, _ = solve_once(s, gids, prob, REF_OPTS, tol=REF_TOL, steps=3)
    asm_ms, asm_bytes = s.time_kernel(2, reps=10)
    fwd_ms, fwd_bytes = s.time_kernel(3, reps=20)
    copy_ms, copy_bytes = s.time_kernel(4, reps=10)
    sp64_ms, sp64_bytes = s.time_kernel(0, reps=50)
    s.assemble(precision=1)
    sp32_ms, sp32_bytes = s.time_kernel(1, reps=50)
    chunks = min(4, max(2, -(-int(s.max_nconn) // 16)))
    value = len(times) / sum(times)
    rec = {
        "workload": "SC 40^3 elastic cube, FP64 drop-in path (BSR3 SpMV + Jacobi-PCG), 20 load steps of 0.0025 mm, "
                    "r=0.1 mm, E=70 GPa, nu=0.33",
        "particles": s.N, "bonds": s.nbond, "bsr_blocks": s.nnzb, "newton_tol": NEWTON_TOL,
        "value": value, "unit": "solves/s", "solves": len(times), "ms_per_solve": sum(times) / len(times) * 1e3,
        "p50_ms": percentile(times, 50) * 1e3, "p99_ms": percentile(times, 99) * 1e3, "setup_ms": setup_ms,
        "cg_iterations_per_solve": float(np.mean(cgs)), "newton_iterations_per_solve": float(np.mean(newtons)),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "kernel": f"cg_spmv_kernel<3,double,{chunks}>",
                     "avg_ms": spmv_ms, "launches": prof["spmv_launches"],
                     "algorithmic_bytes": prof["spmv_bytes_fp64"],
                     "timing": "start/stop events attached to every CG SpMV dispatch of one extra solve "
                               f"({t_prof * 1e3:.0f} ms with events); the 278 MB FP64 matrix is about the size "
                               "of the 256 MB Infinity Cache"},
        "mixed": dict(value=len(mt) / sum(mt), unit="solves/s", p50_ms=percentile(mt, 50) * 1e3,
                      p99_ms=percentile(mt, 99) * 1e3, cg_iterations=float(np.mean(mcg)),
                      speedup_vs_fp64=(len(mt) / sum(mt)) / value,
                      disp_rel_err_vs_fp64=float(np.linalg.norm(um - u64) / np.linalg.norm(u64)),
                      force_rel_err_vs_fp64=float(np.linalg.norm(fm - f64) / np.linalg.norm(f64)),
                      inner_rtol_sq=MIXED_RTOL, predictor=1.0, parts=parts,
                      fp32_share=(len(mt) / sum(mt)) / parts["fp64_predictor_same_inner_tol"]["value"]),
        "bicgstab": dict(bicg, note="FP64 Jacobi-preconditioned BiCGSTAB (solver option 3) at the CG's "
                                   "tolerances; cg_iterations counts BiCGSTAB iterations (two SpMVs each)"),
        "reference_settings_first_3_steps": dict(seconds=t_ref, cg_iterations=cg_ref, newton_tol=REF_TOL,
                                                  opts=REF_OPTS),
        # whole assembly: tangent, symmetrisation, ||diag||, elimination, FP32 copy;
        # bytes = its algorithmic minimum (bond and particle inputs once, the
        # FP64 blocks and block columns once)
        "assembly": {"ms": asm_ms, "algorithmic_bytes": asm_bytes, "GBps": asm_bytes / (asm_ms * 1e-3) / 1e9,
                     "frac_of_hbm_peak": asm_bytes / (asm_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "forward": {"ms": fwd_ms, "algorithmic_bytes": fwd_bytes, "GBps": fwd_bytes / (fwd_ms * 1e-3) / 1e9},
        "stream_copy": {"ms": copy_ms, "bytes": copy_bytes, "GBps": copy_bytes / (copy_ms * 1e-3) / 1e9},
        "spmv_back_to_back": {"fp64_ms": sp64_ms, "fp64_GBps": sp64_bytes / (sp64_ms * 1e-3) / 1e9,
                              "fp32_ms": sp32_ms, "fp32_GBps": sp32_bytes / (sp32_ms * 1e-3) / 1e9},
    }
    s.close()
    return rec, prob


# ---------------------------------------------------------------- CPU baseline

def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    env = os.environ.get("OMP_NUM_THREADS")
    threads = int(env) if env and env.isdigit() and int(env) > 0 else aff
    return dict(cpu_model=model, nproc=os.cpu_count(), affinity=aff, omp_num_threads_env=env, threads=threads,
                mkl="not called: the oracle restates MKL's RCI CG and stopping test in its own FP64 code")


def cpu_run(backend, prob, opts, tol, steps=None):
    from lpmg import problems as P
    if backend.__name__ == "Oracle":  # the oracle solves in FP64 from the reference's start
        opts = {k: v for k, v in opts.items() if k in ("precond", "rtol", "atol", "maxit")}
    s, gids = P.build(backend, prob)
    s.forward()
    t0 = time.perf_counter()
    cg = newton = 0
    ok_all = True
    st_list = []
    if steps is None:
        st_list = P.solve_problem(s, gids, prob, opts)
    else:
        for step in prob.steps[:steps]:
            P.apply_step(s, gids, step)
            ok, st = s.solve_step(prob.undamaged, prob.max_iter, tol, **opts)
            ok_all = ok_all and ok
            st_list.append(st)
    t = time.perf_counter() - t0
    cg = float(sum(st[0] for st in st_list))
    newton = float(sum(st[2] for st in st_list))
    return t, cg, newton, ok_all, s


def cpu_baseline(stream_probs, sc_prob, sc_rec):
    """The oracle on the box's host cores, each sample next to the GPU on the same work."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import lpmg
    from lpmg import problems as P
    info = cpu_info()
    oracle.set_threads(info["threads"])
    # the stream's smallest request among those timed, in full, with the plain
    # path's algorithm (Jacobi-PCG, rtol 1e-12 on squared norms, Newton 1e-9)
    k = int(np.argmin([len(p.xyz) for p in stream_probs]))
    ct = stream_probs[k]
    t_cpu, cg_cpu, nw_cpu, _, o = cpu_run(oracle.Oracle, ct, PLAIN_OPTS, ct.tol)
    u_cpu = o.get("XYZ") - o.get("XYZ_INITIAL")
    o.close()
    t_gpu, cg_gpu, _, _, g = cpu_run(lpmg.System, ct, PLAIN_OPTS, ct.tol)
    u_gpu = g.get("XYZ") - g.get("XYZ_INITIAL")
    g.close()
    log(f"CPU CT request {k} ({len(ct.xyz)} particles): {t_cpu:.2f} s CPU, {t_gpu:.2f} s GPU plain FP64")
    # the hexagonal plate 100x100 in full at the reference's settings
    hexp = P.hex_plate()
    th_cpu, cgh_cpu, _, okh, o = cpu_run(oracle.Oracle, hexp, REF_OPTS, REF_TOL, steps=len(hexp.steps))
    o.close()
    th_gpu, cgh_gpu, _, _, g = cpu_run(lpmg.System, hexp, REF_OPTS, REF_TOL, steps=len(hexp.steps))
    g.close()
    log(f"CPU hex plate 100x100: {th_cpu:.2f} s CPU, {th_gpu:.3f} s GPU")
    # the SC cube's first three load steps at the reference's settings
    ts_cpu, cgs_cpu, _, oks, o = cpu_run(oracle.Oracle, sc_prob, REF_OPTS, REF_TOL, steps=3)
    o.close()
    log(f"CPU SC 40^3, 3 load steps: {ts_cpu:.2f} s CPU")
    rec = dict(
        value=1.0 / t_cpu, unit="solves/s", cores=info["threads"], kind="port",
        sample=f"request {k} of the CT stream ({len(ct.xyz)} particles, the smallest of those timed), all 30 load "
               f"steps, with the plain FP64 path's algorithm (Jacobi-PCG, ||r||^2 <= 1e-12 ||r0||^2, Newton tol "
               f"{ct.tol:g}): {t_cpu:.2f} s, {cg_cpu:.0f} CG iterations; the GPU plain FP64 path on the same "
               f"request: {t_gpu:.2f} s",
        gpu_plain_same_request=dict(value=1.0 / t_gpu, seconds=t_gpu, cg_iterations=cg_gpu,
                                    disp_rel_err_vs_cpu=float(np.linalg.norm(u_gpu - u_cpu) / np.linalg.norm(u_cpu))),
        hex_plate_100x100=dict(workload="hexagonal plate 100x100 (10,000 particles), 10 load steps, the reference's "
                                        "settings (unpreconditioned CG, rtol = atol = 1e-12, tol_iter 1e-5)",
                               cpu_s=th_cpu, cpu_cg_iterations=cgh_cpu, gpu_s=th_gpu, gpu_cg_iterations=cgh_gpu,
                               converged=bool(okh), speedup_gpu_vs_cpu=th_cpu / th_gpu),
        sc_cube_first_3_steps=dict(cpu_s=ts_cpu, cpu_cg_iterations=cgs_cpu,
                                   gpu_s=sc_rec["reference_settings_first_3_steps"]["seconds"],
                                   gpu_cg_iterations=sc_rec["reference_settings_first_3_steps"]["cg_iterations"],
                                   converged=bool(oks),
                                   speedup_gpu_vs_cpu=ts_cpu / sc_rec["reference_settings_first_3_steps"]["seconds"],
                                   settings="the reference's (unpreconditioned CG, rtol = atol = 1e-12, "
                                            "tol_iter 1e-5), both sides"),
        **info)
    return rec


# ---------------------------------------------------------------- back-to-back SpMV on every configuration

def spmv_configs():
    """Back-to-back SpMV (50 launches each, FP64 then FP32) on the largest
    system of each configuration, assembled at rest, and a plain streaming read
    of its FP64 values (the attainable read rate for the same bytes)."""
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
        for prec, which in (("fp64", 0), ("fp32", 1), ("read_fp64_values", 9)):
            s.time_kernel(which, reps=5)
            ms, nbytes = s.time_kernel(which, reps=50)
            rec[prec] = dict(us=ms * 1e3, algorithmic_bytes=nbytes, GBps=nbytes / (ms * 1e-3) / 1e9,
                             frac_of_hbm_peak=nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS)
        rec["fp64_spmv_over_read"] = rec["fp64"]["GBps"] / rec["read_fp64_values"]["GBps"]
        log(f"{name}: {rec}")
        out.append(rec)
        s.close()
    return dict(metric="back-to-back SpMV rate per configuration", results=out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6, help="timed steps: requests of the CT stream served")
    ap.add_argument("--warmup", type=int, default=2, help="requests of a small warm-up stream served first")
    ap.add_argument("--plain-requests", type=int, default=6,
                    help="requests of the stream's prefix also run through the plain FP64 path (minutes each "
                         "above 1e5 particles)")
    ap.add_argument("--time-budget", type=float, default=540.0,
                    help="seconds; sections after the headline are skipped when they would not fit")
    ap.add_argument("--sc-solves", type=int, default=3, help="timed solves of each SC 40^3 variant")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-poly", action="store_true")
    ap.add_argument("--no-sc", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["sc", "poly", "serve", "spmv-configs", "sc-fp64", "poly-fp64", "sc-bicgstab"],
                    help="run one section alone and print its record (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every solve of this workload runs on a single MI355X "
              "(multi-GPU sharding is described in DESIGN.md, not run).", file=sys.stderr)
        sys.exit(2)
    if args.only == "spmv-configs":
        print(json.dumps(spmv_configs()))
        return
    if args.only == "sc-fp64":
        print(json.dumps(sc_fp64_profile()))
        return
    if args.only == "poly-fp64":
        print(json.dumps(polycrystal_fp64_profile()))
        return
    if args.only == "sc-bicgstab":
        print(json.dumps(sc_bicgstab_profile()))
        return
    if args.only == "sc":
        print(json.dumps(sc_bench(args.sc_solves)[0]))
        return
    if args.only == "poly":
        rec, roof = polycrystal_bench()
        print(json.dumps(dict(polycrystal=rec, roofline=roof)))
        return
    if args.only == "serve":
        print(json.dumps(serving_bench(args.steps, args.warmup, min(args.steps, args.plain_requests),
                                       plain=not args.no_plain)[0]))
        return

    K, W = args.steps, args.warmup
    P = 0 if args.no_plain else min(K, args.plain_requests)
    serving, probs = serving_bench(K, W, P)
    served = serving["served_mixed"]
    skipped = []

    def left():
        return args.time_budget - (time.perf_counter() - T0)

    poly = roof = None
    if not args.no_poly:
        if left() > 80:
            poly, roof = polycrystal_bench()
        else:
            skipped.append("polycrystal (time budget)")
    sc = sc_prob = None
    if not args.no_sc:
        if left() > 40:
            sc, sc_prob = sc_bench(args.sc_solves)
        else:
            skipped.append("sc_cube (time budget)")
    cpu = None
    if not args.no_cpu:
        if sc is not None and left() > 45:
            cpu = cpu_baseline(probs[:K], sc_prob, sc)
        else:
            skipped.append("cpu_baseline (time budget or no SC record)")
    if roof is None and sc is not None:
        roof = dict(sc["roofline"], traffic=None)
    traffic_file = os.path.join(ROOT, "profiles", "spmv_fp64_polycrystal_traffic.json")
    if roof is not None and poly is not None and os.path.exists(traffic_file):
        with open(traffic_file) as f:
            roof["traffic"] = json.load(f).get("hbm_bytes_per_launch")
    cmp = serving.get("plain_comparison")
    out = {
        "metric": METRIC,
        "value": served["value"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": served["makespan_s"] / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"CT request stream (config 4): first {K} requests, seeded Poisson arrivals at "
                               f"{SERVE_RATE:g}/s (seed {SERVE_SEED}), 2D hex CT specimens of 5k-250k particles "
                               f"(log-uniform), a/W U[0.2,0.5], critical bond stretch 1e-3, 30 load steps of pin "
                               f"opening {SERVE_OPENING:g} W, Newton tol {SERVE_TOL:g}; served mixed precision "
                               f"(FP32 stiffness/inner CG, FP64 residual and refinement)",
                   "particles": serving["particles"], "slots": SERVE_SLOTS, "fuse_max_rows": SERVE_FUSE_ROWS,
                   "inner_rtol_sq": SERVED_OPTS["rtol"]},
        "p50_ms": served["p50_ms"],
        "p99_ms": served["p99_ms"],
        "speedup_vs_plain_fp64": cmp["speedup_vs_plain_fp64"] if cmp else None,
        "max_disp_rel_err_vs_fp64": cmp["max_disp_rel_err_vs_fp64"] if cmp else None,
        "roofline": roof,
        "serving": serving,
        "polycrystal": poly,
        "sc_cube": sc,
        "cpu_baseline": cpu,
        "skipped": skipped,
        "wall_s": time.perf_counter() - T0,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
