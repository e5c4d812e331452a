"""Sweep benchmark of the A-Phi path on one MI355X.

Workload: the mixed-precision configuration -- a 32^3-hex Kuhn cube (196,608 tets,
464,910 unknowns) with an off-centre 6x3 conductor bar, 64 log-spaced frequencies
from 1 Hz to 100 MHz (full-wave), 16 seeded N(0,1) complex right-hand sides per
frequency, swept in ascending order like the reference's solve_mesh loop. One
step = one frequency point: device refill, equilibration,
complex64 LDL^T on MFMA, 16-RHS solve, FP64 GMRES refinement (the FP32 factor as
preconditioner) until every right-hand side's backward error is <= 1e-12. Where
the points already solved show the FP32 factor would need more refinement
solves than a complex128 factor costs, the point is factored in complex128
instead (aphi_amd.PrecisionPolicy); the counts are reported. The plain FP64 path
(complex128 factor, no refinement) is timed on --fp64-steps of the sweep's points,
spread over the range (its cost does not depend on the frequency).

--pipeline points are in flight at once (default 2), each on its own context and
copy of the mesh, so that one point's refinement solves (latency-bound chains of
small launches) run beside the next point's factorization (MFMA-bound). `value`
and the plain FP64 rate are measured that way; the same sweep one point at a time
is reported beside them (`sequential`), with the latency of a point that has the
device to itself.

Prints one JSON line. Inputs are resident in HBM before the timed region.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hybrid-aphi-solver_amd"))
sys.path.insert(0, ROOT)

METRIC = "sweep solves/s + p99 latency, FP32-MFMA+IR vs FP64 LDLᵀ at 1e-12 backward error"
MFMA_FP32_PEAK_TFLOPS = 157.3
MFMA_FP64_PEAK_TFLOPS = 78.6


RHS_SEED = 20261203


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def update_traffic(n):
    """HBM bytes per k_update<float> launch on this workload, from the committed
    FETCH_SIZE / WRITE_SIZE counter runs (profiles/k_update_f32_traffic.json),
    each scaled by the same counter's reading for a copy of known size; None for
    other mesh sizes or without the file."""
    path = os.path.join(ROOT, "profiles", "k_update_f32_traffic.json")
    if n != 32 or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["traffic_bytes_per_launch"]


def cpu_baseline(A, prob, mesh, freq):
    """The oracle's restatement of the reference's scalar up-looking LDL^T
    (src/factorization.cpp:109-218), one thread, on this box's host: the whole
    factorization of the 6^3 configuration and of a 12^3 cube (no extrapolation
    inside either), the rate of the latter applied to this workload's flop count,
    and the same factorization run on P threads at once for the throughput of P
    independent frequency points (capped by host memory for this workload's factor)."""
    import concurrent.futures as cf
    import platform
    from oracle import oracle as O
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), model)
    except OSError:
        pass
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()

    def full(n):
        p = A.Problem(A.CubeConfig(n=n, bar=(n // 3, n // 3 + 2, n // 3, n // 3 + 2)))
        v, _ = O.refill(p, 2 * np.pi * freq)
        return p, v

    p6, v6 = full(6)
    s6 = O.factor_sample(p6, v6, 1e9)
    p12, v12 = full(12)
    s12 = O.factor_sample(p12, v12, 1e9)
    rate = s12["flops_done"] / s12["seconds"]
    vals, _ = mesh.system()
    s32 = O.factor_sample(prob, vals, 3.0)  # its equilibration, and a first look at its leaf columns
    est = s32["flops_total"] / rate + s32["equilibrate_s"]
    # P independent points at once (the C factorization runs outside the
    # interpreter lock): aggregate rate over the 8^3 cube
    P = max(1, min(16, cores))
    p8, v8 = full(8)
    single = O.factor_sample(p8, v8, 1e9)
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=P) as ex:
        runs = list(ex.map(lambda _: O.factor_sample(p8, v8, 1e9), range(P)))
    wall = time.perf_counter() - t0
    eff = sum(r["flops_done"] for r in runs) / wall / (single["flops_done"] / single["seconds"])
    factor_bytes = 20.0 * prob.counts["predicted_nnz"] + 40.0 * prob.counts["nnz"]  # values + row indices, matrix
    try:
        host = os.sysconf("SC_PAGE_SIZE") * os.sysconf("SC_PHYS_PAGES")
    except (ValueError, OSError):
        host = 256e9
    p_mem = max(1, int(min(host, 270 * 2 ** 30) // factor_bytes))
    parallel = min(eff, p_mem)
    return {"value": 1.0 / est, "unit": "frequency points/s", "cores": 1, "kind": "port",
            "cpu_model": model, "nproc": cores,
            "config1_point_s": s6["seconds"] + s6["equilibrate_s"],
            "rate_gflops_12cube": rate / 1e9, "est_point_s": est,
            "parallel_value": parallel / est, "parallel_processes": P, "parallel_efficiency": eff / P,
            "parallel_memory_cap": p_mem,
            "sample": f"oracle (restated reference up-looking LDL^T, 1 thread, {model}, {cores} cores visible): "
                      f"the 6^3 configuration factored whole in {s6['seconds']:.2f} s; a 12^3 cube ({p12.n} "
                      f"unknowns, {s12['flops_total']:.3e} flops) whole in {s12['seconds']:.1f} s = "
                      f"{rate / 1e9:.2f} GFLOP/s; this workload's {s32['flops_total']:.3e} flops at that rate plus "
                      f"its equilibration ({s32['equilibrate_s']:.1f} s): {est:.0f} s per frequency point (solve "
                      f"excluded). {P} threads factoring the 8^3 cube at once run at {eff:.1f}x one thread; host "
                      f"memory holds {p_mem} of this workload's factors at once, so {parallel:.1f} points at a time: "
                      f"{parallel / est:.2e} points/s"}


BATCH_SEED = 20261204
# FP32 refinement budget of a sweep point: the GMRES restart length, so an attempt
# runs one Krylov cycle at most
REFINE_BUDGET = 48


def batched_leg(A, ctx, count, steps):
    """The batched workload: `count` seeded problems (cube side U[0.05, 0.2] m, 4-12
    hexes per edge, a conductor block of 1-4 hexes per side spanning z, sigma
    log-uniform in [1e6, 6e7] S/m, frequency log-uniform in [1 Hz, 1 MHz]) in one
    device plan: one refill launch set at every problem's own frequency, one FP32
    factorization (a replayed graph), the port right-hand sides solved and refined
    per problem to backward error 1e-12. Timed against the same problems solved one
    mesh at a time."""
    import torch
    configs, freqs = A.batched_cube_configs(BATCH_SEED, count)
    omegas = 2 * np.pi * freqs
    t0 = time.perf_counter()
    problems = A.Problem.cubes(configs)
    build_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    batch = A.Batch(ctx, problems)
    create_s = time.perf_counter() - t0
    batch.refill(omegas)
    table, t_def, t_tuned = batch.tune_update_tiles(repeats=2)
    b = batch.rhs_device()[None, :].contiguous()
    x = torch.empty_like(b)
    batch.solve_symmetric(b, x, precision=A.APHI_FP32, max_refinements=30, tolerance=1e-12)  # warm-up (capture)
    ctx.synchronize()
    lat = []
    for _ in range(steps):
        s0 = time.perf_counter()
        batch.refill(omegas)
        b = batch.rhs_device()[None, :].contiguous()
        _, reps = batch.solve_symmetric(b, x, precision=A.APHI_FP32, max_refinements=30, tolerance=1e-12)
        ctx.synchronize()
        lat.append(time.perf_counter() - s0)
    t_batch = min(lat)
    st = batch.factorize_ldlt(A.APHI_FP32)
    worst = max(r["backward_error"] for r in reps)
    worst_s = max(r["scaled_backward_error"] for r in reps)
    unknowns = int(batch.n)
    batch.close()
    # the same problems one mesh at a time (meshes created untimed)
    meshes = [A.Mesh(ctx, p) for p in problems]
    ctx.synchronize()
    s0 = time.perf_counter()
    for m, w in zip(meshes, omegas):
        m.refill(w)
        m.solve_symmetric(m.rhs_device()[None, :], precision=A.APHI_FP32, max_refinements=30, tolerance=1e-12)
    ctx.synchronize()
    t_single = time.perf_counter() - s0
    for m in meshes:
        m.close()
    out = {"problems": count, "seed": BATCH_SEED, "unknowns": unknowns, "value": count / t_batch,
           "unit": "problems/s (refill + FP32 factor + port solve refined to backward error <= 1e-12)",
           "ms_per_batch": 1e3 * t_batch, "factor_ms": st["milliseconds"],
           "one_at_a_time_problems_per_s": count / t_single, "speedup_vs_one_at_a_time": t_single / t_batch,
           "backward_error_max": worst, "scaled_backward_error_max": worst_s,
           "host_build_s": build_s, "device_create_s": create_s,
           "update_tile_table": table.tolist(), "tune_ms_default": t_def, "tune_ms_tuned": t_tuned}
    log(f"batched: {count} problems ({unknowns} unknowns) {out['value']:.1f} problems/s "
        f"({1e3 * t_batch:.1f} ms per batch, factor {st['milliseconds']:.1f} ms); one at a time "
        f"{out['one_at_a_time_problems_per_s']:.1f} problems/s; max be {worst:.1e} scaled {worst_s:.1e}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--fp64-steps", type=int, default=4)
    ap.add_argument("--pipeline", type=int, default=2,
                    help="frequency points in flight, each on its own context and copy of the mesh (1: one at a time)")
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--nrhs", type=int, default=16)
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU-baseline sampling (0 skips)")
    ap.add_argument("--serving-requests", type=int, default=200,
                    help="requests of the serving run (mixed sizes 8..32, Poisson arrivals; 0 skips)")
    ap.add_argument("--serving-rate", type=float, default=2.0, help="arrival rate of the serving run (requests/s)")
    ap.add_argument("--serving-slots", type=int, default=2)
    ap.add_argument("--serving-only", action="store_true",
                    help="run only the serving workload and print its summary (with latency by mesh size) as JSON")
    ap.add_argument("--batch-problems", type=int, default=256,
                    help="problems of the batched run (seeded draws of 4..12 hexes per edge; 0 skips)")
    ap.add_argument("--batch-steps", type=int, default=3, help="timed passes over the batch")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of known size (HBM counter calibration) and exit")
    ap.add_argument("--factor-only", choices=("fp32", "fp64"), default=None,
                    help="only refill, equilibrate and factor the workload once at one frequency (counter runs) and exit")
    ap.add_argument("--solve-reps", type=int, default=0,
                    help="with --factor-only: that many 16-RHS solves after the factorization")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the sweep's frequency points are independent and would shard across GPUs "
              "without collectives, but this build runs on a single MI355X. Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import aphi_amd as A

    if not torch.cuda.is_available():
        sys.exit("bench.py: no HIP device")
    if a.copy_calibration:
        # 1 GiB read + 1 GiB written by one copy kernel: the scale for FETCH_SIZE / WRITE_SIZE
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4}))
        return
    if a.serving_only:
        # the serving workload alone (a long stream: progress on stderr every 50 requests)
        from aphi_amd.serving import Server, serving_requests

        def progress(done, total):
            if done % 50 == 0 or done == total:
                log(f"serving: {done}/{total} requests done")
        reqs = serving_requests(RHS_SEED, a.serving_requests, a.serving_rate)
        srv = Server(slots=a.serving_slots, hbm_budget=200e9)
        try:
            results, summary = srv.run(reqs, progress=progress)
        finally:
            srv.close()
        by_n = {}
        for r in reqs:
            res = results[r.id]
            by_n.setdefault(r.config.n, []).append(res["latency_ms"])
        summary.update(rate_per_s=a.serving_rate, seed=RHS_SEED,
                       distinct_meshes=len({r.key() for r in reqs}),
                       latency_by_hexes_per_edge={str(k): {"requests": len(v), "p50_ms": float(np.percentile(v, 50)),
                                                           "p99_ms": float(np.percentile(v, 99))}
                                                  for k, v in sorted(by_n.items())},
                       workload="Poisson stream, meshes n in {8,16,24,32} (p = 0.7 repeat), 1 Hz-100 MHz, 1-8 RHS, "
                                "FP32-MFMA + FP64 refinement to 1e-12")
        print(json.dumps({"serving": summary}), flush=True)
        return
    bar = (6, 12, 20, 23) if a.n == 32 else (a.n // 5, a.n // 5 + 6, (2 * a.n) // 3, (2 * a.n) // 3 + 3)
    cfg = A.CubeConfig(n=a.n, side=0.1, bar=bar, sigma=5.8e7)
    t0 = time.time()
    prob = A.Problem(cfg)
    log(f"host symbolic: {time.time() - t0:.1f}s  n={prob.n} nnz={prob.counts['nnz']} "
        f"nnz(L)={prob.counts['predicted_nnz']} flops/factor={prob.counts['flops']:.3e}")
    ctx = A.Context(0)
    mesh = A.Mesh(ctx, prob)
    freqs = A.frequency_sweep(1.0, 1e8, 64)
    if a.factor_only:
        prec = A.APHI_FP32 if a.factor_only == "fp32" else A.APHI_FP64
        mesh.refill(2 * np.pi * freqs[37])
        mesh.equilibrate(10)
        st = mesh.factorize_ldlt(prec)
        ctx.synchronize()
        solve_ms = []
        if a.solve_reps > 0:
            Bs = A.random_complex_rhs(ctx, RHS_SEED, a.nrhs, prob.n)
            Xs = torch.empty_like(Bs)
            for _ in range(a.solve_reps):
                s0 = time.perf_counter()
                mesh.solve(Bs, Xs)
                ctx.synchronize()
                solve_ms.append((time.perf_counter() - s0) * 1e3)
        print(json.dumps({"factor_only": a.factor_only, "frequency": float(freqs[37]),
                          "milliseconds": st["milliseconds"], "ok": bool(st["ok"]), "solve_ms": solve_ms}))
        return
    n = prob.n
    # right-hand sides of frequency f: the library's counter-based generator
    # (Philox-4x32-10 + Box-Muller, the same bits on host and device), stream
    # RHS_SEED, normals from offset 2 f nrhs n on
    B = torch.stack([A.random_complex_rhs(ctx, RHS_SEED, a.nrhs, n, offset=2 * fi * a.nrhs * n)
                     for fi in range(len(freqs))])
    X = torch.empty(a.nrhs, n, dtype=torch.complex128, device="cuda")
    # --pipeline workers: a context, a copy of the mesh and a solution buffer each.
    # Their solves launch kernel by kernel (graph capture on one thread's stream does
    # not tolerate another thread's work on the legacy stream).
    ctxs, meshes, Xs = [ctx], [mesh], [X]
    for _ in range(1, max(1, a.pipeline)):
        c = A.Context(0)
        ctxs.append(c)
        meshes.append(A.Mesh(c, prob))
        Xs.append(torch.empty_like(X))
    if len(meshes) > 1:
        for m in meshes:
            m.set_solve_graphs(False)
    # the sweep in ascending frequency, as the reference's solve_mesh loop runs it
    # (tools/solve_mesh.cpp:207); --steps below 64 takes every (64/steps)-th point
    nsteps = max(1, min(a.steps, len(freqs)))
    order = [int(round(i * len(freqs) / nsteps)) for i in range(nsteps)]
    # FP64 points spread over the whole band, its top included (their cost does
    # not depend on the frequency)
    k64 = min(a.fp64_steps, nsteps)
    order64 = ([order[int(round(i * (nsteps - 1) / max(k64 - 1, 1)))] for i in range(k64)] if k64 else [])

    def step(fi, precision, w=0):
        m, Xw = meshes[w], Xs[w]
        omega = 2 * np.pi * freqs[fi]
        m.refill(omega)
        _, rep = m.solve_symmetric(B[fi], Xw, precision=precision, equilibration_iterations=10,
                                   max_refinements=(REFINE_BUDGET if precision == A.APHI_FP32 else 0),
                                   tolerance=1e-12)
        # the tets' fields (compute_fields, tools/solve_mesh.cpp:296) reduced to the port
        # current, R and L of every right-hand side (postprocess.hpp:348-358)
        rep["current"], rep["R"], rep["L"] = m.terminal_current(Xw, omega, cfg.side)
        return rep

    def sweep(points, choose, record=None, workers=None):
        # the points in order over --pipeline workers, each with its own context and
        # copy of the mesh: a worker takes the next point as soon as its previous
        # one is done, so one point's refinement solves run beside the next point's
        # factorization. The precision of a point is chosen when it is taken, from
        # the points finished by then; latency = a point's own start to finish.
        lat, reps = [None] * len(points), [None] * len(points)
        lock = threading.Lock()
        nxt = [0]
        errors = []

        def worker(w):
            try:
                while True:
                    with lock:
                        i = nxt[0]
                        if i >= len(points) or errors:
                            return
                        nxt[0] += 1
                        fi = points[i]
                        prec = choose(freqs[fi])
                    s0 = time.perf_counter()
                    rep = step(fi, prec, w)
                    lat[i] = time.perf_counter() - s0
                    rep["requested"] = prec
                    reps[i] = rep
                    if record is not None:
                        with lock:
                            record(freqs[fi], prec, rep)
            except BaseException as e:  # noqa: BLE001 (re-raised below)
                errors.append(e)

        nw = len(meshes) if workers is None else workers
        for c in ctxs:
            c.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        threads = [threading.Thread(target=worker, args=(w,)) for w in range(nw)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        for c in ctxs:
            c.synchronize()
        torch.cuda.synchronize()
        if errors:
            raise errors[0]
        return time.perf_counter() - t0, lat, reps

    # mixed-precision sweep: FP32 factor + GMRES refinement, FP64 factor at the
    # frequencies where the points solved so far show FP32 would be slower
    policy = A.PrecisionPolicy(budget=REFINE_BUDGET)

    # warm-up points: the sweep's first frequencies on every worker, through a
    # policy of their own so the timed sweep starts from nothing learnt
    for w in range(len(meshes)):
        for i in range(a.warmup):
            fi = order[i % len(order)]
            warm = A.PrecisionPolicy(budget=REFINE_BUDGET)
            step(fi, warm.choose(freqs[fi]), w)
    elapsed, lat, reps = sweep(order, policy.choose, policy.record)
    value = len(order) / elapsed
    log(f"fp32+IR: {value:.3f} points/s  p50 {np.percentile(lat, 50) * 1e3:.0f} ms  "
        f"p99 {np.percentile(lat, 99) * 1e3:.0f} ms  max be {max(r['backward_error'] for r in reps):.2e}  "
        f"({len(meshes)} in flight)")

    # plain FP64 path on the same frequency points, in flight the same way
    value64, lat64, reps64 = None, [], []
    if order64:
        el64, lat64, reps64 = sweep(order64, lambda f: A.APHI_FP64)
        value64 = len(order64) / el64
        log(f"fp64: {value64:.3f} points/s  max be {max(r['backward_error'] for r in reps64):.2e}")
    # the measurements below use the first worker's mesh alone
    pipeline = len(meshes)
    for m in meshes[1:]:
        m.close()
    for c in ctxs[1:]:
        c.close()
    del meshes[1:], Xs[1:], ctxs[1:]
    torch.cuda.empty_cache()
    mesh.set_solve_graphs(True)
    sequential = None
    if pipeline > 1:
        # the same sweep one point at a time (a fresh policy): the latency of a point
        # that has the device to itself
        pol1 = A.PrecisionPolicy(budget=REFINE_BUDGET)
        el1, lat1, reps1 = sweep(order, pol1.choose, pol1.record, workers=1)
        sequential = {"value": len(order) / el1, "p50_ms": float(np.percentile(lat1, 50) * 1e3),
                      "p99_ms": float(np.percentile(lat1, 99) * 1e3),
                      "fp32_points": sum(1 for r in reps1 if r["precision_used"] == A.APHI_FP32),
                      "backward_error_max": max(r["backward_error"] for r in reps1)}
        log(f"fp32+IR one at a time: {sequential['value']:.3f} points/s  p50 {sequential['p50_ms']:.0f} ms  "
            f"p99 {sequential['p99_ms']:.0f} ms")

    # error of the FP32+IR solution against the complex128 path, per frequency
    # (untimed; both solved again on the same points)
    err_vs_fp64 = []
    # the FP64 points and the FP32 point that needed the most refinement solves
    fp32_reps = [(r["refinement_steps"], fi) for r, fi in zip(reps, order) if r["precision_used"] == A.APHI_FP32]
    worst32 = max(fp32_reps)[1] if fp32_reps else None
    err_points = order64 + ([worst32] if worst32 is not None and worst32 not in order64 else [])
    for fi in err_points:
        mesh.refill(2 * np.pi * freqs[fi])
        x64, _ = mesh.solve_symmetric(B[fi], precision=A.APHI_FP64, equilibration_iterations=10)
        mesh.refill(2 * np.pi * freqs[fi])
        x32, _ = mesh.solve_symmetric(B[fi], precision=A.APHI_FP32, equilibration_iterations=10,
                                      max_refinements=20, tolerance=1e-12)
        e = torch.linalg.vector_norm(x32 - x64, dim=1) / torch.linalg.vector_norm(x64, dim=1)
        err_vs_fp64.append(float(e.max()))
    if err_vs_fp64:
        log("relative error of FP32+IR vs complex128 per frequency:", " ".join(f"{e:.2e}" for e in err_vs_fp64))

    # roofline of the dominant kernel (MFMA trailing update), measured live with
    # HIP events around each of its launches on the library's stream
    mesh.set_profiling(True)
    mesh.refill(2 * np.pi * freqs[order[0]])
    mesh.equilibrate(10)
    st32 = mesh.factorize_ldlt(A.APHI_FP32)
    st64 = mesh.factorize_ldlt(A.APHI_FP64)
    # the same launches one after the other on one stream: the kernel's rate alone
    mesh.set_profiling(True, serial=True)
    st32s = mesh.factorize_ldlt(A.APHI_FP32)
    mesh.set_profiling(False)
    ctx.synchronize()
    s0 = time.perf_counter()
    mesh.solve(B[order[0]], X)
    ctx.synchronize()
    solve64_ms = (time.perf_counter() - s0) * 1e3
    mesh.factorize_ldlt(A.APHI_FP32)
    ctx.synchronize()
    s0 = time.perf_counter()
    mesh.solve(B[order[0]], X)
    ctx.synchronize()
    solve32_ms = (time.perf_counter() - s0) * 1e3
    log(f"{a.nrhs}-RHS solve: fp64 factor {solve64_ms:.1f} ms, fp32 factor {solve32_ms:.1f} ms")
    ach32 = st32["update_flops"] / (st32["update_ms"] * 1e-3) / 1e12
    ach64 = st64["update_flops"] / (st64["update_ms"] * 1e-3) / 1e12
    ach32s = st32s["update_flops"] / (st32s["update_ms"] * 1e-3) / 1e12
    log(f"update kernel alone (launches serialised): {st32s['update_ms']:.1f} ms = {ach32s:.1f} TFLOP/s")
    log(f"factor fp32 {st32['milliseconds']:.1f} ms ({st32['flops'] / st32['milliseconds'] / 1e9:.1f} TFLOP/s alg.), "
        f"update kernel {st32['update_ms']:.1f} ms over {st32['update_launches']} launches = {ach32:.1f} TFLOP/s; "
        f"fp64 {st64['milliseconds']:.1f} ms, update {ach64:.1f} TFLOP/s")

    cpu = None
    if a.cpu_budget > 0:
        cpu = cpu_baseline(A, prob, mesh, freqs[order[0]])
        log("cpu baseline:", cpu["sample"])

    batched = None
    if a.batch_problems > 0:
        batched = batched_leg(A, ctx, a.batch_problems, a.batch_steps)

    serving = None
    if a.serving_requests > 0:
        # serving mode: a seeded Poisson stream of solve requests (meshes of 8..32
        # hexes per edge repeating with p = 0.7, 1 Hz..100 MHz, 1..8 RHS, FP32 +
        # refinement to 1e-12) with continuous batching over slots and a 200 GB HBM
        # budget; the sweep's mesh is released first
        from aphi_amd.serving import Server, serving_requests
        mesh.close()
        del B, X
        torch.cuda.empty_cache()
        reqs = serving_requests(RHS_SEED, a.serving_requests, a.serving_rate)
        srv = Server(slots=a.serving_slots, hbm_budget=200e9)
        try:
            _, serving = srv.run(reqs)
        finally:
            srv.close()
        serving.update(rate_per_s=a.serving_rate, seed=RHS_SEED,
                       workload="Poisson stream, meshes n in {8,16,24,32} (p = 0.7 repeat), 1 Hz-100 MHz, 1-8 RHS, "
                                "FP32-MFMA + FP64 refinement to 1e-12")
        log(f"serving: {serving['requests']} requests, {serving['solves_per_s']:.1f} solves/s, "
            f"p50 {serving['p50_ms']:.0f} ms, p99 {serving['p99_ms']:.0f} ms, evictions {serving['evictions']}")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": f"frequency points/s (each: refill + factor + {a.nrhs}-RHS solve to backward error <= 1e-12)",
        "n_gpus": 1,
        "steps": len(order),
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / len(order),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (FP32 MFMA) factor + c128 GMRES refinement; c128 factor where FP32 is predicted slower",
        "data": "synthetic: seeded N(0,1) complex right-hand sides (Philox-4x32-10, seed %d); generated Kuhn-cube mesh" % RHS_SEED,
        "seed": RHS_SEED,
        "config": {"workload": f"{a.n}^3-hex Kuhn cube, 6x3 off-centre bar, 64 log frequencies 1 Hz-100 MHz swept in "
                               f"ascending order ({len(order)} timed), {a.nrhs} RHS per point, FP32-MFMA LDL^T + "
                               f"FP64 GMRES-IR to 1e-12",
                   "points_in_flight": pipeline,
                   "unknowns": n, "nnz_L": int(prob.counts["predicted_nnz"]),
                   "factor_flops": prob.counts["flops"], "frequencies_timed": [float(freqs[fi]) for fi in order]},
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "rhs_solves_per_s": value * a.nrhs,
        "backward_error_max": max(r["backward_error"] for r in reps),
        "scaled_backward_error_max": max(r["scaled_backward_error"] for r in reps),
        "refinement_steps": [r["refinement_steps"] for r in reps],
        "precision_requested": [r["requested"] for r in reps],
        "precision_used": [r["precision_used"] for r in reps],
        "fp32_points": sum(1 for r in reps if r["precision_used"] == A.APHI_FP32),
        "sequential": sequential,
        "fp64_value": value64,
        "fp64_p99_ms": float(np.percentile(lat64, 99) * 1e3) if lat64 else None,
        "fp64_backward_error_max": max(r["backward_error"] for r in reps64) if reps64 else None,
        "error_vs_complex128": err_vs_fp64,
        "error_vs_complex128_frequencies": [float(freqs[fi]) for fi in err_points],
        "speedup_vs_fp64": (value / value64) if value64 else None,
        "factor_ms_fp32": st32["milliseconds"],
        "factor_ms_fp64": st64["milliseconds"],
        "solve_ms_fp32": solve32_ms,
        "solve_ms_fp64": solve64_ms,
        "roofline": {"bound": "mfma", "achieved": ach32, "peak": MFMA_FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": ach32 / MFMA_FP32_PEAK_TFLOPS, "traffic": update_traffic(a.n),
                     "kernel": "k_update<float, *> (MFMA trailing update, 128x128 and 64x64 tiles together)",
                     "launches": int(st32["update_launches"]),
                     "avg_launch_ms": st32["update_ms"] / max(1, st32["update_launches"]),
                     "flops_per_factor": st32["update_flops"],
                     "timing": ("HIP events around every launch of one FP32 factorization on the stream it runs "
                                "on, durations summed; the look-ahead overlaps launches of two streams, which "
                                "share the compute units, so this is a lower bound of the rate a launch reaches "
                                "alone (achieved_alone: the same launches serialised on one stream)"),
                     "achieved_alone": ach32s, "frac_alone": ach32s / MFMA_FP32_PEAK_TFLOPS},
        "roofline_fp64": {"achieved": ach64, "peak": MFMA_FP64_PEAK_TFLOPS, "frac": ach64 / MFMA_FP64_PEAK_TFLOPS},
        "cpu_baseline": cpu,
        "batched": batched,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)
    mesh.close()
    ctx.close()


if __name__ == "__main__":
    main()
