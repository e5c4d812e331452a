"""Serving loop for the A-Phi solve: a stream of solve requests of mixed mesh
sizes and frequencies, run with continuous batching on one GPU.

Worker slots are host threads; a request is taken by the first free slot once it
has arrived, and the C-ABI releases the interpreter lock, so the slots' refills,
factorizations and solves overlap on the device. Symbolic problems (host) are
built once per distinct mesh, on a host thread pool as soon as a request for a
new mesh arrives, and reused. Device meshes are kept in one cache shared by all
slots, each mesh with a context (its own HIP streams) of its own, so a mesh is
uploaded and held once whichever slot serves it; a mesh serves one request at a
time (its numeric buffers are per mesh), and a free slot takes the earliest
arrived request whose mesh is not busy. New meshes are admitted against an HBM
budget, least recently used idle meshes evicted first. Each distinct mesh keeps
a precision policy (aphi_amd.PrecisionPolicy) over the frequencies it has
served, so a request at a frequency where the FP32 factor would need more
refinement than an FP64 factor costs is factored in FP64 from the start. Every
computation of a request depends only on its own inputs (fixed-order
reductions, no float atomics), so its results are the same alone and in the mix
for the same precision.
"""
import hashlib
import math
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np

from . import (APHI_FP32, APHI_FP64, APHI_ROW_SCALED, Context, CubeConfig, Mesh, Problem, random_complex_rhs,
               random_uniform_host)
from .sweep import PrecisionPolicy


@dataclass
class Request:
    id: int
    arrival: float        # seconds after the stream starts
    config: CubeConfig
    frequency: float
    nrhs: int
    seed: int

    def key(self):
        c = self.config
        return (c.n, c.side, tuple(c.bar), c.sigma)


def serving_requests(seed, count, rate, sizes=(8, 16, 24, 32), repeat=0.7, max_rhs=8, fmin=1.0, fmax=1e8):
    """A seeded Poisson stream: request i uses uniforms 12i .. 12i + 11 of stream
    ``seed``: an exponential gap of mean 1/rate, with probability ``repeat`` a mesh
    drawn earlier (uniformly among them), else a new one (hexes per edge from
    ``sizes``, a conductor block of 1-4 hexes per side (at most n - 1) spanning z,
    sigma log-uniform in [1e6, 6e7] S/m, side 0.1 m), a frequency log-uniform in
    [fmin, fmax] and 1..max_rhs right-hand sides."""
    u = random_uniform_host(seed, 12 * int(count))
    t, meshes, out = 0.0, [], []
    for i in range(int(count)):
        v = u[12 * i: 12 * i + 12]
        t += -math.log1p(-v[0]) / rate
        if meshes and v[1] < repeat:
            cfg = meshes[min(int(v[2] * len(meshes)), len(meshes) - 1)]
        else:
            n = sizes[min(int(v[3] * len(sizes)), len(sizes) - 1)]
            bx, by = min(1 + min(int(v[4] * 4), 3), n - 1), min(1 + min(int(v[5] * 4), 3), n - 1)
            i0, j0 = min(int(v[6] * (n - bx + 1)), n - bx), min(int(v[7] * (n - by + 1)), n - by)
            sigma = math.exp(math.log(1e6) + (math.log(6e7) - math.log(1e6)) * v[8])
            cfg = CubeConfig(n=n, side=0.1, bar=(i0, i0 + bx, j0, j0 + by), sigma=sigma)
            meshes.append(cfg)
        f = math.exp(math.log(fmin) + (math.log(fmax) - math.log(fmin)) * v[9])
        nrhs = 1 + min(int(v[10] * max_rhs), max_rhs - 1)
        out.append(Request(i, t, cfg, f, nrhs, (int(seed) * 1000003 + i) & 0xFFFFFFFFFFFF))
    return out


def _workspace_bytes(p, precision):
    """Device bytes of one request's numeric state on the mesh of problem p:
    factor and contribution blocks in the factor's precision, assembled values,
    element matrices, solve and refinement vectors."""
    c = p.counts
    cs = 16 if precision == APHI_FP64 else 8
    return int(cs * (c["factor_entries"] + c["cb_pool_entries"]) + 16 * 196 * c["tets"] + 256 * c["nnz"]
               + 4096 * c["num_total"])


def _plan_bytes(p):
    """An upper estimate of a mesh's plan on the device (assembly plan, pattern,
    supernodal plan, per-tet unknowns), replaced by the measured size once it is
    uploaded."""
    c = p.counts
    return int(16 * c["contributions"] + 64 * c["nnz"] + 256 * c["num_total"] + 256 * c["tets"] + (64 << 20))


class Server:
    """Runs request streams with ``slots`` concurrent workers on one device.

    precision: APHI_FP32 (FP32 factor + refinement, FP64 where refinement fails),
    APHI_FP64, or "auto" (per mesh, the precision policy's choice by frequency)."""

    def __init__(self, slots=2, hbm_budget=200e9, precision="auto", tolerance=1e-12, max_refinements=30,
                 host_threads=8, device=0, keep_solutions=False):
        self.slots = int(slots)
        self.budget = float(hbm_budget)
        self.precision, self.tolerance, self.max_refinements = precision, tolerance, max_refinements
        self.keep_solutions = keep_solutions
        self.device = device
        import torch  # the device allocator: initialised here, not first in a worker thread
        torch.empty(1, device=f"cuda:{device}")
        self.pool = ThreadPoolExecutor(max_workers=int(host_threads))
        self.problems = {}           # key -> Future[Problem]
        self.meshes = {}             # key -> dict(mesh, ctx, bytes, last, busy)
        self.policies = {}           # key -> PrecisionPolicy
        self.lock = threading.Condition()
        self.reserved = 0
        self.evictions = 0
        self.trims = 0
        self.host_builds = 0
        self.uploads = 0

    def close(self):
        with self.lock:
            for e in self.meshes.values():
                self._drop(e)
            self.meshes.clear()
        self.pool.shutdown(wait=True)

    @staticmethod
    def _drop(e):
        if e.get("mesh") is not None:
            e["mesh"].close()
        if e.get("ctx") is not None:
            e["ctx"].close()
        e["mesh"] = e["ctx"] = None

    # ---------------------------------------------------------------- memory
    def _admit(self, e, need):
        """Reserves ``need`` more bytes for entry ``e`` (busy, in the cache): its
        request's numeric state, and its plan when the mesh is not resident. Idle
        meshes give memory back least recently used first: their per-frequency
        state (aphi_mesh_release_workspace) before their plans, so the symbolic
        factors stay on the device as long as the budget allows; waits while busy
        meshes hold the memory. Every reserved byte belongs to an entry."""
        with self.lock:
            while self.reserved + need > self.budget:
                idle = sorted((x["last"], k) for k, x in self.meshes.items() if not x["busy"])
                trim = [k for _, k in idle
                        if self.meshes[k]["mesh"] is not None and self.meshes[k]["bytes"] > self.meshes[k]["plan"]]
                if trim:
                    x = self.meshes[trim[0]]
                    x["mesh"].release_workspace()
                    self.reserved -= x["bytes"] - x["plan"]
                    x["bytes"] = x["plan"]
                    self.trims += 1
                    continue
                if not idle:
                    if not any(x["busy"] and x is not e for x in self.meshes.values()):
                        raise MemoryError(f"a mesh needs {need / 1e9:.1f} GB, over the HBM budget")
                    self.lock.wait()
                    continue
                k = idle[0][1]
                x = self.meshes.pop(k)
                self._drop(x)
                self.reserved -= x["bytes"]
                self.evictions += 1
            self.reserved += need
            e["bytes"] += need

    def _problem(self, req):
        key = req.key()
        with self.lock:
            fut = self.problems.get(key)
            if fut is None:
                fut = self.pool.submit(Problem, req.config)
                self.problems[key] = fut
                self.host_builds += 1
        return fut

    def _build(self, req, e, p):
        """Uploads the request's mesh into the cache entry ``e`` (a busy
        placeholder put there when the request was taken, its memory admitted)."""
        try:
            e["ctx"] = Context(self.device)
            e["mesh"] = Mesh(e["ctx"], p)
            e["plan"] = e["mesh"].device_bytes()
            # other slots' threads issue GPU work at any time: no graph capture
            e["mesh"].set_solve_graphs(False)
            with self.lock:
                self.uploads += 1
        except Exception:
            with self.lock:
                if self.meshes.get(req.key()) is e:
                    self.meshes.pop(req.key())
                self.reserved -= e["bytes"]
                e["bytes"] = 0
                self._drop(e)
                self.lock.notify_all()
            raise

    def _release(self, key, e):
        with self.lock:
            if self.meshes.get(key) is e:
                actual = e["mesh"].device_bytes() if e.get("mesh") is not None else 0
                self.reserved += actual - e["bytes"]
                e["bytes"] = actual
                e["busy"] = False
                e["last"] = time.perf_counter()
            self.lock.notify_all()

    def _choose(self, req):
        if self.precision != "auto":
            return self.precision
        with self.lock:
            pol = self.policies.setdefault(req.key(), PrecisionPolicy(budget=self.max_refinements))
            return pol.choose(req.frequency)

    def _record(self, req, prec, rep):
        if self.precision != "auto":
            return
        with self.lock:
            self.policies[req.key()].record(req.frequency, prec, rep)

    # ---------------------------------------------------------------- requests
    def _serve(self, req, e, prec):
        m = e["mesh"]
        omega = 2.0 * math.pi * req.frequency
        m.refill(omega, APHI_ROW_SCALED)
        B = random_complex_rhs(e["ctx"], req.seed, req.nrhs, m.n)
        X, rep = m.solve_symmetric(B, precision=prec, equilibration_iterations=10,
                                   max_refinements=self.max_refinements, tolerance=self.tolerance)
        self._record(req, prec, rep)
        xh = X.cpu().numpy()
        be, sbe = rep["backward_error"], rep["scaled_backward_error"]
        out = dict(id=req.id, n=m.n, nrhs=req.nrhs, backward_error=be, scaled_backward_error=sbe,
                   ok=bool(be <= self.tolerance and sbe <= self.tolerance),
                   refinement_steps=rep["refinement_steps"], precision_requested=prec,
                   precision_used=rep["precision_used"], factorize_ms=rep["factorize_ms"],
                   digest=hashlib.sha256(xh.tobytes()).hexdigest())
        if self.keep_solutions:
            out["x"] = xh
        return out

    def run(self, requests, progress=None):
        """Serves ``requests`` in real time (each becomes available at its arrival
        time after the call starts). Returns (results by id, summary). ``progress``,
        if given, is called as progress(done, total) after each request."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        queue, results = [], {}
        state = dict(done=False)
        t0 = time.perf_counter()

        def take():
            """Under the lock: the earliest queued request whose mesh is not busy,
            with its cache entry marked busy (a placeholder if it is not resident)."""
            for r in queue:
                e = self.meshes.get(r.key())
                if e is None or not e["busy"]:
                    queue.remove(r)
                    if e is None:
                        e = dict(mesh=None, ctx=None, bytes=0, plan=0, last=time.perf_counter(), busy=True)
                        self.meshes[r.key()] = e
                    e["busy"] = True
                    return r, e
            return None, None

        def worker(slot):
            while True:
                with self.lock:
                    while True:
                        req, e = take()
                        if req is not None:
                            break
                        if state["done"] and not queue:
                            return
                        self.lock.wait()
                start = time.perf_counter()
                build_ms = 0.0
                try:
                    p = self._problem(req).result()
                    prec = self._choose(req)
                    if e["mesh"] is None:
                        need = _plan_bytes(p) + _workspace_bytes(p, prec)
                    else:  # what its numeric state may still grow by
                        need = max(0, e["plan"] + _workspace_bytes(p, prec) - e["bytes"])
                    self._admit(e, need)
                    if e["mesh"] is None:
                        self._build(req, e, p)
                    build_ms = (time.perf_counter() - start) * 1e3
                    res = self._serve(req, e, prec)
                except Exception as ex:  # reported per request; the loop goes on
                    res = dict(id=req.id, n=0, nrhs=req.nrhs, ok=False, error=f"{type(ex).__name__}: {ex}",
                               backward_error=float("inf"), scaled_backward_error=float("inf"), refinement_steps=0,
                               precision_requested=-1, precision_used=-1, factorize_ms=0.0, digest="")
                finally:
                    self._release(req.key(), e)
                end = time.perf_counter()
                res.update(slot=slot, arrival=req.arrival, start=start - t0, end=end - t0,
                           latency_ms=(end - t0 - req.arrival) * 1e3, build_ms=build_ms,
                           serve_ms=(end - start) * 1e3 - build_ms)
                with self.lock:
                    results[req.id] = res
                    done = len(results)
                if progress is not None:
                    progress(done, len(reqs))

        threads = [threading.Thread(target=worker, args=(s,)) for s in range(self.slots)]
        for th in threads:
            th.start()
        try:
            for req in reqs:
                delay = t0 + req.arrival - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                self._problem(req)  # start its host build at arrival
                with self.lock:
                    queue.append(req)
                    self.lock.notify_all()
        finally:
            with self.lock:
                state["done"] = True
                self.lock.notify_all()
            for th in threads:
                th.join()
        makespan = max(r["end"] for r in results.values()) if results else 0.0
        lat = np.array([results[r.id]["latency_ms"] for r in reqs])
        fp64 = sum(1 for r in reqs if results[r.id]["precision_used"] == APHI_FP64)
        stalled = sum(1 for r in reqs if results[r.id]["precision_requested"] == APHI_FP32
                      and results[r.id]["precision_used"] == APHI_FP64)
        summary = dict(requests=len(reqs), slots=self.slots, makespan_s=makespan,
                       p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                       requests_per_s=len(reqs) / makespan if makespan else 0.0,
                       solves_per_s=sum(r.nrhs for r in reqs) / makespan if makespan else 0.0,
                       max_backward_error=max(results[r.id]["backward_error"] for r in reqs),
                       max_scaled_backward_error=max(results[r.id]["scaled_backward_error"] for r in reqs),
                       all_ok=all(results[r.id]["ok"] for r in reqs), evictions=self.evictions,
                       workspace_releases=self.trims,
                       host_builds=self.host_builds, device_uploads=self.uploads, fp64_requests=fp64,
                       fp32_attempts_refactored=stalled, hbm_budget_gb=self.budget / 1e9,
                       mean_build_ms=float(np.mean([results[r.id].get("build_ms", 0.0) for r in reqs])),
                       mean_serve_ms=float(np.mean([results[r.id].get("serve_ms", 0.0) for r in reqs])),
                       mean_wait_ms=float(np.mean([results[r.id]["start"] - r.arrival for r in reqs]) * 1e3))
        return results, summary
