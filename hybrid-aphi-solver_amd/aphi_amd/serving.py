"""Serving loop for the A-Phi solve: a stream of solve requests of mixed mesh
sizes and frequencies, run with continuous batching on one GPU.

Each worker slot owns a context (its own HIP streams) and is driven by its own
host thread; the C-ABI releases the interpreter lock, so the slots' refills,
factorizations and solves overlap on the device. A request is taken by the
first free slot once it has arrived. Symbolic problems (host) are built once
per distinct mesh, on a host thread pool as soon as a request for a new mesh
arrives, and reused; device meshes are cached per slot and admitted against an
HBM budget, least recently used idle meshes evicted first. Every computation
of a request depends only on its own inputs (fixed-order reductions, no float
atomics), so its results are the same alone and in the mix.
"""
import hashlib
import math
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np

from . import (APHI_FP32, APHI_ROW_SCALED, Context, CubeConfig, Mesh, Problem, random_complex_rhs,
               random_uniform_host)


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


def _estimate_bytes(p):
    """Device bytes of a mesh with an FP32 factor (a fallback to FP64 is accounted
    from the mesh's actual size once its request is done)."""
    c = p.counts
    return int(8 * (c["factor_entries"] + c["cb_pool_entries"]) + 16 * 196 * c["tets"] + 256 * c["nnz"]
               + 4096 * c["num_total"])


class Server:
    """Runs request streams with ``slots`` concurrent workers on one device."""

    def __init__(self, slots=2, hbm_budget=200e9, precision=APHI_FP32, tolerance=1e-12, max_refinements=40,
                 host_threads=8, device=0, keep_solutions=False):
        self.slots = int(slots)
        self.budget = float(hbm_budget)
        self.precision, self.tolerance, self.max_refinements = precision, tolerance, max_refinements
        self.keep_solutions = keep_solutions
        import torch  # the device allocator: initialised here, not first in a worker thread
        torch.empty(1, device=f"cuda:{device}")
        self.ctxs = [Context(device) for _ in range(self.slots)]
        self.pool = ThreadPoolExecutor(max_workers=int(host_threads))
        self.problems = {}           # key -> Future[Problem]
        self.meshes = {}             # (slot, key) -> dict(mesh, bytes, last, busy)
        self.lock = threading.Condition()
        self.reserved = 0
        self.evictions = 0
        self.host_builds = 0

    def close(self):
        with self.lock:
            for e in self.meshes.values():
                if e["mesh"] is not None:
                    e["mesh"].close()
            self.meshes.clear()
        self.pool.shutdown(wait=True)
        for c in self.ctxs:
            c.close()

    # ---------------------------------------------------------------- memory
    def _admit(self, slot, key, need):
        """Reserves ``need`` bytes for a new mesh of ``slot``, evicting idle meshes
        (least recently used first); waits while other slots hold the memory. The
        reservation is entered (busy, no mesh yet) under the lock, so every
        reserved byte belongs to an entry."""
        with self.lock:
            while self.reserved + need > self.budget:
                idle = sorted((e["last"], k) for k, e in self.meshes.items() if not e["busy"])
                if not idle:
                    if not any(e["busy"] for e in self.meshes.values()):
                        raise MemoryError(f"a mesh needs {need / 1e9:.1f} GB, over the HBM budget")
                    self.lock.wait()
                    continue
                k = idle[0][1]
                e = self.meshes.pop(k)
                e["mesh"].close()
                self.reserved -= e["bytes"]
                self.evictions += 1
            self.reserved += need
            e = dict(mesh=None, bytes=need, last=time.perf_counter(), busy=True)
            self.meshes[key] = e
            return e

    def _problem(self, req):
        key = req.key()
        with self.lock:
            fut = self.problems.get(key)
            if fut is None:
                fut = self.pool.submit(Problem, req.config)
                self.problems[key] = fut
                self.host_builds += 1
        return fut

    def _mesh(self, slot, req):
        key = (slot, req.key())
        with self.lock:
            e = self.meshes.get(key)
            if e is not None:
                e["busy"] = True
                return e
        p = self._problem(req).result()
        e = self._admit(slot, key, _estimate_bytes(p))
        try:
            e["mesh"] = Mesh(self.ctxs[slot], p)
            # other slots' threads issue GPU work at any time: no graph capture
            e["mesh"].set_solve_graphs(False)
        except Exception:
            with self.lock:
                self.meshes.pop(key, None)
                self.reserved -= e["bytes"]
                self.lock.notify_all()
            raise
        return e

    def _release(self, e):
        with self.lock:
            actual = e["mesh"].device_bytes()
            self.reserved += actual - e["bytes"]
            e["bytes"] = actual
            e["busy"] = False
            e["last"] = time.perf_counter()
            self.lock.notify_all()

    # ---------------------------------------------------------------- requests
    def _serve(self, slot, req):
        e = self._mesh(slot, req)
        try:
            m = e["mesh"]
            omega = 2.0 * math.pi * req.frequency
            m.refill(omega, APHI_ROW_SCALED)
            B = random_complex_rhs(self.ctxs[slot], req.seed, req.nrhs, m.n)
            X, rep = m.solve_symmetric(B, precision=self.precision, equilibration_iterations=10,
                                       max_refinements=self.max_refinements, tolerance=self.tolerance)
            xh = X.cpu().numpy()
            out = dict(id=req.id, n=m.n, nrhs=req.nrhs, ok=bool(rep["ok"]), backward_error=rep["backward_error"],
                       refinement_steps=rep["refinement_steps"], precision_used=rep["precision_used"],
                       digest=hashlib.sha256(xh.tobytes()).hexdigest())
            if self.keep_solutions:
                out["x"] = xh
            return out
        finally:
            self._release(e)

    def run(self, requests):
        """Serves ``requests`` in real time (each becomes available at its arrival
        time after the call starts). Returns (results by id, summary)."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        queue, results = [], {}
        done = threading.Event()
        qlock = threading.Condition()
        t0 = time.perf_counter()

        def pick(slot):
            """The earliest request whose mesh this slot holds; else the earliest one
            no other slot holds; else the earliest."""
            with self.lock:
                mine = {k for (s_, k) in self.meshes if s_ == slot}
                others = {k for (s_, k) in self.meshes if s_ != slot}
            for r in queue:
                if r.key() in mine:
                    return r
            for r in queue:
                if r.key() not in others:
                    return r
            return queue[0]

        def worker(slot):
            while True:
                with qlock:
                    while not queue and not done.is_set():
                        qlock.wait()
                    if not queue:
                        return
                    req = pick(slot)
                    queue.remove(req)
                start = time.perf_counter()
                try:
                    res = self._serve(slot, req)
                except Exception as ex:  # reported per request; the loop goes on
                    res = dict(id=req.id, n=0, nrhs=req.nrhs, ok=False, error=f"{type(ex).__name__}: {ex}",
                               backward_error=float("inf"), refinement_steps=0, precision_used=-1, digest="")
                end = time.perf_counter()
                res.update(slot=slot, arrival=req.arrival, start=start - t0, end=end - t0,
                           latency_ms=(end - t0 - req.arrival) * 1e3)
                results[req.id] = res

        threads = [threading.Thread(target=worker, args=(s,)) for s in range(self.slots)]
        for th in threads:
            th.start()
        try:
            for req in reqs:
                delay = t0 + req.arrival - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                self._problem(req)  # start its host build at arrival
                with qlock:
                    queue.append(req)
                    qlock.notify()
        finally:
            with qlock:
                done.set()
                qlock.notify_all()
            for th in threads:
                th.join()
        makespan = max(r["end"] for r in results.values()) if results else 0.0
        lat = np.array([results[r.id]["latency_ms"] for r in reqs])
        summary = dict(requests=len(reqs), slots=self.slots, makespan_s=makespan,
                       p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                       requests_per_s=len(reqs) / makespan if makespan else 0.0,
                       solves_per_s=sum(r.nrhs for r in reqs) / makespan if makespan else 0.0,
                       max_backward_error=max(results[r.id]["backward_error"] for r in reqs),
                       all_ok=all(results[r.id]["ok"] for r in reqs), evictions=self.evictions,
                       host_builds=self.host_builds, hbm_budget_gb=self.budget / 1e9)
        return results, summary
