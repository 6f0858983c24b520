"""Continuous-batching serving loop for measurement solves on one GPU.

A request is a seeded configuration (L, T, Ls, beta, am) with 1-12 complex
Gaussian sources; serving it means building the links on the device
(u = 1 + i theta, gauge.F90:12-43), solving (M^+M) x_k = b_k for every source
to FP64 true residual `tol` with the batched mixed-precision solver (the
congrad system of measure_module.F90:9-140, all sources in one multi-RHS
launch), and returning per-source solution norms and true residuals.

Scheduling:
  * `slots` worker threads, each with its own library context (scratch
    buffers + HIP stream), so requests that are admitted together overlap on
    the device;
  * admission by HBM footprint: a request is started only when its estimated
    device bytes fit under `hbm_budget` minus what running requests hold;
    waiting requests are admitted as soon as a slot and memory free up
    (first-fit over the queue, so a large request does not block small ones
    forever: it gets priority once it has waited `max_skip` admissions);
  * fields live on the device for the whole request: theta and the sources
    are drawn on the device by the library's counter-based generator, the
    solution norms are reduced there, and only the result summary returns to
    the host;
  * after each request the worker releases its context's scratch buffers, so
    device memory in use is the sum of the running requests' footprints
    (which bound what the solver allocates) and never a worker's high-water
    mark.

Every request computes exactly the same arithmetic alone or in a mix (its
own context, deterministic reductions), so results are independent of load.
`run_load(..., shed=True)` drops the requests still queued when the arrival
window closes, so a rate sweep stays bounded in time and reports how many
requests each rate could not serve.
"""
import collections
import dataclasses
import queue
import threading
import time

import numpy as np

import hiprt as hr
import thirring_hip as G


@dataclasses.dataclass
class Request:
    rid: int
    L: int
    T: int
    Ls: int
    beta: float
    am: float
    nsrc: int
    seed: int
    tol: float = 1e-10
    arrival: float = 0.0
    start: float = 0.0
    finish: float = 0.0
    result: dict = None

    @property
    def dims(self):
        return (self.L, self.L, self.T, self.Ls)

    def footprint(self):
        """Device bytes the request holds while it runs: theta, links (FP64
        and FP32), sources and solutions, and the batched solver's scratch
        with FP32 Krylov storage below 1e-2 (per source, in FP64 spinors: a
        ring of 32 residuals (a one-shift solve keeps 32-iteration blocks)
        plus h and r0 in FP64 (34) and the same in FP32
        (17), the shift's p (1), the six batches of the FP64 defect
        correction (6) and the three of the reliable updates (3)), plus 1 MiB
        for partial sums and solver state."""
        v4 = self.L * self.L * self.T
        spinor = 4 * v4 * self.Ls * 16
        per_src = spinor * (34 + 1 + 6 + 3) + spinor // 2 * 34 + 2 * spinor  # + source, solution
        return 3 * v4 * (4 + 16 + 8) + self.nsrc * per_src + (1 << 20)


def random_request(rid, rng, arrival=0.0, tol=1e-10):
    L = int(rng.choice([8, 12, 16, 24]))
    T = int(rng.choice([L, 2 * L]))
    Ls = int(rng.choice([8, 16, 32, 48]))
    beta = float(rng.uniform(0.2, 0.5))
    am = float(rng.choice([0.005, 0.01, 0.02, 0.05]))
    nsrc = int(rng.integers(1, 13))
    return Request(rid, L, T, Ls, beta, am, nsrc, int(rng.integers(0, 2**31 - 1)), tol, arrival)


def solve_request(req, stream=None, maxcg=200000, keep_x=False):
    """Runs one request in the calling thread's bound context: theta ~
    N(0, 1/beta) and the sources (unit complex variance) drawn on the device
    from the request's seed, links u = 1 + i theta, the batched solve of
    (M^+M) x_k = b_k to FP64 true residual `tol`, and the solution norms
    reduced on the device.  keep_x also returns the solutions (a copy to the
    host)."""
    lx, ly, lt, ls = req.dims
    n4 = lx * ly * lt * 3
    nsp = 4 * lx * ly * lt * ls
    arrays = []

    def dev(shape, dtype):
        a = hr.DeviceArray(shape, dtype)
        arrays.append(a)
        return a

    try:
        theta = dev((n4,), np.float32)
        G.gaussian_dev(theta, n4, 1.0 / np.sqrt(req.beta), 2 * req.seed, prec="f32",
                       stream=stream)
        l64 = dev((n4, 2), np.float64)
        l32 = dev((n4, 2), np.float32)
        G.coef_dev(l64, theta, req.dims, stream=stream)
        G.links_to_f32_dev(l32, l64, req.dims, stream=stream)
        bd = dev((req.nsrc, nsp, 2), np.float64)
        G.gaussian_dev(bd, req.nsrc * nsp * 2, np.sqrt(0.5), 2 * req.seed + 1, stream=stream)
        x = dev((req.nsrc, 1, nsp, 2), np.float64)
        it, cgr, rel = G.multishift_fast_dev(l64, l32, req.dims, req.am, 3, np.zeros(1), bd, x,
                                             tol=req.tol, maxcg=maxcg, nsrc=req.nsrc, th32=1e-2,
                                             true_residuals=True, stream=stream)
        out = {"iterations": int(it), "true_residual": rel[:, 0].tolist(),
               "x_norm": G.norms_dev(x, req.nsrc, nsp, stream=stream).tolist()}
        if keep_x:
            out["x"] = x.to_host()
        return out
    finally:
        for a in arrays:
            a.free()


class ServingEngine:
    def __init__(self, slots=4, hbm_budget=200e9, max_skip=8, keep_solutions=False):
        self.slots = slots
        self.hbm_budget = hbm_budget
        self.max_skip = max_skip
        self.keep = keep_solutions
        self.pending = collections.deque()
        self.lock = threading.Condition()
        self.in_use = 0
        self.running = 0
        self.done = []
        self.shed = []
        self.stop = False
        self.skips = collections.Counter()
        self.work = queue.Queue()
        self.workers = [threading.Thread(target=self._worker, daemon=True) for _ in range(slots)]
        self.scheduler = threading.Thread(target=self._schedule, daemon=True)
        for w in self.workers:
            w.start()
        self.scheduler.start()

    def submit(self, req):
        """Queues a request; one that could never fit the HBM budget is
        finished at once with an error instead of blocking the queue."""
        with self.lock:
            if req.arrival == 0.0:
                req.arrival = time.perf_counter()
            if req.footprint() > self.hbm_budget:
                req.start = req.finish = time.perf_counter()
                req.result = {"error": f"footprint {req.footprint():.3e} B exceeds the HBM budget "
                                       f"{self.hbm_budget:.3e} B"}
                self.done.append(req)
            else:
                self.pending.append(req)
            self.lock.notify_all()

    def _probe(self):
        """Index of the request to start next, or None (no side effects):
        the oldest request that fits the free budget, except that a request
        passed over max_skip times holds back every younger one."""
        free = self.hbm_budget - self.in_use
        for idx, r in enumerate(self.pending):
            if r.footprint() <= free:
                return idx
            if self.skips[r.rid] >= self.max_skip:
                return None
        return None

    def _commit(self, idx):
        """Starts pending[idx]: every older request it overtakes counts one skip."""
        for older in list(self.pending)[:idx]:
            self.skips[older.rid] += 1
        req = self.pending[idx]
        del self.pending[idx]
        self.in_use += req.footprint()
        self.running += 1
        return req

    def _schedule(self):
        while True:
            with self.lock:
                idx = None
                while True:
                    if self.running < self.slots and self.pending:
                        idx = self._probe()
                        if idx is not None:
                            break
                    if self.stop and not self.pending:
                        for _ in self.workers:
                            self.work.put(None)
                        return
                    self.lock.wait(0.05)
                req = self._commit(idx)
            self.work.put(req)

    def shed_pending(self):
        """Drops every request not yet started (returns how many): the load
        this engine could not take within its arrival window."""
        with self.lock:
            n = len(self.pending)
            now = time.perf_counter()
            while self.pending:
                r = self.pending.popleft()
                r.start = r.finish = now
                r.result = {"shed": True}
                self.shed.append(r)
            self.lock.notify_all()
            return n

    def _worker(self):
        ctx = G.Context()
        ctx.bind()
        stream = hr.Stream()
        while True:
            req = self.work.get()
            if req is None:
                break
            req.start = time.perf_counter()
            try:
                req.result = solve_request(req, stream=stream, keep_x=self.keep)
                req.result["scratch_bytes"] = ctx.scratch_bytes()
            except Exception as e:  # a failed solve still frees its slot and memory
                req.result = {"error": f"{type(e).__name__}: {e}"}
            finally:
                # give the request's scratch back before its footprint is released
                try:
                    ctx.trim()
                    req.result["scratch_after_trim"] = ctx.scratch_bytes()
                except Exception as e:
                    req.result = {"error": f"{type(e).__name__}: {e}"}
            req.finish = time.perf_counter()
            with self.lock:
                self.in_use -= req.footprint()
                self.running -= 1
                self.done.append(req)
                self.lock.notify_all()
        if hasattr(stream, "destroy"):
            stream.destroy()
        G.Context.unbind()
        ctx.destroy()

    def drain(self, timeout=None):
        t0 = time.perf_counter()
        with self.lock:
            self.stop = True
            self.lock.notify_all()
        for w in self.workers:
            w.join(None if timeout is None else max(0.0, timeout - (time.perf_counter() - t0)))
        self.scheduler.join(1.0)
        return self.done


def run_load(rate, duration, seed=0, slots=4, hbm_budget=200e9, tol=1e-10, max_requests=None,
             request_fn=random_request, shed=False):
    """Poisson arrivals at `rate`/s for `duration` s; returns summary metrics.
    shed: drop the requests still queued when the window closes (the engine
    then only finishes those already running)."""
    rng = np.random.default_rng(seed)
    eng = ServingEngine(slots=slots, hbm_budget=hbm_budget)
    t0 = time.perf_counter()
    n = 0
    t_next = t0 + rng.exponential(1.0 / rate)
    while True:
        now = time.perf_counter()
        if now - t0 >= duration or (max_requests and n >= max_requests):
            break
        if now >= t_next:
            eng.submit(request_fn(n, rng, arrival=t_next, tol=tol))
            n += 1
            t_next += rng.exponential(1.0 / rate)
        else:
            time.sleep(min(0.002, t_next - now))
    nshed = eng.shed_pending() if shed else 0
    done = eng.drain()
    wall = time.perf_counter() - t0
    ok = [r for r in done if "error" not in r.result]
    lat = np.array([r.finish - r.arrival for r in ok]) if ok else np.zeros(1)
    solves = sum(r.nsrc for r in ok)
    worst = max((max(r.result["true_residual"]) for r in ok), default=0.0)
    return {"rate_per_s": rate, "submitted": n, "requests": len(ok),
            "failed": len(done) - len(ok), "shed": nshed,
            "sources_solved": int(solves),
            "solves_per_s": round(solves / wall, 3), "requests_per_s": round(len(ok) / wall, 3),
            "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
            "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
            "max_true_residual": worst, "wall_s": round(wall, 2)}


def rate_sweep(rates, window, seed=0, slots=4, hbm_budget=200e9, tol=1e-10,
               request_fn=random_request, progress=None):
    """run_load with shedding at each rate; the highest sustainable rate is the
    largest one at which nothing was shed or failed."""
    out = []
    for k, rate in enumerate(rates):
        if progress:
            progress(f"serving: Poisson rate {rate}/s for {window} s, {slots} slots")
        out.append(run_load(rate, window, seed=seed + k, slots=slots, hbm_budget=hbm_budget,
                            tol=tol, request_fn=request_fn, shed=True))
    ok = [m["rate_per_s"] for m in out if m["shed"] == 0 and m["failed"] == 0]
    return {"sweep": out, "highest_sustainable_rate_per_s": max(ok) if ok else None}
