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
  * fields live on the device for the whole request; only the result summary
    returns to the host.

Every request computes exactly the same arithmetic alone or in a mix (its
own context, deterministic reductions), so results are independent of load.
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
        """Device bytes: links, sources, solutions and the solver scratch."""
        v5 = self.L * self.L * self.T * self.Ls
        spinor = 4 * v5 * 16
        # ring of 32 FP64 residual batches + FP32 copies + p, h, r0 + shift x/p
        per_src = spinor * (32 + 3 + 2) + spinor // 2 * (32 + 3) + 2 * spinor
        return 3 * self.L * self.L * self.T * 24 + self.nsrc * (per_src + 2 * spinor)


def random_request(rid, rng, arrival=0.0, tol=1e-10):
    L = int(rng.choice([8, 12, 16, 24]))
    T = int(rng.choice([L, 2 * L]))
    Ls = int(rng.choice([8, 16, 32, 48]))
    beta = float(rng.uniform(0.2, 0.5))
    am = float(rng.choice([0.005, 0.01, 0.02, 0.05]))
    nsrc = int(rng.integers(1, 13))
    return Request(rid, L, T, Ls, beta, am, nsrc, int(rng.integers(0, 2**31 - 1)), tol, arrival)


def solve_request(req, stream=None, maxcg=200000):
    """Runs one request in the calling thread's bound context."""
    lx, ly, lt, ls = req.dims
    rng = np.random.default_rng(req.seed)
    th = (rng.standard_normal((lx, ly, lt, 3)) / np.sqrt(req.beta)).astype(np.float32)
    nsp = 4 * lx * ly * lt * ls
    b = rng.standard_normal((req.nsrc, nsp, 2)) / np.sqrt(2.0)
    theta = hr.DeviceArray.from_host(th.reshape(-1, order="F"))
    l64 = hr.DeviceArray((lx * ly * lt * 3, 2), np.float64)
    l32 = hr.DeviceArray((lx * ly * lt * 3, 2), np.float32)
    G.coef_dev(l64, theta, req.dims, stream=stream)
    G.links_to_f32_dev(l32, l64, req.dims, stream=stream)
    bd = hr.DeviceArray.from_host(b)
    x = hr.DeviceArray((req.nsrc, 1, nsp, 2), np.float64)
    sig = np.zeros(1)
    it, cgr, rel = G.multishift_fast_dev(l64, l32, req.dims, req.am, 3, sig, bd, x, tol=req.tol,
                                         maxcg=maxcg, nsrc=req.nsrc, th32=1e-4,
                                         true_residuals=True, stream=stream)
    xh = x.to_host()
    for a in (theta, l64, l32, bd, x):
        a.free()
    return {"iterations": int(it), "true_residual": rel[:, 0].tolist(),
            "x_norm": [float(np.linalg.norm(xh[k])) for k in range(req.nsrc)],
            "x": xh}


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

    def _pick(self):
        """First request that fits (oldest first); a request skipped too often blocks others."""
        free = self.hbm_budget - self.in_use
        for idx, r in enumerate(self.pending):
            if r.footprint() <= free:
                for older in list(self.pending)[:idx]:
                    self.skips[older.rid] += 1
                return idx
            if self.skips[r.rid] >= self.max_skip:
                return None  # hold back younger requests until this one fits
        return None

    def _schedule(self):
        while True:
            with self.lock:
                while not self.stop and (self.running >= self.slots or not self.pending
                                         or self._pick() is None):
                    self.lock.wait(0.05)
                if self.stop and not self.pending:
                    for _ in self.workers:
                        self.work.put(None)
                    return
                if self.running >= self.slots or not self.pending:
                    continue
                idx = self._pick()
                if idx is None:
                    continue
                req = self.pending[idx]
                del self.pending[idx]
                self.in_use += req.footprint()
                self.running += 1
            self.work.put(req)

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
                req.result = solve_request(req, stream=stream)
                if not self.keep:
                    req.result.pop("x", None)
            except Exception as e:  # a failed solve still frees its slot and memory
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
             request_fn=random_request):
    """Poisson arrivals at `rate`/s for `duration` s; returns summary metrics."""
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
    done = eng.drain()
    wall = time.perf_counter() - t0
    ok = [r for r in done if "error" not in r.result]
    lat = np.array([r.finish - r.arrival for r in ok]) if ok else np.zeros(1)
    solves = sum(r.nsrc for r in ok)
    worst = max((max(r.result["true_residual"]) for r in ok), default=0.0)
    return {"rate_per_s": rate, "requests": len(ok), "failed": len(done) - len(ok),
            "sources_solved": int(solves),
            "solves_per_s": round(solves / wall, 3), "requests_per_s": round(len(done) / wall, 3),
            "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
            "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
            "max_true_residual": worst, "wall_s": round(wall, 2)}
