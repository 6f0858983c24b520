"""Host logic of the serving loop (no GPU): admission by HBM footprint, slot
limits, first-fit with anti-starvation, latency bookkeeping.  The solve is
replaced by a stub that sleeps in proportion to the request size."""
import threading
import time

import pytest


@pytest.fixture
def serving(monkeypatch):
    import serving as S
    import thirring_hip

    class _Ctx:
        def bind(self):
            pass

        @staticmethod
        def unbind():
            pass

        def destroy(self):
            pass

        def trim(self):
            pass

        def scratch_bytes(self):
            return 0

    class _Stream:
        pass

    monkeypatch.setattr(thirring_hip, "Context", _Ctx)
    monkeypatch.setattr(S.hr, "Stream", _Stream)
    live = {"n": 0, "max": 0, "bytes": 0, "max_bytes": 0}
    lock = threading.Lock()

    def fake_solve(req, stream=None, keep_x=False):
        if req.seed == 666:
            raise RuntimeError("solver failure")
        with lock:
            live["n"] += 1
            live["bytes"] += req.footprint()
            live["max"] = max(live["max"], live["n"])
            live["max_bytes"] = max(live["max_bytes"], live["bytes"])
        time.sleep(0.002 * req.nsrc)
        with lock:
            live["n"] -= 1
            live["bytes"] -= req.footprint()
        return {"iterations": 1, "true_residual": [1e-11] * req.nsrc,
                "x_norm": [1.0] * req.nsrc}

    monkeypatch.setattr(S, "solve_request", fake_solve)
    S._live = live
    return S


def test_slots_and_budget_respected(serving):
    S = serving
    reqs = [S.Request(k, 8, 8, 8, 0.3, 0.05, 1 + k % 4, k) for k in range(20)]
    budget = 2.5 * max(r.footprint() for r in reqs)
    eng = S.ServingEngine(slots=3, hbm_budget=budget)
    for r in reqs:
        eng.submit(r)
    done = eng.drain(timeout=30)
    assert len(done) == 20
    assert S._live["max"] <= 3
    assert S._live["max_bytes"] <= budget
    for r in done:
        assert r.finish >= r.start >= r.arrival


def test_large_request_not_starved(serving):
    """A request that does not fit next to running ones is overtaken by
    exactly max_skip younger requests, then holds the queue until it fits."""
    S = serving
    big = S.Request(0, 24, 48, 48, 0.3, 0.05, 12, 1)
    # small requests take 24 ms each in the stub, so the first is still
    # running when the scheduler looks at the big one
    small = [S.Request(k, 8, 8, 8, 0.3, 0.05, 12, k) for k in range(1, 30)]
    # the big request fits only when nothing else runs
    budget = big.footprint() + 0.5 * small[0].footprint()
    for max_skip in (2, 3):
        eng = S.ServingEngine(slots=4, hbm_budget=budget, max_skip=max_skip)
        eng.submit(small[0])
        eng.submit(big)
        for r in small[1:]:
            eng.submit(r)
        done = eng.drain(timeout=60)
        assert len(done) == 30
        order = [r.rid for r in sorted(done, key=lambda r: r.start)]
        # small[0] (older) first, then max_skip younger ones, then the big one
        assert order.index(0) == 1 + max_skip, (max_skip, order[:8])


def test_probe_has_no_side_effects(serving):
    S = serving
    eng = S.ServingEngine(slots=1, hbm_budget=1e12)
    with eng.lock:
        eng.stop = True  # keep the scheduler idle while the queue is examined
    eng.drain(timeout=5)
    reqs = [S.Request(k, 8, 8, 8, 0.3, 0.05, 1, k) for k in range(3)]
    eng.pending.extend(reqs)
    eng.in_use = eng.hbm_budget  # nothing fits
    for _ in range(5):
        assert eng._probe() is None
    eng.in_use = 0
    for _ in range(5):
        assert eng._probe() == 0
    assert sum(eng.skips.values()) == 0
    eng._commit(2)
    assert eng.skips[0] == 1 and eng.skips[1] == 1 and eng.skips[2] == 0


def test_footprint_counts_solver_scratch(serving):
    """63 FP64 spinors per source (FP64 and FP32 Krylov storage, the
    correction and reliable-update batches, source, solution) plus the link
    fields."""
    S = serving
    r = S.Request(0, 8, 16, 8, 0.3, 0.05, 3, 0)
    spinor = 4 * 8 * 8 * 16 * 8 * 16
    assert r.footprint() == 3 * 8 * 8 * 16 * 28 + 3 * 63 * spinor + (1 << 20)


def test_rate_sweep_sheds_overload(serving):
    S = serving

    def slow(rid, rng, arrival=0.0, tol=1e-10):
        return S.Request(rid, 8, 8, 8, 0.3, 0.05, 12, rid, tol, arrival)  # 24 ms each

    out = S.rate_sweep([5.0, 2000.0], window=0.4, slots=1, request_fn=slow)
    low, high = out["sweep"]
    assert low["shed"] == 0 and low["requests"] == low["submitted"]
    assert high["shed"] > 0 and high["requests"] + high["shed"] == high["submitted"]
    assert out["highest_sustainable_rate_per_s"] == 5.0


def test_run_load_metrics(serving):
    S = serving

    def small(rid, rng, arrival=0.0, tol=1e-10):
        return S.Request(rid, 8, 8, 8, 0.3, 0.05, 1, rid, tol, arrival)

    m = S.run_load(rate=200.0, duration=0.3, slots=4, request_fn=small)
    assert m["requests"] >= 10
    assert m["p99_latency_s"] >= m["p50_latency_s"] > 0
    assert m["solves_per_s"] > 0


def test_failed_solve_frees_its_slot(serving):
    S = serving
    reqs = [S.Request(k, 8, 8, 8, 0.3, 0.05, 1, 666 if k == 2 else k) for k in range(8)]
    eng = S.ServingEngine(slots=1, hbm_budget=1.5 * reqs[0].footprint())
    for r in reqs:
        eng.submit(r)
    done = eng.drain(timeout=30)
    assert len(done) == 8
    failed = [r for r in done if "error" in r.result]
    assert [r.rid for r in failed] == [2] and "solver failure" in failed[0].result["error"]


def test_request_over_budget_is_rejected(serving):
    S = serving
    small = S.Request(0, 8, 8, 8, 0.3, 0.05, 1, 0)
    huge = S.Request(1, 24, 48, 48, 0.3, 0.05, 12, 1)
    eng = S.ServingEngine(slots=2, hbm_budget=2 * small.footprint())
    eng.submit(huge)
    eng.submit(small)
    done = eng.drain(timeout=30)
    assert len(done) == 2
    byid = {r.rid: r for r in done}
    assert "exceeds the HBM budget" in byid[1].result["error"]
    assert "error" not in byid[0].result
