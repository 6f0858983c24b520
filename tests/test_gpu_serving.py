"""Serving loop: requests give the same result alone and in a concurrent mix
(identical to the last bit: each request runs in its own context with
deterministic reductions), every source reaches FP64 true residual 1e-10, and
admission respects the HBM budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _reqs():
    import serving
    return [serving.Request(0, 8, 8, 8, 0.3, 0.05, 2, 11),
            serving.Request(1, 8, 16, 16, 0.4, 0.02, 3, 12),
            serving.Request(2, 12, 12, 8, 0.25, 0.01, 1, 13),
            serving.Request(3, 8, 8, 16, 0.5, 0.005, 4, 14)]


def test_alone_equals_mixed():
    import serving
    import thirring_hip as G
    G.init(0)
    alone = {}
    for r in _reqs():
        res = serving.solve_request(r)
        assert max(res["true_residual"]) <= 1e-10
        alone[r.rid] = res["x"]
    eng = serving.ServingEngine(slots=3, keep_solutions=True)
    for r in _reqs():
        eng.submit(r)
    done = eng.drain(timeout=240)
    assert len(done) == 4
    for r in done:
        assert max(r.result["true_residual"]) <= 1e-10
        assert np.array_equal(r.result["x"], alone[r.rid]), r.rid


def test_hbm_admission_serialises_big_requests():
    import serving
    import thirring_hip as G
    G.init(0)
    # equal footprints: any two together exceed the budget
    reqs = [serving.Request(k, 8, 8, 8, 0.3 + 0.05 * k, 0.05, 2, 20 + k) for k in range(3)]
    budget = reqs[0].footprint() * 1.5
    eng = serving.ServingEngine(slots=3, hbm_budget=budget)
    for r in reqs:
        eng.submit(r)
    done = eng.drain(timeout=240)
    assert len(done) == 3
    # with room for only one request at a time the solves never overlap
    iv = sorted((r.start, r.finish) for r in done)
    for (s0, f0), (s1, f1) in zip(iv, iv[1:]):
        assert s1 >= f0 - 1e-6


def test_poisson_load_summary():
    import serving
    import thirring_hip as G
    G.init(0)

    def small(rid, rng, arrival=0.0, tol=1e-10):
        return serving.Request(rid, 8, 8, 8, float(rng.uniform(0.2, 0.5)), 0.05,
                               int(rng.integers(1, 4)), int(rng.integers(0, 2**31 - 1)), tol, arrival)

    m = serving.run_load(rate=20.0, duration=2.0, slots=4, request_fn=small, max_requests=30)
    assert m["requests"] >= 1
    assert m["max_true_residual"] <= 1e-10
    assert m["p99_latency_s"] >= m["p50_latency_s"] > 0
