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
        res = serving.solve_request(r, keep_x=True)
        assert max(res["true_residual"]) <= 1e-10
        # norms reduced on the device are those of the returned solutions
        xn = [float(np.linalg.norm(res["x"][k])) for k in range(r.nsrc)]
        assert np.allclose(res["x_norm"], xn, rtol=1e-13, atol=0)
        alone[r.rid] = res["x"]
    eng = serving.ServingEngine(slots=3, keep_solutions=True)
    for r in _reqs():
        eng.submit(r)
    done = eng.drain(timeout=240)
    assert len(done) == 4
    for r in done:
        assert max(r.result["true_residual"]) <= 1e-10
        assert np.array_equal(r.result["x"], alone[r.rid]), r.rid


def test_footprint_bounds_real_scratch_and_trim_releases_it():
    """The admission estimate bounds what a request really holds (library
    scratch plus its fields), and a worker gives its scratch back after each
    request, so a big request followed by small ones never leaves the big
    one's scratch allocated."""
    import serving
    import thirring_hip as G
    G.init(0)
    big = serving.Request(0, 16, 16, 32, 0.3, 0.01, 4, 31)
    small = [serving.Request(k, 8, 8, 8, 0.3, 0.05, 1, 31 + k) for k in range(1, 4)]
    eng = serving.ServingEngine(slots=1, hbm_budget=1.2 * big.footprint())
    for r in [big] + small:
        eng.submit(r)
    done = eng.drain(timeout=240)
    assert len(done) == 4
    for r in done:
        assert "error" not in r.result, r.result
        v4 = r.L * r.L * r.T
        fields = 3 * v4 * (4 + 16 + 8) + 2 * r.nsrc * 4 * v4 * r.Ls * 16
        assert r.result["scratch_bytes"] + fields <= r.footprint(), r.rid
        assert r.result["scratch_after_trim"] == 0


def test_device_gaussian_fields():
    """The counter-based generator: reproducible, independent of the launch
    (a prefix of a field is the field of the shorter length), mean 0 and
    the requested variance, and the values of a numpy restatement of the same
    mixing and Box-Muller steps."""
    import hiprt as hr
    import thirring_hip as G
    G.init(0)
    n = 1 << 20
    a = hr.DeviceArray((n,), np.float64)
    b = hr.DeviceArray((n,), np.float64)
    G.gaussian_dev(a, n, 0.7, 12345)
    G.gaussian_dev(b, n - 3, 0.7, 12345)
    ah, bh = a.to_host(), b.to_host()
    assert np.array_equal(ah[:n - 3], bh[:n - 3])
    assert abs(ah.mean()) < 5 * 0.7 / np.sqrt(n)
    assert abs(ah.var() / 0.49 - 1) < 5 * np.sqrt(2.0 / n)
    f = hr.DeviceArray((n,), np.float32)
    G.gaussian_dev(f, n, 0.7, 12345, prec="f32")
    assert np.array_equal(f.to_host(), ah.astype(np.float32))
    # restatement of the generator for the first pairs
    M = (1 << 64) - 1

    def mix64(z):
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)
    for k in range(64):
        h1 = mix64((12345 + (k + 1) * 0x9E3779B97F4A7C15) & M)
        h2 = mix64(h1 ^ 0xD1B54A32D192ED03)
        u1 = ((h1 >> 11) + 1.0) * 2.0 ** -53
        u2 = (h2 >> 11) * 2.0 ** -53
        r = 0.7 * np.sqrt(-2.0 * np.log(u1))
        ref = (r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2))
        assert np.allclose(ah[2 * k:2 * k + 2], ref, rtol=1e-12, atol=1e-14), k


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
