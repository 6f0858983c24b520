"""Argument and size limits of the device entry points: every violation is
refused on the host with its status code, before anything is launched."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import thirring_hip as G
    G.init(0)
    return G


def _status(G, fn, *a, **k):
    with pytest.raises(G.ThrError) as e:
        fn(*a, **k)
    return int(str(e.value).split("status ")[1].split(")")[0])


def test_dslash_dev_limits(G):
    from hiprt import DeviceArray
    buf = DeviceArray((64,), np.float64)
    # 4 * V5 must fit 32-bit site indexing
    assert _status(G, G.dslash_dev, buf, buf, buf, (1024, 1024, 64, 64), 0.05) == 4
    assert _status(G, G.dslash_dev, buf, buf, buf, (4, 4, 0, 8), 0.05) == 1
    assert _status(G, G.dslash_dev, buf, buf, buf, (4, 4, 4, 8), [0.05] * 33, nrhs=33) == 1
    assert _status(G, G.dslash_dev, buf, buf, buf, (4, 4, 4, 8), 0.05, imass=2) == 1
    # FP16 spinors need their scale arrays
    assert _status(G, G.dslash_dev, buf, buf, buf, (4, 4, 4, 8), 0.05, prec="f16") == 1


def test_fast_multishift_limits(G):
    from hiprt import DeviceArray
    buf = DeviceArray((64,), np.float64)
    dims = (4, 4, 4, 8)
    # more than 32 shifts or sources, and more than 16 correction rounds
    assert _status(G, G.multishift_fast_dev, buf, buf, dims, 0.05, 3, np.ones(33), buf, buf,
                   tol=1e-10, maxcg=10) == 1
    assert _status(G, G.multishift_fast_dev, buf, buf, dims, 0.05, 3, np.ones(2), buf, buf,
                   tol=1e-10, maxcg=10, nsrc=33) == 1
    assert _status(G, G.multishift_fast_dev, buf, buf, dims, 0.05, 3, np.ones(2), buf, buf,
                   tol=1e-10, maxcg=10, corrections=17) == 1


def test_autotune_config_limits(G):
    dims = (4, 4, 4, 8)
    assert _status(G, G.dslash_set_config, dims, "f64", 0, 0) == 1
    assert _status(G, G.dslash_set_config, dims, "f64", 9, 0) == 1
    assert _status(G, G.dslash_set_config, dims, "f64", 1, 512) == 1
    # every defined flag (the y exchange is 256) is accepted and read back
    saved = G.dslash_get_config(dims, "f64")
    try:
        G.dslash_set_config(dims, "f64", 1, 256 | 22)
        assert G.dslash_get_config(dims, "f64") == (1, 278)
    finally:
        G.dslash_set_config(dims, "f64", *saved)
    from hiprt import DeviceArray
    buf = DeviceArray((3 * 4 * 4 * 4, 2), np.float64)
    assert _status(G, G.dslash_autotune, buf, dims, "f64", None, 0) == 1
    assert _status(G, G.dslash_autotune, buf, dims, "f64", None, 33) == 1


def test_autotune_over_rhs_batches(G):
    """thr_dslash_autotune_nrhs times shapes on batches of right-hand sides and
    pins a valid shape; the operator is unchanged by it."""
    from hiprt import DeviceArray
    dims = (8, 8, 4, 32)
    lx, ly, lt, ls = dims
    rng = np.random.default_rng(3)
    th = (rng.standard_normal((lx, ly, lt, 3)) / np.sqrt(0.3)).astype(np.float32)
    theta = DeviceArray.from_host(th.reshape(-1, order="F"))
    l64 = DeviceArray((lx * ly * lt * 3, 2), np.float64)
    G.coef_dev(l64, theta, dims)
    nsp = 4 * lx * ly * lt * ls
    xh = rng.standard_normal((4, nsp, 2))
    x = DeviceArray.from_host(xh)
    out = DeviceArray((4, nsp, 2), np.float64)
    before = G.dslash_get_config(dims, "f64")
    G.dslash_dev(out, x, l64, dims, 0.05, imass=3, nrhs=4)
    ref = out.to_host()
    tc, fl = G.dslash_autotune(l64, dims, "f64", nrhs=4)
    # pinned for batches; the single-RHS shape is untouched
    assert (tc, fl) == G.dslash_get_config(dims, "f64", nrhs=4)
    assert (tc, fl) == G.dslash_get_config(dims, "f64", nrhs=32)
    assert G.dslash_get_config(dims, "f64") == before
    assert 1 <= tc <= lt and 0 <= fl < 256
    G.dslash_dev(out, x, l64, dims, 0.05, imass=3, nrhs=4)
    got = out.to_host()
    # same operator (scalar-link shapes may place fused multiply-adds differently)
    assert np.linalg.norm(got - ref) <= 1e-15 * np.linalg.norm(ref) * 10
    G.dslash_set_config(dims, "f64", *before)
