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
    assert _status(G, G.dslash_set_config, dims, "f64", 1, 64) == 1
