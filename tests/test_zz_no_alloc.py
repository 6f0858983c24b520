"""The allocation-free congrad (thr_congrad_dev_noalloc, scratch in the
library's fixed device array) against thr_congrad_dev (measure_module.F90:
9-140 on device fields, library workspace): two calls in a row, the second
on a smaller lattice, each giving the same iteration counts and the same
solution bit for bit (same kernels, same launch shapes, same order of
summation); a lattice needing more scratch than the array holds is refused
on the host with THR_ERR_SIZE."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import thirring_hip as G
    G.init(0)
    return G


def _fields(G, dims, seed):
    from hiprt import DeviceArray
    lx, ly, lt, ls = dims
    rng = np.random.default_rng(seed)
    th = (rng.standard_normal((lx, ly, lt, 3)) / np.sqrt(0.3)).astype(np.float32)
    thd = DeviceArray.from_host(th.reshape(-1, order="F"))
    l64 = DeviceArray((lx * ly * lt * 3, 2), np.float64)
    G.coef_dev(l64, thd, dims)
    n = 4 * lx * ly * lt * ls
    phi = rng.standard_normal((n, 2)) / np.sqrt(2)
    x0 = 0.1 * rng.standard_normal((n, 2))
    return l64, DeviceArray.from_host(phi), x0


@pytest.mark.parametrize("imass", [3])
def test_noalloc_congrad_matches_congrad_dev(G, imass):
    from hiprt import DeviceArray
    assert G.lib().thr_noalloc_scratch_bytes() == 64 << 20
    for dims, seed in (((8, 8, 8, 16), 41), ((4, 4, 8, 8), 42)):
        print("case: congrad_dev_noalloc vs congrad_dev", dims, "seed", seed, flush=True)
        l64, phi, x0 = _fields(G, dims, seed)
        xa = DeviceArray.from_host(x0)
        xb = DeviceArray.from_host(x0)
        ita = G.congrad_dev(phi, xa, l64, dims, 1e-9, 0.05, imass)
        itb = G.congrad_dev_noalloc(phi, xb, l64, dims, 1e-9, 0.05, imass)
        a, b = xa.to_host(), xb.to_host()
        print("itercg", ita, itb, "max |dx|", float(np.abs(a - b).max()), flush=True)
        assert ita == itb
        assert ita[0] > 5
        assert np.array_equal(a, b)


def test_noalloc_refuses_large_lattice(G):
    """32^3 x 64 needs ~3.2 GB of scratch: refused before any device work."""
    import ctypes
    from hiprt import DeviceArray
    dims = (ctypes.c_int * 4)(32, 32, 32, 64)
    it, nx = ctypes.c_int(), ctypes.c_int()
    small = DeviceArray((16, 2), np.float64)  # never read: the size check comes first
    rc = G.lib().thr_congrad_dev_noalloc(G._p(small), G._p(small), G._p(small), dims, 1e-9, 0.05,
                                         3, 1.0, 100, ctypes.byref(it), ctypes.byref(nx), None)
    small.free()
    assert rc == G.STATUS["size"]
    assert b"scratch" in G.lib().thr_last_error()
