"""GPU tests of the persistent (cooperative, one launch per solve) batched
multishift against the oracle and against the streaming blocked solver.

The persistent kernel evaluates the operator per site (its own code, not the
streaming Dslash kernels), so it is checked against the oracle's multishift
for every wall-mass type (imass 1, 3 and 5), on an odd lattice, with each
block sweeping several site ranges, and for several sources with different
masses in one grid.  Every shift must meet the FP64 true residual 1e-10 and
match the oracle's converged solution to 1e-8.
"""
import os

import numpy as np
import pytest

import dense as D
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import thirring_hip as G
    G.init(0)
    return G


@pytest.fixture
def env():
    keys = ("THR_PERSIST", "THR_PERSIST_BLOCKS", "THR_GRAPHS")
    saved = {k: os.environ.get(k) for k in keys}
    yield os.environ
    for k, v in saved.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


def _remez4():
    path = os.path.join(os.path.dirname(__file__), "..", "thirring-rhmc_amd", "data", "remez4")
    rows = [l.replace(",", " ").split() for l in open(path) if l.strip()]
    blk = rows[26:52]
    return np.array([float(r[1]) for r in blk[1:]])


def _setup(dims, seed, nsrc):
    from hiprt import DeviceArray
    import thirring_hip as G
    lx, ly, lt, ls = dims
    rng = np.random.default_rng(seed)
    th = (rng.standard_normal((lx, ly, lt, 3)) / np.sqrt(0.3)).astype(np.float32)
    theta = DeviceArray.from_host(th.reshape(-1, order="F"))
    l64 = DeviceArray((lx * ly * lt * 3, 2), np.float64)
    l32 = DeviceArray((lx * ly * lt * 3, 2), np.float32)
    G.coef_dev(l64, theta, dims)
    G.links_to_f32_dev(l32, l64, dims)
    nsp = 4 * lx * ly * lt * ls
    bh = rng.standard_normal((nsrc, nsp, 2)) / np.sqrt(2)
    return th, l64, l32, bh, DeviceArray.from_host(bh)


def _oracle_check(th, dims, am, imass, sig, bh, xs, shifts):
    u = O.coef(th, dims)
    c = bh[..., 0] + 1j * bh[..., 1]
    out, _, _ = O.multishift(u, dims, am, imass, sig, np.ones(len(sig)), D.from_vec(c, dims),
                             res=1e-12, maxcg=20000)
    for i in shifts:
        xi = xs[i, :, 0] + 1j * xs[i, :, 1]
        oi = out[..., i].reshape(-1, order="F")
        err = float(np.linalg.norm(xi - oi) / np.linalg.norm(oi))
        assert err <= 1e-8, (imass, i, err)


@pytest.mark.parametrize("imass", [1, 3, 5])
def test_persistent_against_oracle(G, env, imass):
    from hiprt import DeviceArray
    dims = (4, 4, 6, 8)
    sig = _remez4()
    th, l64, l32, bh, b = _setup(dims, 11, 1)
    n, nsp = len(sig), bh.shape[1]
    env["THR_PERSIST"] = "1"
    x = DeviceArray((1, n, nsp, 2), np.float64)
    st = {}
    it, cgr, rel = G.multishift_fast_dev(l64, l32, dims, 0.05, imass, sig, b, x, tol=5e-11,
                                         maxcg=20000, stats=st)
    print("imass", imass, "iterations", int(cgr[0]), "max relres", rel.max(), flush=True)
    assert st["switches"] == [0, -1, -1] and st["corr_iterations"] == 0
    assert rel.max() <= 1e-10
    _oracle_check(th, dims, 0.05, imass, sig, bh[0], x.to_host()[0], (0, 6, 12, 24))


def test_persistent_equals_streaming_solver(G, env):
    """Three sources with their own masses in one grid, on an odd lattice,
    each block sweeping several site ranges: the same iteration counts as the
    streaming FP64 solver (same recurrences; the sums differ in order only)
    and the same solutions to 1e-8."""
    from hiprt import DeviceArray
    dims = (3, 5, 7, 6)
    sig = _remez4()[::2]
    th, l64, l32, bh, b = _setup(dims, 12, 3)
    n, nsp = len(sig), bh.shape[1]
    masses = [0.01, 0.05, 0.02]
    res = {}
    for label, e in (("persistent", {"THR_PERSIST": "1", "THR_PERSIST_BLOCKS": "1"}),
                     ("persistent_wide", {"THR_PERSIST": "1"}),
                     ("streaming", {"THR_PERSIST": "0", "THR_GRAPHS": "0"})):
        env.pop("THR_PERSIST_BLOCKS", None)
        env.update(e)
        x = DeviceArray((3, n, nsp, 2), np.float64)
        it, cgr, rel = G.multishift_fast_dev(l64, l32, dims, masses, 3, sig, b, x, tol=5e-11,
                                             maxcg=20000, nsrc=3)
        print(label, "cg_return", list(cgr), "max relres", rel.max(), flush=True)
        assert rel.max() <= 1e-10
        res[label] = (np.array(cgr), x.to_host())
    for label in ("persistent", "persistent_wide"):
        assert np.abs(res[label][0] - res["streaming"][0]).max() <= 1
        d = float(np.linalg.norm(res[label][1] - res["streaming"][1])
                  / np.linalg.norm(res["streaming"][1]))
        assert d <= 1e-8, (label, d)
    # one block per source and many blocks per source sum the partials in
    # different orders: the recurrences agree to rounding, so a stopping test
    # may flip by one iteration
    assert np.abs(res["persistent"][0] - res["persistent_wide"][0]).max() <= 1
    _oracle_check(th, dims, masses[1], 3, sig, bh[1], res["persistent"][1][1], (0, len(sig) - 1))


def test_persistent_nonfinite_and_cap(G, env):
    """A NaN in one source stops that source (ThrNonFinite) while the grid
    still drains; an iteration cap below convergence reports the unconverged
    shifts (ThrNoConvergence)."""
    from hiprt import DeviceArray
    dims = (4, 4, 4, 4)
    sig = _remez4()[::4]
    th, l64, l32, bh, b = _setup(dims, 13, 2)
    n, nsp = len(sig), bh.shape[1]
    env["THR_PERSIST"] = "1"
    bad = bh.copy()
    bad[1, 7, 0] = np.nan
    bb = DeviceArray.from_host(bad)
    x = DeviceArray((2, n, nsp, 2), np.float64)
    with pytest.raises(G.ThrNonFinite):
        G.multishift_fast_dev(l64, l32, dims, 0.05, 3, sig, bb, x, tol=1e-10, maxcg=2000, nsrc=2,
                              true_residuals=False)
    with pytest.raises(G.ThrNoConvergence):
        G.multishift_fast_dev(l64, l32, dims, 0.05, 3, sig, b, x, tol=1e-10, maxcg=5, nsrc=2,
                              true_residuals=False)
    # the library still works after both
    it, cgr, rel = G.multishift_fast_dev(l64, l32, dims, 0.05, 3, sig, b, x, tol=1e-10,
                                         maxcg=2000, nsrc=2)
    assert rel.max() <= 1e-10
