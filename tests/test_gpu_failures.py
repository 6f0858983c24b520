"""GPU tests of failure reporting.

The reference runs a solver on a NaN source to its iteration cap and leaves
the caller to notice (measure_module.F90:124-135; dwf3d_lib.F90:860-881 aborts
when qmrherm hits max_qmr_iters).  Here every solver stops at the iteration
whose dot product is non-finite and returns THR_ERR_NONFINITE, and the solvers
whose contract is an FP64 true residual return THR_ERR_NOCONV when a pair
stays above it.  Each case prints its name before it runs.
"""
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


def _fields(dims, seed, nsrc=1):
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
    return th, l64, l32, bh


def _sigmas():
    return np.array([1e-4, 1e-3, 1e-2, 1e-1, 1.0])


@pytest.fixture(params=["0", "1"], ids=["streaming", "persistent"])
def solver_kind(request):
    """The batched multishift as streaming kernels (THR_PERSIST=0) and as the
    persistent cooperative kernel (THR_PERSIST=1)."""
    import os
    saved = os.environ.get("THR_PERSIST")
    os.environ["THR_PERSIST"] = request.param
    yield request.param
    if saved is None:
        os.environ.pop("THR_PERSIST", None)
    else:
        os.environ["THR_PERSIST"] = saved


def test_nan_source_fast_multishift_stops_within_one_block(G, solver_kind):
    from hiprt import DeviceArray
    print("case: NaN source, batched multishift (2 sources, one poisoned), THR_PERSIST",
          solver_kind, flush=True)
    dims = (4, 4, 4, 8)
    _, l64, l32, bh = _fields(dims, 11, nsrc=2)
    bh[1, 17, 0] = np.nan
    b = DeviceArray.from_host(bh)
    sig = _sigmas()
    x = DeviceArray((2, len(sig), bh.shape[1], 2), np.float64)
    with pytest.raises(G.ThrNonFinite) as e:
        G.multishift_fast_dev(l64, l32, dims, [0.05, 0.05], 3, sig, b, x, tol=1e-10,
                              maxcg=50000, nsrc=2)
    it, cgr, _ = e.value.result
    # |b|^2 is already NaN for the poisoned source: it never iterates, and
    # the call ends after the healthy source converges, far below the cap
    assert cgr[1] == 0
    assert it < 50000


def test_nan_link_fast_multishift(G, solver_kind):
    """A non-finite link makes alpha NaN in the first iteration: the solve
    ends within its first 32-iteration block."""
    from hiprt import DeviceArray
    print("case: NaN link, batched multishift, THR_PERSIST", solver_kind, flush=True)
    dims = (4, 4, 4, 8)
    th, l64, l32, bh = _fields(dims, 12)
    lh = l64.to_host()
    lh[5, 0] = np.inf
    l64b = DeviceArray.from_host(lh)
    l32b = DeviceArray.from_host(lh.astype(np.float32))
    b = DeviceArray.from_host(bh)
    sig = _sigmas()
    x = DeviceArray((1, len(sig), bh.shape[1], 2), np.float64)
    with pytest.raises(G.ThrNonFinite) as e:
        G.multishift_fast_dev(l64b, l32b, dims, 0.05, 3, sig, b, x, tol=1e-10, maxcg=50000)
    assert e.value.status == G.STATUS["nonfinite"]
    it, cgr, _ = e.value.result
    assert it <= 32 and cgr[0] <= 1


def test_nan_source_congrad(G):
    from hiprt import DeviceArray
    print("case: NaN source, reference congrad on device fields", flush=True)
    dims = (4, 4, 4, 8)
    _, l64, _, bh = _fields(dims, 13)
    bh[0, 3, 1] = np.nan
    phi = DeviceArray.from_host(bh[0])
    x = DeviceArray.from_host(np.zeros_like(bh[0]))
    with pytest.raises(G.ThrNonFinite):
        G.congrad_dev(phi, x, l64, dims, res=1e-8, am=0.05, imass=3)


def test_nan_source_congrad_host_array(G):
    """The drop-in congrad (haloed host arrays) reports the same error."""
    print("case: NaN source, drop-in congrad", flush=True)
    dims = (4, 4, 4, 8)
    rng = np.random.default_rng(14)
    th = (rng.standard_normal((4, 4, 4, 3)) / np.sqrt(0.3)).astype(np.float32)
    u = O.coef(th, dims)
    phi = D.rand_spinor(dims, rng)
    phi[2, 2, 2, 2, 1] = np.nan
    x = np.zeros_like(phi)
    with pytest.raises(G.ThrNonFinite):
        G.congrad(phi, x, u, dims, res=1e-8, am=0.05, imass=3)


def test_nan_source_reference_multishift(G):
    print("case: NaN source, reference multishift (FP64 and FP32 vectors)", flush=True)
    dims = (4, 4, 4, 8)
    rng = np.random.default_rng(15)
    th = (rng.standard_normal((4, 4, 4, 3)) / np.sqrt(0.3)).astype(np.float32)
    u = O.coef(th, dims)
    phi = D.rand_spinor(dims, rng)
    phi[1, 1, 1, 1, 0] = np.nan
    sig = _sigmas()
    for sp in (False, True):
        with pytest.raises(G.ThrNonFinite):
            G.multishift(u, dims, 0.05, 3, sig, np.ones(len(sig)), phi, 1e-8, 50000, sp=sp)


def test_iteration_cap_reports_no_convergence(G):
    from hiprt import DeviceArray
    print("case: batched multishift capped at 40 iterations", flush=True)
    dims = (4, 4, 4, 8)
    _, l64, l32, bh = _fields(dims, 16)
    b = DeviceArray.from_host(bh)
    sig = _sigmas()
    x = DeviceArray((1, len(sig), bh.shape[1], 2), np.float64)
    # with true residuals: the FP64 check finds the pairs above tol
    with pytest.raises(G.ThrNoConvergence) as e:
        G.multishift_fast_dev(l64, l32, dims, 0.01, 3, sig, b, x, tol=1e-10, maxcg=40)
    it, _, rel = e.value.result
    assert it <= 40 + 32 and rel.max() > 1e-10
    # without them: shifts still above the iterated target at maxcg
    with pytest.raises(G.ThrNoConvergence):
        G.multishift_fast_dev(l64, l32, dims, 0.01, 3, sig, b, x, tol=1e-10, maxcg=40,
                              true_residuals=False)
    # the same call with room to converge succeeds
    it, _, rel = G.multishift_fast_dev(l64, l32, dims, 0.01, 3, sig, b, x, tol=1e-10,
                                       maxcg=50000)
    assert rel.max() <= 1e-10


def test_mixed_qmrherm_reports_unconverged_shifts(G):
    """qmrherm with use_sp = 2 (the batched mixed solver with FP64 correction)
    must not hand an unconverged shift to the force: with max_qmr_iters too
    small for any correction round to reach res it returns THR_ERR_NOCONV."""
    print("case: mixed qmrherm capped at 3 iterations", flush=True)
    dims = (4, 4, 4, 8)
    rng = np.random.default_rng(17)
    th = (rng.standard_normal((4, 4, 4, 3)) / np.sqrt(0.3)).astype(np.float32)
    u = O.coef(th, dims)
    phi = D.rand_spinor(dims, rng)
    sig = _sigmas()
    anum = np.concatenate([[0.5], np.ones(len(sig))])
    with pytest.raises(G.ThrNoConvergence):
        G.qmrherm(phi, 1e-9, 0.01, 3, anum, sig, 0, u, dims, use_sp="mixed", max_qmr_iters=3)
    x, it, *_ = G.qmrherm(phi, 1e-9, 0.01, 3, anum, sig, 0, u, dims, use_sp="mixed")
    assert it > 3
