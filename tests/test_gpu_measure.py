"""psibarpsi measurement (measure_shamir, measure_module.F90:170-423) on the
GPU: the batched form (all wall sources of a measurement in multi-RHS
launches of the batched solver, FP64 true residual 1e-10) against the
reference-order form (one congrad per source, converged tightly).  Same
noise draws, same contractions: psibarpsi agrees to 1e-6 relative.  The
reference-order form itself reproduces the reference's known answer
(test_gpu_parity.py::test_measure_known_answers)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import thirring_hip as G
    G.init(0)
    return G


@pytest.mark.parametrize("imass,knoise", [(3, 10), (1, 3)])
def test_measure_batched_matches_reference_order(G, imass, knoise):
    import measure
    from hiprt import DeviceArray
    from site_random import SiteRandom
    dims = (8, 8, 8, 16)
    lx, ly, lt, ls = dims
    rng = np.random.default_rng(5)
    th = (rng.standard_normal((lx, ly, lt, 3)) / np.sqrt(0.4)).astype(np.float32)
    u = O.coef(th, dims)
    thd = DeviceArray.from_host(th.reshape(-1, order="F"))
    l64 = DeviceArray((lx * ly * lt * 3, 2), np.float64)
    l32 = DeviceArray((lx * ly * lt * 3, 2), np.float32)
    G.coef_dev(l64, thd, dims)
    G.links_to_f32_dev(l32, l64, dims)
    am = np.float32(0.05)
    ref, aviter = measure.measure_shamir(u, dims, np.float32(1e-10), am, imass,
                                         SiteRandom(lx, ly, lt, 7.0), knoise=knoise)
    got, info = measure.measure_shamir_batched(l64, l32, dims, am, imass,
                                               SiteRandom(lx, ly, lt, 7.0), knoise=knoise)
    print("psibarpsi", ref, got, "aviter", aviter, info, flush=True)
    assert info["sources"] == 4 * knoise
    assert info["max_true_residual"] <= 1e-10
    assert abs(got - ref) <= 1e-6 * abs(ref)


@pytest.mark.parametrize("imass", [1, 3, 5])
def test_measure_noise_entry_point_matches_oracle(G, imass):
    """thr_measure_shamir_noise (one noise vector's four wall solves on the
    device) against the same measurement composed from the oracle's dslashd
    and congrad (measure_module.F90:210-372): equal psibarpsi and iteration
    counts at most one apart per solve on average (the two sum the CG dot
    products in different orders, which can move a stopping iteration by
    one), for every imass (imass 5 reads at the source wall)."""
    import measure
    from site_random import SiteRandom
    dims = (4, 4, 8, 8)
    lx, ly, lt, ls = dims
    print("case: measure_shamir_noise imass", imass, "4x4x8 Ls 8 seed 31", flush=True)
    rng = np.random.default_rng(31)
    # theta std 0.5: ~134 iterations per solve (rough fields at beta 0.4 take
    # ~650, where the oracle's own OpenMP sums move the count by a few)
    th = (rng.standard_normal((lx, ly, lt, 3)) / np.sqrt(4.0)).astype(np.float32)
    u = O.coef(th, dims)
    res, am = np.float32(1e-8), np.float32(0.05)
    got, itg = measure.measure_shamir(u, dims, res, am, imass, SiteRandom(lx, ly, lt, 9.0),
                                      knoise=2)
    ref, ito = measure.measure_shamir(u, dims, res, am, imass, SiteRandom(lx, ly, lt, 9.0),
                                      knoise=2, ops=O)
    print("psibarpsi", got, ref, "aviter", itg, ito, flush=True)
    assert abs(itg - ito) <= max(1.0, 0.01 * ito)  # aviter: iterations per solve
    assert abs(got - ref) <= 1e-6 * abs(ref)


@pytest.mark.parametrize("imass", [1, 3])
def test_measure_wilson_entry_point_matches_oracle(G, imass):
    """thr_measure_wilson_noise (measure_wilson's eight wall solves per noise
    vector with the DWilson projection of the read slices, on the device)
    against the same measurement composed from the oracle's dslashd (Wilson
    kernel), congrad and DWilson slice (tests/measure_ref.py,
    measure_module.F90:472-606): equal psibarpsi, iteration counts at most
    one apart per solve on average."""
    import measure
    import measure_ref
    from site_random import SiteRandom
    dims = (4, 4, 8, 8)
    lx, ly, lt, ls = dims
    print("case: measure_wilson_noise imass", imass, "4x4x8 Ls 8 seed 33", flush=True)
    rng = np.random.default_rng(33)
    th = (rng.standard_normal((lx, ly, lt, 3)) / np.sqrt(4.0)).astype(np.float32)
    u = O.coef(th, dims)
    res, am = np.float32(1e-8), np.float32(0.05)
    got, itg = measure.measure_wilson(u, dims, res, am, imass, SiteRandom(lx, ly, lt, 9.0),
                                      knoise=2)

    def dslashd_w(r, uu, d, m, im):
        return O.dslashd(r, uu, d, m, im, kernel="wilson")

    ref, ito = measure_ref.measure_wilson(dslashd_w, O.congrad, O.dwilson_slice, u, dims, res, am,
                                          imass, SiteRandom(lx, ly, lt, 9.0), knoise=2)
    print("psibarpsi", got, ref, "aviter", itg, ito, flush=True)
    assert abs(itg - ito) <= max(1.0, 0.01 * ito)
    assert abs(got - ref) <= 1e-6 * abs(ref)
