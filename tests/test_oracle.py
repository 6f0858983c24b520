"""Pins the CPU oracle (oracle/oracle.c) without the reference binary.

The reference Fortran cannot be built here under the project's rules (its
MPI build pulls in sed-generated *_sp modules and a hand-made `mpi` module),
so every oracle routine is pinned by what the mathematics fixes, chosen so a
plausible slip (sign, index, conjugation, projector, coefficient) fails:

  * reference known answer: congrad itercg == 27 (tests/test_congrad.F90:47)
  * Clifford algebra of the gamma tables (gammamatrices.F90:27-125)
  * dense brute-force assembly of M from those tables vs. the oracle
  * imass = 1 wall term vs. the reference's operator definition
    (notes/fermions-on-lattice.tex:196-199); imass = 5 condensate == imass = 3
    condensate on every gauge field (chiral rotation), with imass = 3's sign
    fixed by the psibarpsi known answer
  * adjointness <a, M b> = <M^+ a, b> (Shamir and Wilson, imass 1/3/5)
  * U(1) gauge covariance (link direction and conjugation)
  * Wilson-kernel DWF == Shamir DWF dressed with (1 - D_W) on every s-coupling
  * derivs == exact finite difference of Re<R, M^(+) X2> (M is affine in u)
  * multishift iterate k == plain CG iterate k of each shifted system; dense solves
  * qmrherm: rational function via dense solves; iflag 2/3 forces vs. finite
    differences of the pseudofermion action
  * remez4 block 2 approximates x^(-1/4) on [1e-4, 64] to < 5e-10
"""
import numpy as np
import pytest

import dense as D
import oracle as O
import measure
import measure_ref
import refstate


def ip(a, b):
    return np.vdot(D.interior_vec(a), D.interior_vec(b))


TINY = (2, 2, 2, 4)
SMALL = (3, 2, 4, 4)


def test_congrad_reference_known_answer():
    """tests/test_congrad.F90:30-47: 12^3 x 24, res 0.1, am 0.05, imass 3 -> 27."""
    dims = (12, 12, 12, 24)
    st = refstate.starting_state(dims)
    it, nx = O.congrad(st["phi"], st["x"], st["u"], dims, res=0.1, am=0.05, imass=3)
    assert it == 27 and nx == 27
    # the stopping rule is absolute: |r|^2 < 4*V5*res^2 (measure_module.F90:41)
    r = st["phi"].copy(order="F")
    mx = O.dirac_op_shifted(st["x"], st["u"], dims, 0.05, 3)
    resid2 = np.sum(np.abs(D.interior_vec(r) - D.interior_vec(mx)) ** 2)
    assert resid2 < 4 * 12 * 12 * 12 * 24 * np.float32(0.1) ** 2


def test_gamma_clifford():
    G = D.gamma_tables()
    for m in range(4):
        assert np.allclose(G[m], G[m].conj().T)
        for n in range(4):
            assert np.allclose(G[m] @ G[n] + G[n] @ G[m], 2 * np.eye(4) * (m == n))
    assert np.allclose(G[4], G[0] @ G[1] @ G[2] @ G[3])
    assert np.allclose(G[5], G[3] @ G[4])


def _shift_matrix(dims, mu, u):
    """T+ (psi)(x) = U(x,mu) psi(x+mu) on the 4-D interior, periodic."""
    lx, ly, lt, ls = dims
    n4 = lx * ly * lt
    T = np.zeros((n4, n4), dtype=np.complex128)
    step = [(1, 0, 0), (0, 1, 0), (0, 0, 1)][mu]
    for t in range(lt):
        for y in range(ly):
            for x in range(lx):
                i = x + lx * (y + ly * t)
                xn, yn, tn = (x + step[0]) % lx, (y + step[1]) % ly, (t + step[2]) % lt
                j = xn + lx * (yn + ly * tn)
                T[i, j] += u[x + 1, y + 1, t + 1, mu]
    return T


def _dense_shamir(dims, u, am, imass, am3=1.0, dag=False):
    """Independent dense assembly (Kronecker form, ordering s fastest, then x,y,t, spin)."""
    lx, ly, lt, ls = dims
    G = D.gamma_tables()
    kap = 0.5
    n4 = lx * ly * lt
    I4s, Is, Ix = np.eye(4), np.eye(ls), np.eye(n4)
    diag = float(np.float32(np.float32(3.0) - np.float32(am3)) + np.float32(1.0))
    M = diag * np.kron(I4s, np.kron(Ix, Is)).astype(np.complex128)
    sgn = -1.0 if dag else 1.0
    for mu in range(3):
        Tp = _shift_matrix(dims, mu, u)
        Tm = Tp.conj().T
        M += np.kron(I4s, np.kron(-kap * (Tp + Tm), Is))
        M += sgn * np.kron(kap * G[mu], np.kron(Tp - Tm, Is))
    Pp = np.diag([1.0, 1.0, 0, 0])
    Pm = np.diag([0, 0, 1.0, 1.0])
    up = np.eye(ls, k=1)   # (S+ v)(s) = v(s+1)
    dn = np.eye(ls, k=-1)  # (S- v)(s) = v(s-1)
    e1Ls = np.zeros((ls, ls)); e1Ls[0, ls - 1] = 1.0   # s=1 <- s=Ls
    eLs1 = np.zeros((ls, ls)); eLs1[ls - 1, 0] = 1.0   # s=Ls <- s=1
    e11 = np.zeros((ls, ls)); e11[0, 0] = 1.0
    eLL = np.zeros((ls, ls)); eLL[ls - 1, ls - 1] = 1.0
    m = float(np.float32(am))
    P_hi_in, P_lo_in = (Pp, Pm) if not dag else (Pm, Pp)
    M -= np.kron(P_lo_in, np.kron(Ix, up)) + np.kron(P_hi_in, np.kron(Ix, dn))
    swap = np.zeros((4, 4)); swap[0, 2] = swap[1, 3] = 1.0   # spins 1:2 <- 3:4
    if imass == 1:
        M += m * (np.kron(P_lo_in, np.kron(Ix, eLs1)) + np.kron(P_hi_in, np.kron(Ix, e1Ls)))
    elif imass == 3:
        c = 1j * m if not dag else 1j * m
        if not dag:
            M += c * np.kron(Pm, np.kron(Ix, eLs1)) - c * np.kron(Pp, np.kron(Ix, e1Ls))
        else:
            M += c * np.kron(Pp, np.kron(Ix, eLs1)) - c * np.kron(Pm, np.kron(Ix, e1Ls))
    elif imass == 5:
        c = 1j * m
        if not dag:
            M += c * (np.kron(swap.T, np.kron(Ix, eLL)) + np.kron(swap, np.kron(Ix, e11)))
        else:
            M -= c * (np.kron(swap, np.kron(Ix, eLL)) + np.kron(swap.T, np.kron(Ix, e11)))
    return M


@pytest.mark.parametrize("imass", [1, 3, 5])
def test_dense_assembly(imass):
    rng = np.random.default_rng(10 + imass)
    dims = TINY
    u = D.rand_links(dims, rng)
    for dag, f in ((False, O.dslash), (True, O.dslashd)):
        Mo = D.dense(lambda v: f(v, u, dims, 0.3, imass), dims)
        Md = _dense_shamir(dims, u, 0.3, imass, dag=dag)
        assert np.abs(Mo - Md).max() < 1e-14
    # M^dagger is the conjugate transpose of M
    Mo = D.dense(lambda v: O.dslash(v, u, dims, 0.3, imass), dims)
    Md = D.dense(lambda v: O.dslashd(v, u, dims, 0.3, imass), dims)
    assert np.abs(Md - Mo.conj().T).max() < 1e-14


def _s_shift(a):
    """Cyclic shift of a haloed spinor in s: (S a)(s) = a(s + 1 mod Ls)."""
    b = np.array(a, copy=True, order="F")
    b[1:-1] = np.roll(a[1:-1], -1, axis=0)
    return refstate.halo5(b)


def test_imass1_wall_term_matches_notes():
    """The imass = 1 mass term against the reference's own definition of the
    domain-wall operator (notes/fermions-on-lattice.tex:196-199):
        D_{d+1}(s;r) = delta_sr - (1 - delta_{s,Ls-1}) P- delta_{s+1,r}
                       - (1 - delta_{s0}) P+ delta_{s-1,r}
                       + m (P- delta_{s,Ls-1} delta_{0r} + P+ delta_{s0} delta_{Ls-1,r}),
    with the domain-wall chiral projectors P+- = (1 +- gamma_4)/2 of the
    reference's gamma table (gammamatrices.F90:27-125).  Two checks:
    M(m) - M(0) is m times that wall term (built here from the table, not
    from spin ranges), and at m = -1 the wall term is exactly the missing
    wrap of the s hops, so M(-1) commutes with the cyclic shift in s.  A
    flipped sign (planted on the oracle's output) fails the second check."""
    rng = np.random.default_rng(21)
    dims = TINY
    lx, ly, lt, ls = dims
    u = D.rand_links(dims, rng)
    G = D.gamma_tables()
    Pp = 0.5 * (np.eye(4) + G[3])
    Pm = 0.5 * (np.eye(4) - G[3])
    n4 = lx * ly * lt
    w_lo = np.zeros((ls, ls)); w_lo[ls - 1, 0] = 1.0   # s = Ls-1 <- r = 0
    w_hi = np.zeros((ls, ls)); w_hi[0, ls - 1] = 1.0   # s = 0 <- r = Ls-1
    W = np.kron(Pm, np.kron(np.eye(n4), w_lo)) + np.kron(Pp, np.kron(np.eye(n4), w_hi))
    M0 = D.dense(lambda v: O.dslash(v, u, dims, 0.0, 1), dims)
    m = 0.25  # exact in float32
    Mm = D.dense(lambda v: O.dslash(v, u, dims, m, 1), dims)
    assert np.abs((Mm - M0) - m * W).max() < 1e-15
    r = D.rand_spinor(dims, rng)

    def commutator(op):
        return np.linalg.norm(D.interior_vec(op(_s_shift(r)) - _s_shift(op(r))))
    good = commutator(lambda v: O.dslash(v, u, dims, -1.0, 1))
    assert good < 1e-13 * np.linalg.norm(D.interior_vec(r))
    # the same operator with the wall term's sign flipped is antiperiodic in s
    flipped = commutator(lambda v: 2 * O.dslash(v, u, dims, 0.0, 1) - O.dslash(v, u, dims, -1.0, 1))
    assert flipped > 0.1 * np.linalg.norm(D.interior_vec(r))


def _exact_condensate(M, dims, imass):
    """The noise-free value of measure_shamir's psibarpsi for one gauge field:
    its wall-to-wall contractions (measure_module.F90:240-347) with the noise
    average taken exactly (sum over every wall site and spin of the inverse),
    finished as at measure_module.F90:368-380."""
    lx, ly, lt, ls = dims
    kvol = lx * ly * lt
    Mi = np.linalg.inv(M)

    def ix(s, x, y, t, a):  # interior vector index, Fortran order (s, x, y, t, spin)
        return s + ls * (x + lx * (y + ly * (t + lt * a)))
    p1 = p2 = 0j
    for t in range(lt):
        for y in range(ly):
            for x in range(lx):
                for ids in (0, 1):
                    if imass != 5:
                        # source spin ids on s = 1 read at s = Ls; spin ids+2 on s = Ls read at s = 1
                        p1 += Mi[ix(ls - 1, x, y, t, ids), ix(0, x, y, t, ids)]
                        p2 += Mi[ix(0, x, y, t, ids + 2), ix(ls - 1, x, y, t, ids + 2)]
                    else:
                        # xi(1, 1:2) = -eta placed from spins 3:4, read at s = 1 spin ids+2
                        # (245-257, 281-291); xi(Ls, 3:4) = -eta, read at s = Ls spin ids
                        # (302-315, 341-347)
                        p1 -= Mi[ix(0, x, y, t, ids + 2), ix(0, x, y, t, ids)]
                        p2 -= Mi[ix(ls - 1, x, y, t, ids), ix(ls - 1, x, y, t, ids + 2)]
    if imass == 1:
        return (p1 / kvol + p2 / kvol).real
    if imass == 3:
        return (-1j * p1 / kvol + 1j * p2 / kvol).real
    return (-1j * p1 / kvol - 1j * p2 / kvol).real


def test_imass5_condensate_equals_imass3():
    """imass = 3 and imass = 5 are the two twisted masses of 2+1-D domain-wall
    fermions, related by a chiral rotation: on every gauge field their
    condensates, as measure_shamir contracts them, are equal (the spectra of
    M and M^+M are equal too, but those are the same for m and -m, so they
    cannot fix a sign).  imass = 3's sign is pinned by the reference's
    psibarpsi known answer (test_measure_shamir_known_answer), so this pins
    the sign of the imass = 5 wall term of dirac.F90:158-176 together with the
    reference's imass = 5 contraction: with that term flipped (planted on the
    oracle's output) the condensate changes sign."""
    rng = np.random.default_rng(22)
    dims = TINY
    u = D.rand_links(dims, rng, kind="theta")
    m = 0.25
    M3 = D.dense(lambda v: O.dslash(v, u, dims, m, 3), dims)
    M5 = D.dense(lambda v: O.dslash(v, u, dims, m, 5), dims)
    c3, c5 = _exact_condensate(M3, dims, 3), _exact_condensate(M5, dims, 5)
    assert abs(c3) > 0.05
    assert abs(c5 - c3) < 1e-12 * abs(c3), (c3, c5)
    M50 = D.dense(lambda v: O.dslash(v, u, dims, 0.0, 5), dims)
    c5_flipped = _exact_condensate(2 * M50 - M5, dims, 5)
    assert abs(c5_flipped + c3) < 1e-12 * abs(c3)


@pytest.mark.parametrize("imass", [1, 3, 5])
def test_measure_wall_table_matches_dense_condensate(imass):
    """The wall sources and reads of measure_shamir's noise loop
    (thirring-rhmc_amd/measure.py, the table the GPU entry point
    thr_measure_shamir_noise follows) against the dense contraction of
    _exact_condensate, for one noise vector: the solved composition (oracle
    dslashd + congrad to res 1e-9) equals sum over wall sites of
    conj(eta) M^-1 xi built from the dense inverse, for every imass.
    _exact_condensate's imass 5 table is itself pinned by the imass 3 = 5
    condensate equality above."""
    from site_random import SiteRandom
    rng = np.random.default_rng(23)
    dims = TINY
    lx, ly, lt, ls = dims
    u = D.rand_links(dims, rng, kind="theta")
    m = np.float32(0.25)
    Mi = np.linalg.inv(D.dense(lambda v: O.dslash(v, u, dims, m, imass), dims))
    gen = SiteRandom(lx, ly, lt, 5.0)
    got, _ = measure.measure_shamir(u, dims, np.float32(1e-9), m, imass,
                                    SiteRandom(lx, ly, lt, 5.0), knoise=1, ops=O, niterc=1000)
    ps, pt = gen.gauss0(), gen.gauss0()
    e1 = (ps[..., 0] + 1j * ps[..., 1].astype(np.float64)).ravel(order="F")  # (x, y, t)
    e2 = (pt[..., 0] + 1j * pt[..., 1].astype(np.float64)).ravel(order="F")
    sites = np.arange(lx * ly * lt)

    def wall(s, a):  # interior indices of (s, x, y, t, spin a) over the 4-D sites
        return s + ls * (sites + lx * ly * lt * a)

    def contract(eta, read, src):  # sum_{x, x'} conj(eta(x)) M^-1[read(x), src(x')] eta(x')
        return np.conj(eta) @ Mi[np.ix_(read, src)] @ eta
    p1 = p2 = 0j
    for ids in (0, 1):
        if imass != 5:
            p1 += contract(e1, wall(ls - 1, ids), wall(0, ids))
            p2 += contract(e2, wall(0, ids + 2), wall(ls - 1, ids + 2))
        else:
            p1 -= contract(e1, wall(0, ids + 2), wall(0, ids))
            p2 -= contract(e2, wall(ls - 1, ids), wall(ls - 1, ids + 2))
    want = float(np.float32(measure._finish(p1, p2, imass, lx * ly * lt)))
    assert abs(got - want) <= 1e-6 * abs(want), (got, want)


@pytest.mark.parametrize("kernel,imass", [("shamir", 1), ("shamir", 3), ("shamir", 5),
                                          ("wilson", 1), ("wilson", 3)])
def test_adjoint(kernel, imass):
    rng = np.random.default_rng(imass)
    dims = SMALL
    u = D.rand_links(dims, rng)
    a, b = D.rand_spinor(dims, rng), D.rand_spinor(dims, rng)
    lhs = ip(a, O.dslash(b, u, dims, 0.3, imass, kernel=kernel))
    rhs = ip(O.dslashd(a, u, dims, 0.3, imass, kernel=kernel), b)
    assert abs(lhs - rhs) < 1e-13 * abs(lhs)


@pytest.mark.parametrize("kernel", ["shamir", "wilson"])
def test_gauge_covariance(kernel):
    rng = np.random.default_rng(5)
    dims = SMALL
    lx, ly, lt, ls = dims
    u = D.rand_links(dims, rng)
    b = D.rand_spinor(dims, rng)
    g = np.exp(1j * rng.uniform(0, 2 * np.pi, (lx, ly, lt)))
    up = u.copy(order="F")
    steps = [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
    for mu in range(3):
        gn = np.roll(g, tuple(-s for s in steps[mu]), axis=(0, 1, 2))
        up[1:-1, 1:-1, 1:-1, mu] = g * u[1:-1, 1:-1, 1:-1, mu] * np.conj(gn)
    refstate.halo4(up)
    gb = b.copy(order="F")
    gb[1:-1, 1:-1, 1:-1, 1:-1, :] *= g[None, :, :, :, None]
    refstate.halo5(gb)
    for f in (O.dslash, O.dslashd):
        lhs = O.interior(f(gb, up, dims, 0.3, 3, kernel=kernel))
        rhs = O.interior(f(b, u, dims, 0.3, 3, kernel=kernel)) * g[None, :, :, :, None]
        assert np.abs(lhs - rhs).max() < 1e-13


def test_wilson_kernel_identity():
    """dirac.F90:226-299 == Shamir with every s-coupling N replaced by (1 - D_W) N.

    D_W is taken from the Shamir operator on one s-slice (Ls = 1, am = 0).  The
    reference's imass=3 wall term carries the opposite sign in the Wilson
    kernel, so it matches Shamir at -am."""
    rng = np.random.default_rng(7)
    dims = SMALL
    lx, ly, lt, ls = dims
    u = D.rand_links(dims, rng)
    b = D.rand_spinor(dims, rng)

    def DW(v):
        out = np.zeros_like(v)
        d1 = (lx, ly, lt, 1)
        for s in range(1, ls + 1):
            sl = np.zeros((3, lx + 2, ly + 2, lt + 2, 4), complex, order="F")
            sl[1] = v[s]
            refstate.halo5(sl)
            out[s] = O.dslash(sl, u, d1, 0.0, 1)[1] - v[s]
        return refstate.halo5(out)

    for imass, sgn in ((1, 1.0), (3, -1.0)):
        dwb = DW(b)
        N = refstate.halo5(O.dslash(b, u, dims, sgn * 0.3, imass) - dwb - b)
        rhs = dwb + b + N - DW(N)
        lhs = O.dslash(b, u, dims, 0.3, imass, kernel="wilson")
        assert np.abs(O.interior(lhs) - O.interior(rhs)).max() < 1e-13


@pytest.mark.parametrize("kernel", ["shamir", "wilson"])
@pytest.mark.parametrize("iflag", [0, 1])
def test_derivs_exact_finite_difference(kernel, iflag):
    """derivs(R,X2,a,0) = fl32(2a) d/du_imag Re<R, M^+ X2>; iflag 1 uses M."""
    rng = np.random.default_rng(20 + iflag)
    dims = (2, 3, 2, 4)
    lx, ly, lt, ls = dims
    u = D.rand_links(dims, rng)
    R, X2 = D.rand_spinor(dims, rng), D.rand_spinor(dims, rng)
    anum = 0.37
    ds = np.full((lx, ly, lt, 3), 0.25, order="F")  # accumulates on top
    O.derivs(R, X2, anum, iflag, dims, ds, kernel=kernel, am=0.3, imass=3)
    c = float(np.float32(2 * np.float32(anum)))
    f = O.dslashd if iflag == 0 else O.dslash
    base = ip(R, f(X2, u, dims, 0.3, 3, kernel=kernel)).real
    fd = np.zeros_like(ds)
    for mu in range(3):
        for x in range(lx):
            for y in range(ly):
                for t in range(lt):
                    up = u.copy(order="F")
                    up[x + 1, y + 1, t + 1, mu] += 1j
                    refstate.halo4(up)
                    fd[x, y, t, mu] = c * (ip(R, f(X2, up, dims, 0.3, 3, kernel=kernel)).real - base)
    assert np.abs(ds - 0.25 - fd).max() < 1e-12 * np.abs(fd).max()


def test_coef_closed_form():
    rng = np.random.default_rng(3)
    dims = (3, 4, 5, 4)
    lx, ly, lt, ls = dims
    th = rng.standard_normal((lx, ly, lt, 3)).astype(np.float32)
    u = O.coef(th, dims, ibound=-1)
    ref = (1.0 + 1j * th.astype(np.float64))
    ref[:, :, lt - 1, 2] *= -1
    assert np.array_equal(u[1:-1, 1:-1, 1:-1, :], ref)
    assert np.array_equal(u[0, 1:-1, 1:-1, :], ref[lx - 1])
    assert np.array_equal(u[1:-1, 1:-1, 0, :], ref[:, :, lt - 1])
    u2 = O.coef(th, dims, ibound=1)
    assert np.array_equal(u2[1:-1, 1:-1, lt, 2], 1.0 + 1j * th[:, :, lt - 1, 2].astype(np.float64))


def _dense_MdM(dims, u, am, imass, kernel="shamir"):
    M = D.dense(lambda v: O.dslash(v, u, dims, am, imass, kernel=kernel), dims)
    return M.conj().T @ M


@pytest.mark.parametrize("kernel", ["shamir", "wilson"])
def test_multishift_iterates_equal_shifted_cg(kernel):
    """After k iterations x_i equals the k-th CG iterate of (A + sigma_i) x = b."""
    rng = np.random.default_rng(30)
    dims = TINY
    u = D.rand_links(dims, rng, kind="theta")
    b = D.rand_spinor(dims, rng)
    A = _dense_MdM(dims, u, 0.2, 3, kernel)
    bv = D.interior_vec(b)
    sig = np.array([0.05, 0.3, 1.7, 6.0])
    anum = np.array([0.1, 0.2, 0.3, 0.4])
    for k in (1, 3, 6):
        out, cgr, _ = O.multishift(u, dims, 0.2, 3, sig, anum, b, res=1e-30, maxcg=k, kernel=kernel)
        assert cgr == k
        for i, s in enumerate(sig):
            As = A + s * np.eye(A.shape[0])
            x = np.zeros_like(bv)
            r = bv.copy()
            p = r.copy()
            for _ in range(k):
                Ap = As @ p
                al = np.vdot(r, r).real / np.vdot(p, Ap).real
                x = x + al * p
                rn = r - al * Ap
                p = rn + (np.vdot(rn, rn).real / np.vdot(r, r).real) * p
                r = rn
            xi = out[..., i].reshape(-1, order="F")
            assert np.linalg.norm(xi - x) < 1e-10 * np.linalg.norm(x)


def test_multishift_converged_and_sp():
    rng = np.random.default_rng(31)
    dims = TINY
    u = D.rand_links(dims, rng, kind="theta")
    b = D.rand_spinor(dims, rng)
    A = _dense_MdM(dims, u, 0.2, 3)
    bv = D.interior_vec(b)
    sig = np.array([0.01, 0.3, 2.0])
    anum = np.array([0.5, 0.5, 0.5])
    out, cgr, cgrs = O.multishift(u, dims, 0.2, 3, sig, anum, b, res=1e-9, maxcg=5000)
    for i, s in enumerate(sig):
        x = np.linalg.solve(A + s * np.eye(A.shape[0]), bv)
        xi = out[..., i].reshape(-1, order="F")
        assert np.linalg.norm(xi - x) < 1e-7 * np.linalg.norm(x)
    # a large shift converges first and stops being updated
    assert cgrs[2] < cgrs[0] == cgr
    outs, cgrsp, _ = O.multishift(u, dims, 0.2, 3, sig, anum, b, res=1e-5, maxcg=5000, sp=True)
    for i in range(3):
        x = np.linalg.solve(A + sig[i] * np.eye(A.shape[0]), bv)
        xi = outs[..., i].reshape(-1, order="F")
        assert np.linalg.norm(xi - x) < 1e-4 * np.linalg.norm(x)


def _remez(path, block):
    with open(path) as fh:
        rows = [l.replace(",", " ").split() for l in fh if l.strip()]
    n = (len(rows) - 2) // 2
    blk = rows[block * (n + 1):(block + 1) * (n + 1)]
    a0 = float(blk[0][0])
    a = np.array([float(r[0]) for r in blk[1:]])
    s = np.array([float(r[1]) for r in blk[1:]])
    return a0, a, s


def test_remez4_block2_accuracy():
    import os
    path = os.path.join(os.path.dirname(__file__), "..", "thirring-rhmc_amd", "data", "remez4")
    a0, a, s = _remez(path, 1)
    assert len(a) == 25 and np.all(a > 0) and np.all(s > 0)
    x = np.logspace(-4, np.log10(64.0), 4001)
    r = a0 + np.sum(a[None, :] / (x[:, None] + s[None, :]), axis=1)
    assert np.max(np.abs(r * x ** 0.25 - 1.0)) < 5e-10


@pytest.mark.parametrize("iflag", [0, 1])
def test_qmrherm_rational_function(iflag):
    rng = np.random.default_rng(40)
    dims = TINY
    u = D.rand_links(dims, rng, kind="theta")
    phi = D.rand_spinor(dims, rng)
    A = _dense_MdM(dims, u, 0.2, 3)
    sig = np.array([0.02, 0.4, 3.0])
    anum = np.array([0.7, 0.25, 0.5, 1.5])
    x, it, x1, phi0, _, _ = O.qmrherm(phi, 1e-10, 0.2, 3, anum, sig, iflag, u, dims)
    pv = D.interior_vec(phi)
    ref = anum[0] * pv
    for i, s in enumerate(sig):
        xi = np.linalg.solve(A + s * np.eye(A.shape[0]), pv)
        ref = ref + anum[i + 1] * xi
        if iflag == 1:
            assert np.linalg.norm(D.interior_vec(phi0[..., i]) - xi) < 1e-8 * np.linalg.norm(xi)
    assert np.linalg.norm(D.interior_vec(x) - ref) < 1e-8 * np.linalg.norm(ref)
    # x comes back with periodic halos
    assert np.array_equal(x[0], x[-2]) and np.array_equal(x[:, :, :, 0], x[:, :, :, -2])


def _theta_links(th, dims):
    return O.coef(th, dims, ibound=1)


def test_qmrherm_forces_vs_action():
    """iflag 2 == -dS/dtheta and iflag 3 == 2 dF/dtheta by central differences.

    S(theta) = Re sum_i a_i <X, (A+s_i)^-1 X>, F = Re sum_i a_i <Phi, (A(theta)+s_i)^-1 X>,
    A = M^+ M with u = 1 + i theta (periodic, so du/dtheta = i on every link)."""
    rng = np.random.default_rng(41)
    dims = TINY
    lx, ly, lt, ls = dims
    th = (0.4 * rng.standard_normal((lx, ly, lt, 3))).astype(np.float32)
    u = _theta_links(th, dims)
    X = D.rand_spinor(dims, rng)
    Phi = D.rand_spinor(dims, rng)
    sig = np.array([0.05, 0.6])
    anum = np.array([0.0, 0.375, 0.5])  # exact in float32 after doubling
    Xv, Pv = D.interior_vec(X), D.interior_vec(Phi)

    def S_F(thd):
        uu = O.coef(thd.astype(np.float32), dims, ibound=1) if thd.dtype == np.float32 else None
        return uu

    def dense_A(thd):
        uu = np.zeros((lx + 2, ly + 2, lt + 2, 3), complex, order="F")
        uu[1:-1, 1:-1, 1:-1, :] = 1.0 + 1j * thd
        refstate.halo4(uu)
        return _dense_MdM(dims, uu, 0.2, 3)

    # iflag 1 to fill Phi0 from Phi (as force_module.F90:80 does), then iflag 2 and 3 on X
    _, _, x1, phi0, _, _ = O.qmrherm(Phi, 1e-11, 0.2, 3, anum, sig, 1, u, dims)
    _, _, _, _, ds2, _ = O.qmrherm(X, 1e-11, 0.2, 3, anum, sig, 2, u, dims)
    _, _, _, _, ds3, _ = O.qmrherm(X, 1e-11, 0.2, 3, anum, sig, 3, u, dims, x1=x1, phi0=phi0)
    h = 1e-5
    thd = th.astype(np.float64)
    for (x, y, t, mu) in [(0, 0, 0, 0), (1, 0, 1, 1), (0, 1, 1, 2), (1, 1, 0, 2)]:
        vals = []
        for sgn in (1, -1):
            tp = thd.copy()
            tp[x, y, t, mu] += sgn * h
            A = dense_A(tp)
            S = F = 0.0
            for i, s in enumerate(sig):
                As = A + s * np.eye(A.shape[0])
                S += anum[i + 1] * np.vdot(Xv, np.linalg.solve(As, Xv)).real
                F += anum[i + 1] * np.vdot(Pv, np.linalg.solve(As, Xv)).real
            vals.append((S, F))
        dS = (vals[0][0] - vals[1][0]) / (2 * h)
        dF = (vals[0][1] - vals[1][1]) / (2 * h)
        assert abs(ds2[x, y, t, mu] + dS) < 1e-6 * max(1.0, abs(dS))
        assert abs(ds3[x, y, t, mu] - 2 * dF) < 1e-6 * max(1.0, abs(dF))


def test_congrad_wilson_true_residual():
    rng = np.random.default_rng(50)
    dims = (4, 4, 4, 4)
    u = D.rand_links(dims, rng, kind="theta")
    phi = D.rand_spinor(dims, rng)
    x = np.zeros_like(phi)
    it, _ = O.congrad(phi, x, u, dims, res=1e-6, am=0.1, imass=1, kernel="wilson")
    r = D.interior_vec(phi) - D.interior_vec(O.dirac_op_shifted(x, u, dims, 0.1, 1, kernel="wilson"))
    assert np.sum(np.abs(r) ** 2) < 4 * 4 ** 4 * np.float32(1e-6) ** 2


def _remez_data(name, block):
    import os
    return _remez(os.path.join(os.path.dirname(__file__), "..", "thirring-rhmc_amd", "data", name),
                  block)


@pytest.mark.parametrize("kernel", ["shamir", "wilson"])
def test_force_is_derivative_of_hamilton(kernel):
    """force_module.F90 against dwf3d_lib.F90 hamilton: with the same rational
    approximations in both, dSdpi - beta*theta is the theta-derivative of the
    pseudofermion action hf = <R, r2(K) R>, R = r4(K_PV) Phi (central
    differences; u = 1 + i theta is linear in theta, antiperiodic in t, so the
    t-boundary links enter with the opposite sign).  Pins the composition of
    the four qmrherm calls, the sign flip at the t boundary and the gauge term."""
    a0_4, a_4, s_4 = _remez_data("remez4", 0)   # x^(1/4)
    b0_2, b_2, t_2 = _remez_data("remez2", 1)   # x^(-1/2)
    a0_2, a_2, s_2 = _remez_data("remez2", 0)   # x^(1/2)
    x = np.logspace(-3, 1, 200)
    r = lambda c0, c, s: c0 + np.sum(c[None, :] / (x[:, None] + s[None, :]), axis=1)
    assert np.max(np.abs(r(a0_4, a_4, s_4) / x ** 0.25 - 1)) < 1e-6
    assert np.max(np.abs(r(b0_2, b_2, t_2) * x ** 0.5 - 1)) < 1e-6
    assert np.max(np.abs(r(a0_2, a_2, s_2) / x ** 0.5 - 1)) < 1e-6
    anum4, aden4 = np.concatenate([[a0_4], a_4]), s_4
    bnum2, bden2 = np.concatenate([[b0_2], b_2]), t_2
    anum2, aden2 = np.concatenate([[a0_2], a_2]), s_2

    rng = np.random.default_rng(60)
    dims = (4, 4, 4, 4)
    lx, ly, lt, ls = dims
    am, imass, beta = 0.2, 3, 0.6
    th = (0.3 * rng.standard_normal((lx, ly, lt, 3))).astype(np.float32)
    phi = D.rand_spinor(dims, rng)
    pp = np.zeros((lx, ly, lt, 3), dtype=np.float32)
    ham = (anum4, aden4, bnum2, bden2)
    frc = (anum4, aden4, bnum2, bden2, anum2, aden2)
    res = np.float32(1e-11)

    def hf(t32):
        h, _ = O.hamilton(phi, res, am, imass, beta, t32, pp, O.coef(t32, dims), dims, ham,
                          kernel=kernel)
        return h["hf"]

    f = O.force(phi, res, am, imass, beta, th, O.coef(th, dims), dims, frc, kernel=kernel)
    gauge = (np.float32(beta) * th).astype(np.float64)  # d/dtheta of hg = beta/2 sum theta^2
    for (xx, yy, tt, mu) in [(0, 0, 0, 0), (1, 2, 1, 1), (2, 1, 2, 2), (3, 3, lt - 1, 2),
                             (0, 3, lt - 1, 0)]:
        tp, tm = th.copy(), th.copy()
        tp[xx, yy, tt, mu] += np.float32(2e-3)
        tm[xx, yy, tt, mu] -= np.float32(2e-3)
        step = float(tp[xx, yy, tt, mu]) - float(tm[xx, yy, tt, mu])
        dS = (hf(tp) - hf(tm)) / step
        got = f[xx, yy, tt, mu] - gauge[xx, yy, tt, mu]
        assert abs(got - dS) < 1e-6 * max(1.0, abs(dS)), ((xx, yy, tt, mu), got, dS)


def test_site_random_gauss_known_answers():
    """tests/test_gauss0.F90:83-86 and tests/test_gaussp.F90:83-86 (SITE_RANDOM
    build, seed 1.0, 12^3 sites): sum, max and min of one noise field."""
    from site_random import SiteRandom
    ps = SiteRandom(12, 12, 12, 1.0).gauss0()
    for got, want in ((ps.sum(dtype=np.float32), 14.460), (ps.max(), 2.59536),
                      (ps.min(), -2.441575)):
        assert abs(got - want) <= 1e-3 * abs(want), (got, want)
    pp = SiteRandom(12, 12, 12, 1.0).gaussp()
    for got, want in ((pp.sum(dtype=np.float32), 20.4506), (pp.max(), 3.67040),
                      (pp.min(), -3.452905)):
        assert abs(got - want) <= 1e-3 * abs(want), (got, want)


def test_measure_shamir_known_answer():
    """tests/test_measure_shamir.F90:24-46 (12^3 x Ls 24, closed-form links,
    SITE_RANDOM noise seeded 4139764973254, res 0.1, am 0.05, imass 3):
    psibarpsi == 2.504295e-04 and 5 CG iterations per solve, to 0.1% (the
    oracle gives 2.5042944e-04).  Flipping the sign of the wall-mass term flips
    the sign of psibarpsi, so this pins that sign."""
    from site_random import SiteRandom
    dims = (12, 12, 12, 24)
    st = refstate.starting_state(dims, want=("u",))
    # `seed = 4139764973254.0` is a default-real literal: the seed is its float32 value
    rng = SiteRandom(12, 12, 12, float(np.float32(4139764973254.0)))
    pbp, aviter = measure.measure_shamir(st["u"], dims, np.float32(0.1), np.float32(0.05), 3, rng,
                                         ops=O)
    assert abs(pbp - 2.504295e-04) <= 1e-3 * 2.504295e-04, pbp
    assert abs(aviter - 5.0) <= 1e-3 * 5.0, aviter


def test_measure_wilson_known_answer():
    """tests/test_measure_wilson.F90:24-47 (same inputs, Wilson-kernel wall
    sources, DWilson at the projected walls, congrad with the Shamir
    operator the reference's test values were generated with): aviter == 10.9
    and |psibarpsi| == 5.4226685e-02 to 0.1% (the oracle gives 5.4226685e-02).
    The reference's value is negative: it predates the swap of the imass = 3
    phases in measure_wilson (measure_module.F90:612-618, "The signs here are
    swapped compared to measure_shamir, fixed by SJH 7/6/24"): the phases
    before that swap negate every noise term, giving -5.4226685e-02.  Flipping
    the wall-mass sign of the Wilson kernel instead gives -8.05e-02, so the
    value pins that sign."""
    from site_random import SiteRandom
    dims = (12, 12, 12, 24)
    st = refstate.starting_state(dims, want=("u",))
    seed = float(np.float32(4139764973254.0))

    def dslashd_w(r, u, d, am, imass):
        return O.dslashd(r, u, d, am, imass, kernel="wilson")

    pbp, aviter = measure_ref.measure_wilson(dslashd_w, O.congrad, O.dwilson_slice, st["u"], dims,
                                             np.float32(0.1), np.float32(0.05), 3,
                                             SiteRandom(12, 12, 12, seed))
    assert abs(pbp - 5.4226685e-02) <= 1e-3 * 5.4226685e-02, pbp
    assert abs(aviter - 10.9) <= 1e-3 * 10.9, aviter


def _md_setup(seed=61):
    """Small lattice, rational approximations shared by force and hamilton (so
    H is the action whose gradient the force is)."""
    a0_4, a_4, s_4 = _remez_data("remez4", 0)
    b0_2, b_2, t_2 = _remez_data("remez2", 1)
    a0_2, a_2, s_2 = _remez_data("remez2", 0)
    anum4, aden4 = np.concatenate([[a0_4], a_4]), s_4
    bnum2, bden2 = np.concatenate([[b0_2], b_2]), t_2
    anum2, aden2 = np.concatenate([[a0_2], a_2]), s_2
    rng = np.random.default_rng(seed)
    dims = (4, 4, 4, 4)
    lx, ly, lt, ls = dims
    th = np.asfortranarray((0.3 * rng.standard_normal((lx, ly, lt, 3))).astype(np.float32))
    pp = np.asfortranarray(rng.standard_normal((lx, ly, lt, 3)).astype(np.float32))
    phi = D.rand_spinor(dims, rng)
    return dict(dims=dims, th=th, pp=pp, phi=phi, am=0.2, imass=3, beta=0.6,
                ham=(anum4, aden4, bnum2, bden2),
                frc=(anum4, aden4, bnum2, bden2, anum2, aden2), res=np.float32(1e-11))


def _stop_after(n, iterl):
    """ytest sequence that ends the trajectory after MD step n (evolution.F90:159-168)."""
    seq = iter([np.float32(1.0)] * (n - 1) + [np.float32(0.5) / np.float32(iterl)])
    return lambda: next(seq)


def test_leapfrog_is_reversible():
    """evolution.F90:113-179: half step, full steps, closing half step.  With
    the momenta negated the integrator retraces its path (a misplaced half
    step or a force taken at the wrong theta breaks this)."""
    m = _md_setup()
    dims, iterl, dt, n = m["dims"], 4, np.float32(0.1), 5
    th1, pp1, _, _, k = O.evolve_theta_pp(iterl, dt, m["pp"], m["phi"], m["res"], m["am"],
                                          m["imass"], m["beta"], m["th"], dims, m["frc"],
                                          _stop_after(n, iterl))
    assert k == n
    assert np.max(np.abs(th1 - m["th"])) > 0.05  # it moved
    th2, pp2, _, _, k2 = O.evolve_theta_pp(iterl, dt, -pp1, m["phi"], m["res"], m["am"],
                                           m["imass"], m["beta"], th1, dims, m["frc"],
                                           _stop_after(n, iterl))
    assert k2 == n
    assert np.max(np.abs(th2 - m["th"])) < 2e-5
    assert np.max(np.abs(pp2 + m["pp"])) < 2e-5


def test_leapfrog_energy_error_is_second_order():
    """dH over a fixed trajectory length falls as dt^2 for the leapfrog
    (evolution.F90:113-179 with hamilton, dwf3d_lib.F90:511-616): halving dt
    divides dH by ~4.  A wrong force sign or factor leaves dH of order 1."""
    m = _md_setup()
    dims = m["dims"]

    def dH(dt, n):
        h0, _ = O.hamilton(m["phi"], m["res"], m["am"], m["imass"], m["beta"], m["th"], m["pp"],
                           O.coef(m["th"], dims), dims, m["ham"])
        th, pp, u, _, k = O.evolve_theta_pp(n, np.float32(dt), m["pp"], m["phi"], m["res"],
                                            m["am"], m["imass"], m["beta"], m["th"], dims,
                                            m["frc"], _stop_after(n, n))
        assert k == n
        h1, _ = O.hamilton(m["phi"], m["res"], m["am"], m["imass"], m["beta"], th, pp, u, dims,
                           m["ham"])
        return h1["h"] - h0["h"]

    d1, d2 = dH(0.08, 5), dH(0.04, 10)
    assert abs(d1) < 0.5
    assert 3.0 < d1 / d2 < 5.0, (d1, d2)


def test_site_random_single_site_stream():
    """rano(yran, idum, 1, 1, 1) advances only the generator of site (1,1,1)
    (site_random.f90:33-60): drawing from it leaves every other site's next
    value unchanged, and that site's values follow its own stream."""
    from site_random import SiteRandom
    a, b = SiteRandom(4, 4, 4, 3.0), SiteRandom(4, 4, 4, 3.0)
    seq = [a.rano_at(0, 0, 0) for _ in range(5)]
    full = [b.rano()[0, 0, 0] for _ in range(5)]
    assert seq == full
    na, nb = a.rano(), b.rano()
    assert na[0, 0, 0] == nb[0, 0, 0]
    na[0, 0, 0] = nb[0, 0, 0] = 0
    # b advanced all sites 5 times; a only site (0,0,0): the others differ
    assert not np.array_equal(na, nb)
    c = SiteRandom(4, 4, 4, 3.0)
    nc = c.rano()
    nc[0, 0, 0] = 0
    assert np.array_equal(na, nc)


def test_trajectory_composition_on_the_oracle():
    """thirring-rhmc_amd/md.py (dwf3d_lib.F90:263-318) run over the oracle on
    the reference's end-to-end sample (con_4x4x4, midout, Ls = 8): the sweep
    is reproducible from the seed (to the oracle's OpenMP summation order),
    conserves energy to |dH| < 0.5, and accepts whenever dH >= 0."""
    import os
    import confio
    import md
    from md_oracle_ops import OracleOps
    from site_random import SiteRandom
    gold = os.path.join(os.path.dirname(__file__), "golden")
    th, seed, _ = confio.read_con(os.path.join(gold, "con_4x4x4_serial"), 4, 4, 4)
    dims = (4, 4, 4, 8)
    runs = []
    for _ in range(2):
        rng = SiteRandom(4, 4, 4, seed)
        r = md.trajectory(th, rng, dims, 0.01, 3, 0.4, 7, 0.14, np.float32(1e-8),
                          np.float32(1e-6), ops=OracleOps)
        runs.append(r)
    a, b = runs
    assert a["steps"] == b["steps"] and a["accepted"] == b["accepted"]
    assert abs(a["dH"] - b["dH"]) < 1e-6
    assert np.max(np.abs(a["theta"] - b["theta"])) < 1e-5
    assert 1 <= a["steps"] <= 28 and abs(a["dH"]) < 0.5
    if a["dH"] >= 0:
        assert a["accepted"]
    assert a["H0"]["hg"] > 0 and a["H0"]["hp"] > 0
