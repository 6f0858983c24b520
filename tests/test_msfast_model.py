"""CPU model of the batched multishift's blocked shift updates (no GPU).

The device solver never updates x_i and p_i per iteration.  Each keeps the
block-start values X_i, P_i and FP64 coefficients, p_i = ca P_i + sum_q cb_q R_q
and x_i = X_i + cc P_i + sum_q cd_q R_q over the block's residuals R_q, which
the per-iteration recurrences of multishift_module.F90:118-176 (x_i -= omega_i
p_i; p_i = gamma_i p_i + zeta_i''' r) update.  A flush applies them once per
block, in one pass or in two passes over halves of the ring (the FP64 ring
of a 64-iteration block).  This model checks that algebra against the
per-iteration recurrence on a small SPD system, at block lengths 32 and 64,
with the split flush, and with a last block that is only partly filled.
"""
import numpy as np
import pytest


def _system(n=48, seed=3):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.geomspace(1e-2, 4.0, n)
    a = (q * lam) @ q.T
    b = rng.standard_normal(n)
    sig = np.array([0.0, 1e-3, 1e-2, 0.1, 1.0])
    return a, b, sig


def _scalars(a, b, sig, iters, rtol=0.0):
    """Base CG and the shifted zeta/omega/gamma recurrences (the finalize
    kernels' formulas), with the base residuals r_k after each iteration;
    stops early once |r|/|b| <= rtol."""
    b2 = b @ b
    r = b.copy()
    p = b.copy()
    delta = r @ r
    omega, gam = 1.0, 0.0
    z_i = np.ones_like(sig)
    z_ii = np.ones_like(sig)
    steps = []
    for _ in range(iters):
        ap = a @ p
        alpha = p @ ap
        os_, om = omega, -delta / alpha
        z_iii = (z_i * z_ii * os_) / (om * gam * (z_i - z_ii) + z_i * os_ * (1.0 - sig * om))
        om_i = om * z_iii / z_ii
        r = r + om * ap
        lam = r @ r
        gam_new = lam / delta
        g_i = gam_new * z_iii * om_i / (z_ii * om)
        steps.append((om_i.copy(), g_i.copy(), z_iii.copy(), r.copy()))
        z_i, z_ii = z_ii, z_iii
        p = r + gam_new * p
        delta, omega, gam = lam, om, gam_new
        if lam <= rtol * rtol * b2:
            break
    return steps


def _per_iteration(b, sig, steps):
    x = np.zeros((len(sig), len(b)))
    p = np.tile(b, (len(sig), 1))
    for om_i, g_i, z_iii, r in steps:
        x -= om_i[:, None] * p
        p = g_i[:, None] * p + z_iii[:, None] * r
    return x


def _blocked(b, sig, steps, msb, split=False):
    ns, n = len(sig), len(b)
    X = np.zeros((ns, n))
    P = np.tile(b, (ns, 1))
    for t0 in range(0, len(steps), msb):
        blk = steps[t0:t0 + msb]
        jc = len(blk)  # a last block may be partly filled
        ca, cc = np.ones(ns), np.zeros(ns)
        cb, cd = np.zeros((ns, msb)), np.zeros((ns, msb))
        ring = np.zeros((msb, n))
        for j, (om_i, g_i, z_iii, r) in enumerate(blk):
            # x_i -= omega_i p_i ; p_i = gamma_i p_i + zeta_i''' R_j
            cc -= om_i * ca
            cd[:, :j] -= om_i[:, None] * cb[:, :j]
            ca *= g_i
            cb[:, :j] *= g_i[:, None]
            cb[:, j] = z_iii
            cd[:, j] = 0.0
            ring[j] = r
        halves = [(0, msb // 2), (msb // 2, msb)] if split else [(0, msb)]
        for h, (q0, q1) in enumerate(halves):
            if q0 >= jc:
                break
            q1 = min(q1, jc)
            if h == 0:
                base = P.copy()
                X = X + cc[:, None] * base
                P = ca[:, None] * base
            X = X + cd[:, q0:q1] @ ring[q0:q1]
            P = P + cb[:, q0:q1] @ ring[q0:q1]
    return X


@pytest.mark.parametrize("msb,split", [(32, False), (64, False), (64, True)])
@pytest.mark.parametrize("iters", [64, 96, 150])
def test_blocked_updates_equal_per_iteration(msb, split, iters):
    a, b, sig = _system()
    steps = _scalars(a, b, sig, iters, rtol=1e-13)
    assert len(steps) > 32
    ref = _per_iteration(b, sig, steps)
    got = _blocked(b, sig, steps, msb, split)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err <= 1e-12, err


def test_model_solves_the_shifted_systems():
    """The recurrences themselves: after enough iterations every x_i solves
    (A + sigma_i) x_i = b (so the model above checks a working solver)."""
    a, b, sig = _system()
    steps = _scalars(a, b, sig, 400, rtol=1e-13)
    assert len(steps) > 64  # more than one 64-iteration block
    x = _blocked(b, sig, steps, 64, split=True)
    for i, s in enumerate(sig):
        res = np.linalg.norm(b - (a + s * np.eye(len(b))) @ x[i]) / np.linalg.norm(b)
        assert res <= 1e-10, (s, res)


# ---------------------------------------------------------------------------
# reliable updates with low-precision Krylov storage
# ---------------------------------------------------------------------------
def _f32(v):
    return v.astype(np.float32).astype(np.float64)


def _lowp_solve(m, a, b, sig, reliable, fix_directions, tol=1e-10, maxit=20000, blk=32,
                rdelta=0.1, th32=1.0, return_x=False):
    """The h-recurrence multishift (h = M p, r_new = r + omega M^T h) with h
    and r rounded to FP32 every iteration once the storage switched (from
    the start when th32 >= 1, else at the end of the first block with
    |r|/|b| < th32), shift vectors in FP64.  With
    `reliable`, at the end of a block where the smallest active shift's
    iterated residual fell by rdelta, the base residual becomes that shift's
    true residual over its zeta, and with `fix_directions` every active p_j
    gets zeta_j (r' - r) added.  Returns the FP64 true residual per shift."""
    n, ns = len(b), len(sig)
    x = np.zeros((ns, n))
    p = np.tile(b, (ns, 1))
    r, h = b.copy(), None
    delta = b2 = r @ r
    z_i, z_ii = np.ones(ns), np.ones(ns)
    omega, gam = 1.0, 0.0
    active = np.ones(ns, bool)
    r_ref = 1.0
    low = th32 >= 1.0

    def rnd(v):
        return _f32(v) if low else v

    for k in range(maxit):
        mr = m @ r
        h = rnd(mr if k == 0 else mr + gam * h)
        alpha = h @ h
        os_, om = omega, -delta / alpha
        z_iii = np.where(active, z_i * z_ii * os_ / (om * gam * (z_i - z_ii)
                                                     + z_i * os_ * (1 - sig * om)), z_ii)
        om_i = om * z_iii / z_ii
        rn = rnd(r + om * (m @ h))
        lam = rn @ rn
        g = lam / delta
        g_i = g * z_iii * om_i / (z_ii * om)
        x[active] -= om_i[active, None] * p[active]
        p[active] = g_i[active, None] * p[active] + z_iii[active, None] * rn
        conv = z_iii ** 2 * lam <= (0.75 * tol) ** 2 * b2
        z_i = np.where(active, z_ii, z_i)
        z_ii = np.where(active, z_iii, z_ii)
        active &= ~conv
        omega, gam, delta, r = om, g, lam, rn
        if not active.any():
            break
        if not low and (k + 1) % blk == 0 and np.sqrt(lam / b2) < th32:
            low = True
            h, r = _f32(h), _f32(r)
            continue
        if low and reliable and (k + 1) % blk == 0:
            i = int(np.argmin(np.where(active, sig, np.inf)))
            worst = abs(z_ii[i]) * np.sqrt(lam / b2)
            if worst < rdelta * r_ref:
                rnew = _f32((b - (a + sig[i] * np.eye(n)) @ x[i]) / z_ii[i])
                if fix_directions:
                    p[active] += z_ii[active, None] * (rnew - r)
                r, delta, r_ref = rnew, rnew @ rnew, worst
    tr = np.array([np.linalg.norm(b - (a + s * np.eye(n)) @ x[i]) / np.sqrt(b2)
                   for i, s in enumerate(sig)])
    return (tr, x) if return_x else tr


def _spectrum_system(n=300, seed=1):
    rng = np.random.default_rng(seed)
    ev = np.exp(rng.uniform(np.log(1e-4), np.log(8.0), n))
    ev[0] = 1e-4
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    m = (q * np.sqrt(ev)) @ q.T
    return m, m @ m, rng.standard_normal(n), np.array([1e-5, 1e-4, 1e-3, 1e-2, 1e-1, 1.0])


def test_drift_is_set_where_the_precision_drops():
    """Without reliable updates, a shift's residual gap is the sum of its
    rounding errors weighted by zeta, so it is set by |r| when the storage
    drops to FP32: switching two decades later shrinks the smallest shift's
    gap by more than 10x each time, and shifts that converged before the
    switch keep FP64 accuracy (the reason the mixed solves run FP64 down to
    |r|/|b| = 1e-2 and correct the rest)."""
    m, a, b, sig = _spectrum_system()
    gaps = {th: _lowp_solve(m, a, b, sig, False, False, th32=th) for th in (1.0, 1e-2, 1e-4)}
    assert gaps[1.0][0] > 10 * gaps[1e-2][0] > 100 * gaps[1e-4][0]
    assert gaps[1e-2][2:].max() <= 1e-10  # sigma >= 1e-3 converged in FP64


def test_reliable_updates_model():
    """FP32 Krylov storage from the first iteration, 6 shifts on a spectrum
    [1e-4, 8] (condition 8e4): without reliable updates the smallest shift
    ends with a drift of ~6e-5; replacing the base residual without moving
    the shift directions breaks the shifted recurrences (~0.3); with the
    directions moved the smallest shift meets 1e-10, and the heavier ones keep
    the drift they took while converging in FP32 (for the FP64 correction)."""
    m, a, b, sig = _spectrum_system()
    plain = _lowp_solve(m, a, b, sig, reliable=False, fix_directions=False)
    broken = _lowp_solve(m, a, b, sig, reliable=True, fix_directions=False)
    fixed = _lowp_solve(m, a, b, sig, reliable=True, fix_directions=True)
    assert plain[0] > 1e-6
    assert broken[0] > 1e-2
    assert fixed[0] <= 1e-10
    assert fixed[1:].max() > 1e-8  # the heavier shifts' drift stays


def test_defect_correction_model():
    """The FP64 defect correction after a drifted low-precision solve (FP32
    storage from the start): for every shift above tol, solve
    (A + sigma_i) c = r_i with r_i = b - (A + sigma_i) x_i, itself in FP32
    storage, to tol/2 relative to |b| (tol |b| / (2 |r_i|) relative to
    |r_i|), and add c.  One round brings every shift to FP64 true residual
    <= tol: the drift of the correction solve is relative to |r_i|, not
    |b|."""
    m, a, b, sig = _spectrum_system()
    tol = 1e-10
    tr, x = _lowp_solve(m, a, b, sig, True, True, return_x=True)
    assert tr.max() > 1e-6  # the heavier shifts drifted
    bn = np.linalg.norm(b)
    for i, s in enumerate(sig):
        if tr[i] <= tol:
            continue
        ri = b - (a + s * np.eye(len(b))) @ x[i]
        tl = 0.5 * tol * bn / np.linalg.norm(ri)
        _, c = _lowp_solve(m, a, ri, sig[i:i + 1], True, True, tol=tl, return_x=True)
        x[i] += c[0]
    final = np.array([np.linalg.norm(b - (a + s * np.eye(len(b))) @ x[i]) / bn
                      for i, s in enumerate(sig)])
    assert final.max() <= tol, final
