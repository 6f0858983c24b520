"""One RHMC trajectory with every fermion solve and the leapfrog on the GPU.

Follows the reference's sweep (dwf3d_lib.F90:263-318) for Nf = 1 with the
rational (non-HMC) branch:

  * thetat = theta; u = coef(thetat)
  * initialise_phi_1flavour (evolution.F90:7-74): R from gauss0 per
    (spin, s), Phi = r_PV(K(1)) r(K(m)) R through two qmrherm calls with the
    remez4 blocks (x^(1/4) at am, then block 2 at am = 1, imass = 1)
  * initialise_pp (evolution.F90:76-108): pp(:,:,:,mu) from gaussp
  * H0 = hamilton; evolve_theta_pp (evolution.F90:113-179); H1 = hamilton
  * dH = H0 - H1; accept when dH >= 0, else when rano <= exp(dH)

Random numbers come from the reference's per-site generator
(`site_random.SiteRandom`): gauss0 and gaussp draw on every site, the
trajectory end and the Metropolis test on site (1,1,1), in the reference's
order, so a trajectory is reproducible from the seed alone.
"""
import numpy as np

import confio
import thirring_hip as G


def coefficient_sets(data=None):
    """The rational approximations the sweep uses (remez.F90, remezg.F90):
    pseudofermion initialisation, Hamiltonian and force ("guidance")."""
    rd = confio.remez if data is None else (lambda n: confio.read_remez(f"{data}/{n}"))
    (anum4, aden4), (bnum4, bden4) = rd("remez4")
    (anum2, aden2), (bnum2, bden2) = rd("remez2")
    (anum4g, aden4g), _ = rd("remez4g")
    (anum2g, aden2g), (bnum2g, bden2g) = rd("remez2g")
    return {"phi": (anum4, aden4, bnum4, bden4),
            # dwf3d_lib.F90:584-590: anum4/aden4, then bnum2/bden2
            "hamilton": (anum4, aden4, bnum2, bden2),
            # force_module.F90:80-96: anum4g/aden4g, bnum2g/bden2g, anum2g/aden2g
            "force": (anum4g, aden4g, bnum2g, bden2g, anum2g, aden2g)}


def initialise_pp(rng):
    """evolution.F90:76-108: one gaussp call per direction, real part kept."""
    lx, ly, lt = rng.shape
    pp = np.zeros((lx, ly, lt, 3), dtype=np.float32, order="F")
    for mu in range(3):
        pp[:, :, :, mu] = rng.gaussp()[..., 0]
    return pp


def initialise_phi_1flavour(rng, u, dims, am, imass, rescga, coeffs, am3=1.0, ops=G,
                            kernel="shamir"):
    """evolution.F90:7-74 (rational branch).  coeffs = (anum4, aden4, bnum4,
    bden4); returns (Phi haloed, (itercg of the two qmrherm calls))."""
    lx, ly, lt, ls = dims
    anum4, aden4, bnum4, bden4 = coeffs
    r = np.zeros((ls + 2, lx + 2, ly + 2, lt + 2, 4), dtype=np.complex128, order="F")
    for idirac in range(4):  # spin outer, s inner (40-51)
        for ithird in range(ls):
            ps = rng.gauss0()
            r[ithird + 1, 1:-1, 1:-1, 1:-1, idirac] = ps[..., 0] + 1j * ps[..., 1]
    G.halo5(r)
    x, it1, *_ = ops.qmrherm(r, rescga, am, imass, anum4, aden4, 0, u, dims, am3=am3,
                             kernel=kernel)
    phi, it2, *_ = ops.qmrherm(x, rescga, 1.0, 1, bnum4, bden4, 0, u, dims, am3=am3,
                               kernel=kernel)
    return phi, (it1, it2)


def trajectory(theta, rng, dims, am, imass, beta, iterl, dt, rescga, rescgg, coeffs=None,
               ibound=-1, am3=1.0, max_qmr_iters=50000, ops=G, kernel="shamir",
               use_sp=False):
    """One sweep of dwf3d_lib.F90:263-318.  theta (Lx,Ly,Lt,3) float32 is not
    modified; returns a dict with the new theta (the trial theta if accepted,
    else a copy of the old one), accepted, dH = H0 - H1, dS = S0 - S1, the MD
    step count, and H0/H1 (h, hg, hp, s, hf).  `ops` supplies coef, qmrherm,
    hamilton and evolve_theta_pp with this module's signatures (the GPU
    library by default); kernel "shamir" or "wilson" is the generation kernel
    (GENERATE_WITH_SHAMIR / GENERATE_WITH_WILSON); use_sp selects the force's
    solver (False: FP64 as the reference, True: its spmd FP32 solver,
    "mixed": the batched mixed-precision solver with FP64 correction).  The
    Hamiltonians that feed the Metropolis test always solve in FP64."""
    c = coefficient_sets() if coeffs is None else coeffs
    thetat = np.array(theta, dtype=np.float32, order="F")
    u = ops.coef(thetat, dims, ibound)
    phi, it_phi = initialise_phi_1flavour(rng, u, dims, am, imass, rescga, c["phi"], am3, ops,
                                          kernel)
    pp = initialise_pp(rng)
    h0, _ = ops.hamilton(phi, rescga, am, imass, beta, thetat, pp, u, dims, c["hamilton"], am3,
                         max_qmr_iters, kernel=kernel)
    _, steps = ops.evolve_theta_pp(iterl, np.float32(dt), pp, phi, rescgg, am, imass, beta,
                                   thetat, u, dims, c["force"], lambda: rng.rano_at(0, 0, 0),
                                   ibound, am3, use_sp=use_sp, max_qmr_iters=max_qmr_iters,
                                   kernel=kernel)
    h1, _ = ops.hamilton(phi, rescga, am, imass, beta, thetat, pp, u, dims, c["hamilton"], am3,
                         max_qmr_iters, kernel=kernel)
    dH = h0["h"] - h1["h"]
    dS = h0["s"] - h1["s"]
    y = np.exp(np.float32(dH))  # y = exp(real(dH)) (301)
    accepted = True
    if dH < 0.0:
        accepted = not (rng.rano_at(0, 0, 0) > y)  # (308-314)
    return {"theta": thetat if accepted else np.array(theta, dtype=np.float32, order="F"),
            "accepted": bool(accepted), "dH": dH, "dS": dS, "steps": steps, "pp": pp,
            "H0": h0, "H1": h1, "itercg_phi": it_phi}
