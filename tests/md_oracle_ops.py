"""The oracle behind the call signatures `thirring-rhmc_amd/md.py` uses, so the
same trajectory code runs on the CPU restatement (test infrastructure)."""

import oracle as O


class OracleOps:
    coef = staticmethod(O.coef)
    qmrherm = staticmethod(O.qmrherm)
    hamilton = staticmethod(O.hamilton)

    @staticmethod
    def evolve_theta_pp(iterl, dt, pp, phi, res, am, imass, beta, theta, u, dims, coeffs_g, ytest,
                        ibound=-1, am3=1.0, use_sp=False, max_qmr_iters=50000, kernel="shamir"):
        th, p, uu, ds, k = O.evolve_theta_pp(iterl, dt, pp, phi, res, am, imass, beta, theta, dims,
                                             coeffs_g, ytest, ibound, am3, use_sp, max_qmr_iters,
                                             kernel)
        theta[...] = th
        pp[...] = p
        u[...] = uu
        return ds, k
