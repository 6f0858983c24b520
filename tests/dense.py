"""Dense-matrix helpers for brute-force checks on tiny lattices (test-only)."""
import numpy as np


def rand_spinor(dims, rng, dtype=np.complex128):
    import refstate
    lx, ly, lt, ls = dims
    a = np.zeros((ls + 2, lx + 2, ly + 2, lt + 2, 4), dtype=dtype, order="F")
    v = rng.standard_normal((ls, lx, ly, lt, 4)) + 1j * rng.standard_normal((ls, lx, ly, lt, 4))
    a[1:-1, 1:-1, 1:-1, 1:-1, :] = v / np.sqrt(2.0)
    return refstate.halo5(a)


def rand_links(dims, rng, kind="generic"):
    """Random haloed links.  kind='generic': arbitrary complex; 'theta': 1 + i theta."""
    import refstate
    lx, ly, lt, ls = dims
    u = np.zeros((lx + 2, ly + 2, lt + 2, 3), dtype=np.complex128, order="F")
    if kind == "theta":
        th = rng.standard_normal((lx, ly, lt, 3))
        u[1:-1, 1:-1, 1:-1, :] = 1.0 + 1j * th
    else:
        u[1:-1, 1:-1, 1:-1, :] = (rng.standard_normal((lx, ly, lt, 3))
                                  + 1j * rng.standard_normal((lx, ly, lt, 3)))
    return refstate.halo4(u)


def interior_vec(a):
    return np.asarray(a[1:-1, 1:-1, 1:-1, 1:-1, :]).reshape(-1, order="F")


def from_vec(v, dims):
    import refstate
    lx, ly, lt, ls = dims
    a = np.zeros((ls + 2, lx + 2, ly + 2, lt + 2, 4), dtype=np.complex128, order="F")
    a[1:-1, 1:-1, 1:-1, 1:-1, :] = v.reshape((ls, lx, ly, lt, 4), order="F")
    return refstate.halo5(a)


def dense(apply, dims):
    """Dense matrix of a linear map given as apply(haloed spinor) -> haloed spinor."""
    lx, ly, lt, ls = dims
    n = ls * lx * ly * lt * 4
    M = np.zeros((n, n), dtype=np.complex128)
    for k in range(n):
        e = np.zeros(n, dtype=np.complex128)
        e[k] = 1.0
        M[:, k] = interior_vec(apply(from_vec(e, dims)))
    return M


def gamma_tables():
    """Dirac algebra of the reference as 4x4 matrices G[mu] (mu = 0..5), unscaled.

    Table from gammamatrices.F90:27-125: row alpha has the single entry
    gamval(mu, alpha) in column gamin(mu, alpha)."""
    val = [[-1j, -1j, 1j, 1j], [-1, 1, 1, -1], [-1j, 1j, 1j, -1j], [1, 1, -1, -1],
           [-1, -1, -1, -1], [-1, -1, 1, 1]]
    col = [[4, 3, 2, 1], [4, 3, 2, 1], [3, 4, 1, 2], [1, 2, 3, 4], [3, 4, 1, 2], [3, 4, 1, 2]]
    G = np.zeros((6, 4, 4), dtype=np.complex128)
    for mu in range(6):
        for a in range(4):
            G[mu, a, col[mu][a] - 1] = val[mu][a]
    return G
