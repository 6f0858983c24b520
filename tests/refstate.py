"""Closed-form test fields of the reference's unit tests, restated in numpy.

Follows `generate_starting_state` (tests/test_utils.F90:35-151) including its
mixed-kind arithmetic: the literals 1.1, 1.3, 0.5*exp(1.0) and 1.7*exp(1.0)
are default-real (float32) values promoted to float64, and
`iunit*idx*tau/idxmax` is evaluated as a float64 phase.
"""
import numpy as np

TAU = 8.0 * np.arctan(1.0)


def _f32(x):
    return float(np.float32(x))


def starting_state(dims, want=("phi", "r", "x", "u")):
    """Returns a dict of haloed Fortran-order arrays (periodic halos filled)."""
    lx, ly, lt, ls = dims
    ksize, ksizet, kthird = lx, lt, ls
    assert lx == ly, "the reference's tests use ksize for both x and y"
    idxmax = 4 * ksize * ksize * ksizet * kthird
    s = np.arange(1, ls + 1)[:, None, None, None, None]
    x = np.arange(1, lx + 1)[None, :, None, None, None]
    y = np.arange(1, ly + 1)[None, None, :, None, None]
    t = np.arange(1, lt + 1)[None, None, None, :, None]
    j = np.arange(1, 5)[None, None, None, None, :]
    idx = (s + (x - 1) * kthird + (y - 1) * kthird * ksize + (t - 1) * kthird * ksize * ksize
           + (j - 1) * kthird * ksize * ksize * ksizet)
    phase = np.exp(1j * (idx.astype(np.float64) * TAU / idxmax))
    e32 = float(np.exp(np.float32(1.0)))  # exp(1.0) in default real
    out = {}

    def haloed(vals):
        a = np.zeros((ls + 2, lx + 2, ly + 2, lt + 2, 4), dtype=np.complex128, order="F")
        a[1:-1, 1:-1, 1:-1, 1:-1, :] = vals
        return halo5(a)

    if "phi" in want:
        out["phi"] = haloed(_f32(1.1) * phase)
    if "r" in want:
        out["r"] = haloed(_f32(1.3) * phase)
    if "x" in want:
        out["x"] = haloed(_f32(np.float32(0.5) * np.float32(e32)) * phase)
    if "phi0" in want:
        c = _f32(np.float32(1.7) * np.float32(e32))
        p0 = np.zeros((ls + 2, lx + 2, ly + 2, lt + 2, 4, 25), dtype=np.complex128, order="F")
        for l in range(1, 26):
            p0[1:-1, 1:-1, 1:-1, 1:-1, :, l - 1] = c * phase + l
            p0[..., l - 1] = halo5(np.asfortranarray(p0[..., l - 1]))
        out["phi0"] = p0
    if "u" in want or "dsdpi" in want:
        xx = np.arange(1, lx + 1)[:, None, None, None]
        yy = np.arange(1, ly + 1)[None, :, None, None]
        tt = np.arange(1, lt + 1)[None, None, :, None]
        jj = np.arange(1, 4)[None, None, None, :]
        idx4 = xx + (yy - 1) * ksize + (tt - 1) * ksize * ksize + (jj - 1) * ksize * ksize * ksizet
        ph4 = np.exp(1j * (idx4.astype(np.float64) * TAU / idxmax))
        if "u" in want:
            u = np.zeros((lx + 2, ly + 2, lt + 2, 3), dtype=np.complex128, order="F")
            u[1:-1, 1:-1, 1:-1, :] = ph4
            out["u"] = halo4(u)
        if "dsdpi" in want:
            out["dsdpi"] = np.asfortranarray((TAU * ph4).real)
    return out


def halo5(a):
    """Periodic halo of a haloed spinor (uncomms5.f90:6-19 semantics)."""
    a[0] = a[-2]
    a[-1] = a[1]
    a[:, 0] = a[:, -2]
    a[:, -1] = a[:, 1]
    a[:, :, 0] = a[:, :, -2]
    a[:, :, -1] = a[:, :, 1]
    a[:, :, :, 0] = a[:, :, :, -2]
    a[:, :, :, -1] = a[:, :, :, 1]
    return a


def halo4(u):
    u[0] = u[-2]
    u[-1] = u[1]
    u[:, 0] = u[:, -2]
    u[:, -1] = u[:, 1]
    u[:, :, 0] = u[:, :, -2]
    u[:, :, -1] = u[:, :, 1]
    return u
