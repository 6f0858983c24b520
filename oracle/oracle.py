"""ctypes front end of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker and never as the thing
measured or shipped.

Arrays follow the reference's haloed column-major layouts, held here as numpy
arrays in Fortran order:

    spinor  (Ls+2, Lx+2, Ly+2, Lt+2, 4)  complex128
    links   (Lx+2, Ly+2, Lt+2, 3)        complex128
    theta   (Lx, Ly, Lt, 3)              float32
    dSdpi   (Lx, Ly, Lt, 3)              float64
    shifts  (Ls, Lx, Ly, Lt, 4, n)       complex128 (halo-free)

`dims` is always (Lx, Ly, Lt, Ls).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "liborcthirring.so")
_lib = None

F32 = ctypes.c_float
I32 = ctypes.c_int
F64 = ctypes.c_double
P = ctypes.c_void_p
IP = ctypes.POINTER(ctypes.c_int)


def build():
    subprocess.check_call(["make", "-s", "-B", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            build()
        L = ctypes.CDLL(_LIBPATH)
        sig = {
            "orc_num_threads": ([], I32),
            "orc_set_threads": ([I32], None),
            "orc_halo5": ([P, P, I32], None),
            "orc_dwilson_slice": ([P, P, P, P, F32, I32], None),
            "orc_halo4": ([P, P], None),
            "orc_coef": ([P, P, P, I32], None),
            "orc_dslash_shamir": ([P, P, P, P, F32, I32, F32], None),
            "orc_dslashd_shamir": ([P, P, P, P, F32, I32, F32], None),
            "orc_dslash_wilson": ([P, P, P, P, F32, I32, F32], None),
            "orc_dslashd_wilson": ([P, P, P, P, F32, I32, F32], None),
            "orc_dslash_shamir_sp": ([P, P, P, P, F32, I32, F32], None),
            "orc_dslashd_shamir_sp": ([P, P, P, P, F32, I32, F32], None),
            "orc_dirac_op_shifted": ([P, P, P, P, F32, I32, F32, F64, I32], None),
            "orc_congrad": ([P, P, P, P, F32, F32, I32, F32, I32, I32, IP, IP], None),
            "orc_multishift": ([P, P, F32, I32, F32, I32, P, P, P, P, F32, I32, I32, IP, P], None),
            "orc_multishift_sp": ([P, P, F32, I32, F32, I32, P, P, P, P, F32, I32, IP, P], None),
            "orc_derivs_shamir": ([P, P, F64, I32, P, P], None),
            "orc_derivs_wilson": ([P, P, F64, I32, P, F32, I32, F32, P], None),
            "orc_qmrherm": ([P, P, F32, IP, F32, I32, F32, P, P, I32, I32, I32, P, P, I32, I32,
                             P, P, P, P], None),
            "orc_force": ([P, F32, F32, I32, F32, F32, I32, P, P, P, P, P, P, P, P, P, I32, I32,
                           I32, I32, P], None),
            "orc_hamilton": ([P, F32, F32, I32, F32, F32, P, P, P, P, P, P, P, P, I32, I32, I32, P,
                              P], None),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["F_CONTIGUOUS"] or a.ndim == 1, "oracle arrays must be Fortran-contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


def _dims(dims):
    return (ctypes.c_int * 4)(*[int(d) for d in dims])


def spinor(dims, dtype=np.complex128):
    lx, ly, lt, ls = dims
    return np.zeros((ls + 2, lx + 2, ly + 2, lt + 2, 4), dtype=dtype, order="F")


def links(dims, dtype=np.complex128):
    lx, ly, lt, ls = dims
    return np.zeros((lx + 2, ly + 2, lt + 2, 3), dtype=dtype, order="F")


def interior(a):
    """Interior view of a haloed spinor (Ls, Lx, Ly, Lt, 4)."""
    return a[1:-1, 1:-1, 1:-1, 1:-1, :]


def num_threads():
    return lib().orc_num_threads()


def set_threads(n):
    lib().orc_set_threads(int(n))


def halo5(a, dims):
    lib().orc_halo5(_p(a), _dims(dims), a.shape[4])
    return a


def dwilson_slice(r, u, dims, mass, dag=False):
    """DWilson / DWilsonD on a haloed slice (Lx+2, Ly+2, Lt+2, 4), Fortran order."""
    r = np.asfortranarray(r, dtype=np.complex128)
    out = np.zeros_like(r)
    lib().orc_dwilson_slice(_p(out), _p(r), _p(u), _dims(dims), mass, int(bool(dag)))
    return out


def halo4(u, dims):
    lib().orc_halo4(_p(u), _dims(dims))
    return u


def coef(theta, dims, ibound=-1):
    u = links(dims)
    th = np.asfortranarray(theta, dtype=np.float32)
    lib().orc_coef(_p(u), _p(th), _dims(dims), int(ibound))
    return u


def _op(name, r, u, dims, am, imass, am3):
    phi = np.zeros_like(r)
    getattr(lib(), name)(_p(phi), _p(r), _p(u), _dims(dims), am, int(imass), am3)
    return phi


def dslash(r, u, dims, am, imass, am3=1.0, kernel="shamir"):
    if r.dtype == np.complex64:
        return _op("orc_dslash_shamir_sp", r, u, dims, am, imass, am3)
    return _op("orc_dslash_" + kernel, r, u, dims, am, imass, am3)


def dslashd(r, u, dims, am, imass, am3=1.0, kernel="shamir"):
    if r.dtype == np.complex64:
        return _op("orc_dslashd_shamir_sp", r, u, dims, am, imass, am3)
    return _op("orc_dslashd_" + kernel, r, u, dims, am, imass, am3)


def _kid(kernel):
    return {"shamir": 0, "wilson": 1}[kernel]


def dirac_op_shifted(xin, u, dims, am, imass, shift=0.0, am3=1.0, kernel="shamir"):
    xout = np.zeros_like(xin)
    lib().orc_dirac_op_shifted(_p(xout), _p(xin), _p(u), _dims(dims), am, int(imass), am3,
                               float(shift), _kid(kernel))
    return xout


def congrad(phi, x, u, dims, res, am, imass, am3=1.0, niterc=None, kernel="shamir"):
    """Runs congrad in place on x; returns (itercg, iterations)."""
    lx, ly, lt, ls = dims
    if niterc is None:
        niterc = ls * lx * ly * lt  # params.F90:57
    it = ctypes.c_int()
    nx = ctypes.c_int()
    lib().orc_congrad(_p(phi), _p(x), _p(u), _dims(dims), res, am, int(imass), am3, int(niterc),
                      _kid(kernel), ctypes.byref(it), ctypes.byref(nx))
    return it.value, nx.value


def multishift(u, dims, am, imass, aden, anum, inp, res, maxcg, am3=1.0, sp=False,
               kernel="shamir"):
    lx, ly, lt, ls = dims
    n = len(aden)
    out = np.zeros((ls, lx, ly, lt, 4, n), dtype=np.complex128, order="F")
    aden = np.ascontiguousarray(aden, dtype=np.float64)
    anum = np.ascontiguousarray(anum, dtype=np.float64)
    cgr = ctypes.c_int()
    cgrs = np.zeros(n, dtype=np.int32)
    if sp:
        lib().orc_multishift_sp(_p(u), _dims(dims), am, int(imass), am3, n, _p(aden), _p(anum),
                                _p(out), _p(inp), res, int(maxcg), ctypes.byref(cgr), _p(cgrs))
    else:
        lib().orc_multishift(_p(u), _dims(dims), am, int(imass), am3, n, _p(aden), _p(anum),
                             _p(out), _p(inp), res, int(maxcg), _kid(kernel), ctypes.byref(cgr),
                             _p(cgrs))
    return out, cgr.value, cgrs


def derivs(r, x2, anum, iflag, dims, dsdpi, kernel="shamir", am=0.0, imass=3, am3=1.0):
    """Accumulates into dsdpi (Lx, Ly, Lt, 3) float64, Fortran order."""
    assert dsdpi.flags["F_CONTIGUOUS"] and dsdpi.dtype == np.float64
    if kernel == "wilson":
        lib().orc_derivs_wilson(_p(r), _p(x2), float(anum), int(iflag), _dims(dims), am,
                                int(imass), am3, _p(dsdpi))
    else:
        lib().orc_derivs_shamir(_p(r), _x(x2), float(anum), int(iflag), _dims(dims), _p(dsdpi))
    return dsdpi


def _x(a):
    return _p(a)


def qmrherm(phi, res, am, imass, anum, aden, iflag, u, dims, x1=None, phi0=None, dsdpi=None,
            use_sp=False, max_qmr_iters=50000, am3=1.0, kernel="shamir"):
    """Returns (x, itercg, x1, phi0, dsdpi, cg_returns); module arrays may be passed in."""
    lx, ly, lt, ls = dims
    n = len(aden)
    x = np.zeros_like(phi)
    if x1 is None:
        x1 = np.zeros((ls, lx, ly, lt, 4, n), dtype=np.complex128, order="F")
    if phi0 is None:
        phi0 = np.zeros((ls + 2, lx + 2, ly + 2, lt + 2, 4, 25), dtype=np.complex128, order="F")
    if dsdpi is None:
        dsdpi = np.zeros((lx, ly, lt, 3), dtype=np.float64, order="F")
    anum = np.ascontiguousarray(anum, dtype=np.float64)
    aden = np.ascontiguousarray(aden, dtype=np.float64)
    it = ctypes.c_int()
    cgrs = np.zeros(n, dtype=np.int32)
    lib().orc_qmrherm(_p(phi), _p(x), res, ctypes.byref(it), am, int(imass), am3, _p(anum),
                      _p(aden), n, int(iflag), int(bool(use_sp)), _p(u), _dims(dims),
                      int(max_qmr_iters), _kid(kernel), _p(x1), _p(phi0), _p(dsdpi), _p(cgrs))
    return x, it.value, x1, phi0, dsdpi, cgrs


def force(phi, res1, am, imass, beta, theta, u, dims, coeffs_g, ibound=-1, am3=1.0, use_sp=False,
          max_qmr_iters=50000, kernel="shamir"):
    """coeffs_g = (anum4g, aden4g, bnum2g, bden2g, anum2g, aden2g)."""
    lx, ly, lt, ls = dims
    a4, d4, b2, e2, a2, d2 = [np.ascontiguousarray(c, dtype=np.float64) for c in coeffs_g]
    dsdpi = np.zeros((lx, ly, lt, 3), dtype=np.float64, order="F")
    th = np.asfortranarray(theta, dtype=np.float32)
    lib().orc_force(_p(phi), res1, am, int(imass), am3, beta, int(ibound), _p(th), _p(u),
                    _dims(dims), _p(a4), _p(d4), _p(b2), _p(e2), _p(a2), _p(d2), len(d4),
                    int(bool(use_sp)), int(max_qmr_iters), _kid(kernel), _p(dsdpi))
    return dsdpi


def evolve_theta_pp(iterl, dt, pp, phi, res, am, imass, beta, theta, dims, coeffs_g, ytest,
                    ibound=-1, am3=1.0, use_sp=False, max_qmr_iters=50000, kernel="shamir"):
    """evolution.F90:113-179 (Nf = 1) composed from the oracle's force and coef.

    Half step of pp (126-130); then per MD step theta += dt*pp in float32
    (152), u = coef(theta) (155), force (156); ytest() < 1/iterl ends the
    trajectory with a closing half step (159-168), else a full step (170).
    pp updates are real(c*dSdpi) with c = 0.5*dt or dt a float32 product.
    Returns (theta, pp, u, dSdpi, steps) with steps = min(iter, 4*iterl) (178).
    """
    f32 = np.float32
    theta = np.array(theta, dtype=np.float32, order="F")
    pp = np.array(pp, dtype=np.float32, order="F")
    half, full = np.float64(f32(0.5) * f32(dt)), np.float64(f32(dt))
    u = coef(theta, dims, ibound)
    ds = force(phi, res, am, imass, beta, theta, u, dims, coeffs_g, ibound, am3, use_sp,
               max_qmr_iters, kernel)
    pp = np.asfortranarray(pp - (half * ds).astype(np.float32))
    proby = f32(1.0) / f32(iterl)
    steps = 4 * iterl
    for it in range(1, 4 * iterl + 1):
        theta = np.asfortranarray(theta + f32(dt) * pp)
        u = coef(theta, dims, ibound)
        ds = force(phi, res, am, imass, beta, theta, u, dims, coeffs_g, ibound, am3, use_sp,
                   max_qmr_iters, kernel)
        if f32(ytest()) < proby:
            pp = np.asfortranarray(pp - (half * ds).astype(np.float32))
            steps = it
            break
        pp = np.asfortranarray(pp - (full * ds).astype(np.float32))
    return theta, pp, u, ds, steps


def hamilton(phi, res2, am, imass, beta, theta, pp, u, dims, coeffs, am3=1.0,
             max_qmr_iters=50000, kernel="shamir"):
    """coeffs = (anum4, aden4, bnum2, bden2); returns ({h,hg,hp,s,hf}, (it1, it2))."""
    a4, d4, b2, e2 = [np.ascontiguousarray(c, dtype=np.float64) for c in coeffs]
    th = np.asfortranarray(theta, dtype=np.float32)
    ppf = np.asfortranarray(pp, dtype=np.float32)
    out = np.zeros(5)
    its = np.zeros(2, dtype=np.int32)
    lib().orc_hamilton(_p(phi), res2, am, int(imass), am3, beta, _p(th), _p(ppf), _p(u),
                       _dims(dims), _p(a4), _p(d4), _p(b2), _p(e2), len(d4), int(max_qmr_iters),
                       _kid(kernel), _p(out), _p(its))
    keys = ("h", "hg", "hp", "s", "hf")
    return dict(zip(keys, out.tolist())), (int(its[0]), int(its[1]))
