"""Python front end of libthirring_hip.so (the gfx950 domain-wall hot path).

Mirrors the reference's Fortran subroutine API (dirac.F90, measure_module.F90
congrad, multishift_module.F90, qmrherm_module.F90, derivs_module.F90,
gauge.F90 coef) on the reference's own haloed column-major arrays, held as
numpy arrays in Fortran order, and exposes the device-resident entry points
for callers that keep fields on the GPU.

There is no CPU fallback: if the shared library is missing or fails to load,
every call raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libthirring_hip.so")

F32 = ctypes.c_float
I32 = ctypes.c_int
F64 = ctypes.c_double
P = ctypes.c_void_p
IP = ctypes.POINTER(ctypes.c_int)
UFN = ctypes.CFUNCTYPE(ctypes.c_float, ctypes.c_void_p)  # thr_uniform_fn

PREC = {"f64": 0, "f32": 1, "f16": 2}
MAXRHS = 32  # right-hand sides per multi-RHS launch (nrhs 1..32 in the C ABI)
KERNEL = {"shamir": 0, "wilson": 1}

_SIGS = {
    "thr_init": ([I32], I32),
    "thr_finalize": ([], I32),
    "thr_last_error": ([], ctypes.c_char_p),
    "thr_context_create": ([ctypes.POINTER(P)], I32),
    "thr_context_destroy": ([P], I32),
    "thr_context_bind": ([P], I32),
    "thr_context_bytes": ([ctypes.POINTER(ctypes.c_ulonglong)], I32),
    "thr_context_trim": ([], I32),
    "thr_context_set_reliable": ([ctypes.c_double], I32),
    "thr_context_reliable_updates": ([ctypes.POINTER(ctypes.c_int)], I32),
    "thr_coef": ([P, P, P, I32], I32),
    "thr_dslash": ([P, P, P, P, F32, I32, F32, I32], I32),
    "thr_dslashd": ([P, P, P, P, F32, I32, F32, I32], I32),
    "thr_dirac_op_shifted": ([P, P, P, P, F32, I32, F32, F64, I32], I32),
    "thr_congrad": ([P, P, P, P, F32, F32, I32, F32, I32, I32, IP, IP], I32),
    "thr_noalloc_scratch_bytes": ([], ctypes.c_ulonglong),
    "thr_congrad_dev_noalloc": ([P, P, P, P, F32, F32, I32, F32, I32, IP, IP, P], I32),
    "thr_derivs_dev": ([P, P, F64, I32, P, F32, I32, I32, P, P], I32),
    "thr_derivs_batch_dev": ([P, P, P, P, I32, P, P, P], I32),
    "thr_measure_shamir_noise": ([P, P, P, P, F32, F32, I32, F32, I32, I32, P, P, IP], I32),
    "thr_measure_wilson_noise": ([P, P, P, P, F32, F32, I32, F32, I32, I32, P, P, IP], I32),
    "thr_multishift": ([P, P, F32, I32, F32, I32, P, P, P, P, F32, I32, I32, I32, IP, P], I32),
    "thr_derivs": ([P, P, F64, I32, P, F32, I32, F32, I32, P], I32),
    "thr_qmrherm": ([P, P, F32, IP, F32, I32, F32, P, P, I32, I32, I32, P, P, I32, I32, P, P, P,
                     P], I32),
    "thr_force": ([P, F32, F32, I32, F32, F32, I32, P, P, P, P, P, P, P, P, P, I32, I32, I32, I32,
                   P, P], I32),
    "thr_hamilton": ([P, F32, F32, I32, F32, F32, P, P, P, P, P, P, P, P, I32, I32, I32, I32, P,
                      P], I32),
    "thr_evolve_theta_pp": ([I32, F32, P, P, F32, F32, I32, F32, F32, I32, P, P, P, P, P, P, P,
                             P, P, I32, I32, I32, I32, UFN, P, P, IP], I32),
    "thr_evolve_theta_pp_dev": ([I32, F32, P, P, F32, F32, I32, F32, F32, I32, P, P, P, P, P, P,
                                 P, P, P, I32, I32, I32, I32, UFN, P, P, IP, P], I32),
    "thr_coef_dev": ([P, P, P, I32, P], I32),
    "thr_dslash_autotune": ([P, P, I32, IP, IP, P], I32),
    "thr_dslash_autotune_nrhs": ([P, P, I32, I32, IP, IP, P], I32),
    "thr_dslash_set_config": ([P, I32, I32, I32], I32),
    "thr_dslash_get_config": ([P, I32, IP, IP], I32),
    "thr_dslash_get_config_nrhs": ([P, I32, I32, IP, IP], I32),
    "thr_links_to_f32_dev": ([P, P, P, P], I32),
    "thr_device_copy": ([P, P, ctypes.c_ulonglong, P], I32),
    "thr_gaussian_dev": ([P, ctypes.c_ulonglong, I32, F64, ctypes.c_ulonglong, P], I32),
    "thr_norms_dev": ([P, I32, ctypes.c_ulonglong, P, P], I32),
    "thr_dslash_dev": ([P, P, P, P, P, P, I32, P, I32, F32, I32, I32, P], I32),
    "thr_congrad_dev": ([P, P, P, P, F32, F32, I32, F32, I32, I32, IP, IP, P], I32),
    "thr_multishift_dev": ([P, P, F32, I32, F32, I32, P, P, P, P, F32, I32, I32, I32, IP, P, P],
                           I32),
    "thr_true_residuals_dev": ([P, P, I32, P, I32, F32, I32, P, P, P, P, P], I32),
    "thr_multishift_plain_dev": ([P, P, F32, I32, F32, I32, P, P, P, F32, I32, IP, P, P], I32),
    "thr_multishift_fast_dev": ([P, P, P, I32, P, I32, F32, I32, P, P, P, F64, I32, F64, F64, I32,
                                 IP, P, IP, P, P, P], I32),
}

_lib = None


class ThrError(RuntimeError):
    """A library call returned a THR_ERR_* status (`status`)."""

    def __init__(self, msg, status=None):
        super().__init__(msg)
        self.status = status


class ThrNoConvergence(ThrError):
    """THR_ERR_NOCONV: a solver left a residual above its target."""


class ThrNonFinite(ThrError):
    """THR_ERR_NONFINITE: a solver's dot products came out NaN or Inf."""


STATUS = {"ok": 0, "arg": 1, "hip": 2, "noconv": 3, "size": 4, "state": 5, "nonfinite": 6}


def lib():
    """Loads the HIP library (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ThrError(f"{LIBPATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIBPATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


_initialised = False


def init(device=0):
    global _initialised
    if not _initialised:
        _check(lib().thr_init(int(device)), "thr_init")
        _initialised = True


def _check(rc, what):
    if rc != 0:
        msg = lib().thr_last_error()
        cls = {STATUS["noconv"]: ThrNoConvergence, STATUS["nonfinite"]: ThrNonFinite}.get(rc,
                                                                                         ThrError)
        raise cls(f"{what} failed (status {rc}): {msg.decode() if msg else ''}", rc)


def _p(a):
    if a is None:
        return None
    if isinstance(a, int):
        return ctypes.c_void_p(a)
    if hasattr(a, "data_ptr"):
        return ctypes.c_void_p(a.data_ptr())
    assert a.flags["F_CONTIGUOUS"] or a.ndim == 1, "arrays must be Fortran-contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


def _dims(dims):
    return (ctypes.c_int * 4)(*[int(d) for d in dims])


def halo5(a):
    """Fills, in place, the one-cell halo of a haloed spinor
    (Ls+2, Lx+2, Ly+2, Lt+2, 4) with periodic copies of the interior: x, y, t
    as update_halo_5 (uncomms5.f90:6-19), then Ls as the third-axis update.
    The library's kernels wrap indices themselves; routines that read the
    haloed array directly (the reference's own) need the halo."""
    for ax in (1, 2, 3, 0):
        n = a.shape[ax]
        lo, hi, first, last = ([slice(None)] * a.ndim for _ in range(4))
        lo[ax], hi[ax], first[ax], last[ax] = 0, n - 1, 1, n - 2
        a[tuple(lo)] = a[tuple(last)]
        a[tuple(hi)] = a[tuple(first)]
    return a


def set_reliable_updates(delta):
    """Reliable-update factor of the mixed multishift in the bound context
    (0 disables, default 0.1): while the Krylov storage is FP32/FP16 the base
    residual is replaced by the FP64 true residual of the smallest active
    shift each time the residual has fallen by this factor."""
    init()
    _check(lib().thr_context_set_reliable(float(delta)), "thr_context_set_reliable")


def reliable_updates():
    """Reliable updates taken by the bound context's last multishift solve."""
    init()
    v = ctypes.c_int()
    _check(lib().thr_context_reliable_updates(ctypes.byref(v)), "thr_context_reliable_updates")
    return int(v.value)


class Context:
    """Private scratch buffers + stream for one thread of work."""

    def __init__(self):
        init()
        h = ctypes.c_void_p()
        _check(lib().thr_context_create(ctypes.byref(h)), "thr_context_create")
        self.handle = h.value

    def bind(self):
        _check(lib().thr_context_bind(ctypes.c_void_p(self.handle)), "thr_context_bind")

    @staticmethod
    def unbind():
        _check(lib().thr_context_bind(None), "thr_context_bind")

    def scratch_bytes(self):
        v = ctypes.c_ulonglong()
        _check(lib().thr_context_bytes(ctypes.byref(v)), "thr_context_bytes")
        return int(v.value)

    def trim(self):
        """Releases this context's scratch buffers (it must be the bound one)."""
        _check(lib().thr_context_trim(), "thr_context_trim")

    def destroy(self):
        if self.handle:
            _check(lib().thr_context_destroy(ctypes.c_void_p(self.handle)), "thr_context_destroy")
            self.handle = None


# ---- host-array drop-in API (reference layouts) ---------------------------

def coef(theta, dims, ibound=-1):
    init()
    lx, ly, lt, ls = dims
    u = np.zeros((lx + 2, ly + 2, lt + 2, 3), dtype=np.complex128, order="F")
    th = np.asfortranarray(theta, dtype=np.float32)
    _check(lib().thr_coef(_p(u), _p(th), _dims(dims), int(ibound)), "coef")
    return u


def dslash(r, u, dims, am, imass, am3=1.0, kernel="shamir"):
    init()
    phi = np.zeros_like(r)
    _check(lib().thr_dslash(_p(phi), _p(r), _p(u), _dims(dims), am, int(imass), am3,
                            KERNEL[kernel]), "dslash")
    return phi


def dslashd(r, u, dims, am, imass, am3=1.0, kernel="shamir"):
    init()
    phi = np.zeros_like(r)
    _check(lib().thr_dslashd(_p(phi), _p(r), _p(u), _dims(dims), am, int(imass), am3,
                             KERNEL[kernel]), "dslashd")
    return phi


def dirac_op_shifted(xin, u, dims, am, imass, shift=0.0, am3=1.0, kernel="shamir"):
    init()
    xout = np.zeros_like(xin)
    _check(lib().thr_dirac_op_shifted(_p(xout), _p(xin), _p(u), _dims(dims), am, int(imass), am3,
                                      float(shift), KERNEL[kernel]), "dirac_op_shifted")
    return xout


def congrad(phi, x, u, dims, res, am, imass, am3=1.0, niterc=None, kernel="shamir"):
    """In place on x (vector::X); returns (itercg, iterations)."""
    init()
    lx, ly, lt, ls = dims
    if niterc is None:
        niterc = ls * lx * ly * lt
    it, nx = ctypes.c_int(), ctypes.c_int()
    _check(lib().thr_congrad(_p(phi), _p(x), _p(u), _dims(dims), res, am, int(imass), am3,
                             int(niterc), KERNEL[kernel], ctypes.byref(it), ctypes.byref(nx)),
           "congrad")
    return it.value, nx.value


def _measure_noise(fn, what, ps, pt, u, dims, res, am, imass, am3, niterc, kernel):
    init()
    lx, ly, lt, ls = dims
    if niterc is None:
        niterc = ls * lx * ly * lt
    for a in (ps, pt):
        if a.dtype != np.float32 or a.shape != (lx + 2, ly + 2, lt + 2, 2) or \
                not a.flags.f_contiguous:
            raise ValueError("noise fields must be float32 (Lx+2, Ly+2, Lt+2, 2) in Fortran order")
    p1 = np.zeros(2)
    p2 = np.zeros(2)
    it = ctypes.c_int()
    _check(fn(_p(ps), _p(pt), _p(u), _dims(dims), res, am, int(imass), am3, int(niterc),
              KERNEL[kernel], _p(p1), _p(p2), ctypes.byref(it)), what)
    return complex(p1[0], p1[1]), complex(p2[0], p2[1]), it.value


def measure_shamir_noise(ps, pt, u, dims, res, am, imass, am3=1.0, niterc=None,
                         kernel="shamir"):
    """One noise vector of measure_shamir (measure_module.F90:210-372) on the
    GPU: ps, pt are gauss0 fields, float32 haloed (Lx+2, Ly+2, Lt+2, 2) in
    Fortran order.  Returns (psibarpsi1, psibarpsi2, iters): the complex sums
    before the division by kvol and the total congrad iterations of the four
    solves."""
    return _measure_noise(lib().thr_measure_shamir_noise, "measure_shamir_noise", ps, pt, u,
                          dims, res, am, imass, am3, niterc, kernel)


def measure_wilson_noise(ps, pt, u, dims, res, am, imass, am3=1.0, niterc=None,
                         kernel="shamir"):
    """One noise vector of measure_wilson (measure_module.F90:472-606) on the
    GPU: eight wall-source solves (dslashd_wilson, congrad with `kernel`) and
    the DWilson projection of the read slices; returns (psibarpsi1,
    psibarpsi2, iters) as measure_shamir_noise."""
    return _measure_noise(lib().thr_measure_wilson_noise, "measure_wilson_noise", ps, pt, u,
                          dims, res, am, imass, am3, niterc, kernel)


def multishift(u, dims, am, imass, aden, anum, inp, res, maxcg, am3=1.0, sp=False,
               kernel="shamir"):
    init()
    lx, ly, lt, ls = dims
    n = len(aden)
    out = np.zeros((ls, lx, ly, lt, 4, n), dtype=np.complex128, order="F")
    aden = np.ascontiguousarray(aden, dtype=np.float64)
    anum = np.ascontiguousarray(anum, dtype=np.float64)
    cgr = ctypes.c_int()
    cgrs = np.zeros(n, dtype=np.int32)
    _check(lib().thr_multishift(_p(u), _dims(dims), am, int(imass), am3, n, _p(aden), _p(anum),
                                _p(out), _p(inp), res, int(maxcg), int(bool(sp)), KERNEL[kernel],
                                ctypes.byref(cgr), _p(cgrs)), "multishift")
    return out, cgr.value, cgrs


def derivs(r, x2, anum, iflag, dims, dsdpi, kernel="shamir", am=0.0, imass=3, am3=1.0):
    init()
    assert dsdpi.flags["F_CONTIGUOUS"] and dsdpi.dtype == np.float64
    _check(lib().thr_derivs(_p(r), _p(x2), float(anum), int(iflag), _dims(dims), am, int(imass),
                            am3, KERNEL[kernel], _p(dsdpi)), "derivs")
    return dsdpi


def qmrherm(phi, res, am, imass, anum, aden, iflag, u, dims, x1=None, phi0=None, dsdpi=None,
            use_sp=False, max_qmr_iters=50000, am3=1.0, kernel="shamir"):
    """Returns (x, itercg, x1, phi0, dsdpi, cg_returns) like the oracle."""
    init()
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
    _check(lib().thr_qmrherm(_p(phi), _p(x), res, ctypes.byref(it), am, int(imass), am3,
                             _p(anum), _p(aden), n, int(iflag), _sp(use_sp), _p(u),
                             _dims(dims), int(max_qmr_iters), KERNEL[kernel], _p(x1), _p(phi0),
                             _p(dsdpi), _p(cgrs)), "qmrherm")
    return x, it.value, x1, phi0, dsdpi, cgrs


def force(phi, res1, am, imass, beta, theta, u, dims, coeffs_g, ibound=-1, am3=1.0, use_sp=False,
          max_qmr_iters=50000, kernel="shamir"):
    """force_module.F90:9-113; coeffs_g = (anum4g, aden4g, bnum2g, bden2g, anum2g, aden2g)."""
    init()
    lx, ly, lt, ls = dims
    a4, d4, b2, e2, a2, d2 = [np.ascontiguousarray(c, dtype=np.float64) for c in coeffs_g]
    dsdpi = np.zeros((lx, ly, lt, 3), dtype=np.float64, order="F")
    th = np.asfortranarray(theta, dtype=np.float32)
    its = np.zeros(4, dtype=np.int32)
    _check(lib().thr_force(_p(phi), res1, am, int(imass), am3, beta, int(ibound), _p(th), _p(u),
                           _dims(dims), _p(a4), _p(d4), _p(b2), _p(e2), _p(a2), _p(d2), len(d4),
                           _sp(use_sp), int(max_qmr_iters), KERNEL[kernel], _p(dsdpi),
                           _p(its)), "force")
    return dsdpi, its


def _sp(use_sp):
    """use_sp argument: False/0 FP64, True/1 the reference's FP32 multishift,
    "mixed"/2 the batched mixed-precision solver with FP64 correction."""
    if isinstance(use_sp, str):
        return {"fp64": 0, "sp": 1, "mixed": 2}[use_sp]
    return int(use_sp)


def _uniform_source(ytest):
    """ctypes callback over a Python callable; an exception ends the
    trajectory at once (the callback returns 0) and is re-raised afterwards."""
    err = []

    def cb(_ctx):
        if err:
            return 0.0
        try:
            return float(ytest())
        except BaseException as e:  # re-raised by the caller of the C entry point
            err.append(e)
            return 0.0
    return UFN(cb), err


def evolve_theta_pp(iterl, dt, pp, phi, res, am, imass, beta, theta, u, dims, coeffs_g, ytest,
                    ibound=-1, am3=1.0, use_sp=False, max_qmr_iters=50000, kernel="shamir"):
    """evolution.F90:113-179 (Nf = 1).  theta, pp ((Lx,Ly,Lt,3) float32,
    Fortran order) and the haloed links u are updated in place; `ytest()`
    returns the next uniform of the caller's generator (the reference's
    rano(yran, idum, 1, 1, 1)), called once per MD step.  Returns
    (dSdpi of the last step, number of MD steps)."""
    init()
    lx, ly, lt, ls = dims
    for name, a, dt_ in (("theta", theta, np.float32), ("pp", pp, np.float32),
                         ("u", u, np.complex128)):
        if not (isinstance(a, np.ndarray) and a.dtype == dt_ and a.flags.f_contiguous):
            raise TypeError(f"{name} must be a Fortran-ordered {np.dtype(dt_).name} array "
                            "(updated in place)")
    a4, d4, b2, e2, a2, d2 = [np.ascontiguousarray(c, dtype=np.float64) for c in coeffs_g]
    dsdpi = np.zeros((lx, ly, lt, 3), dtype=np.float64, order="F")
    itot = ctypes.c_int(0)
    fn, err = _uniform_source(ytest)
    rc = lib().thr_evolve_theta_pp(int(iterl), dt, _p(pp), _p(phi), res, am, int(imass), am3,
                                   beta, int(ibound), _p(theta), _p(u), _dims(dims), _p(a4),
                                   _p(d4), _p(b2), _p(e2), _p(a2), _p(d2), len(d4),
                                   _sp(use_sp), int(max_qmr_iters), KERNEL[kernel], fn,
                                   None, _p(dsdpi), ctypes.byref(itot))
    if err:
        raise err[0]
    _check(rc, "evolve_theta_pp")
    return dsdpi, itot.value


def evolve_theta_pp_dev(iterl, dt, pp, phi, res, am, imass, beta, theta, links64, dims, coeffs_g,
                        ytest, dsdpi, ibound=-1, am3=1.0, use_sp=False, max_qmr_iters=50000,
                        kernel="shamir", stream=None):
    """evolve_theta_pp on device buffers (theta, pp [3 V4] float32, links64
    [3 V4] complex128, phi [4 V5] complex128, dsdpi [3 V4] float64); returns
    the number of MD steps."""
    init()
    a4, d4, b2, e2, a2, d2 = [np.ascontiguousarray(c, dtype=np.float64) for c in coeffs_g]
    itot = ctypes.c_int(0)
    fn, err = _uniform_source(ytest)
    rc = lib().thr_evolve_theta_pp_dev(int(iterl), dt, _p(pp), _p(phi), res, am, int(imass), am3,
                                       beta, int(ibound), _p(theta), _p(links64), _dims(dims),
                                       _p(a4), _p(d4), _p(b2), _p(e2), _p(a2), _p(d2), len(d4),
                                       _sp(use_sp), int(max_qmr_iters), KERNEL[kernel], fn,
                                       None, _p(dsdpi), ctypes.byref(itot), _stream(stream))
    if err:
        raise err[0]
    _check(rc, "evolve_theta_pp_dev")
    return itot.value


def hamilton(phi, res2, am, imass, beta, theta, pp, u, dims, coeffs, am3=1.0,
             max_qmr_iters=50000, kernel="shamir", use_sp=False):
    """dwf3d_lib.F90:511-616; coeffs = (anum4, aden4, bnum2, bden2); use_sp
    as for qmrherm (the reference's hamilton solves in FP64)."""
    init()
    a4, d4, b2, e2 = [np.ascontiguousarray(c, dtype=np.float64) for c in coeffs]
    th = np.asfortranarray(theta, dtype=np.float32)
    ppf = np.asfortranarray(pp, dtype=np.float32)
    out = np.zeros(5)
    its = np.zeros(2, dtype=np.int32)
    _check(lib().thr_hamilton(_p(phi), res2, am, int(imass), am3, beta, _p(th), _p(ppf), _p(u),
                              _dims(dims), _p(a4), _p(d4), _p(b2), _p(e2), len(d4), _sp(use_sp),
                              int(max_qmr_iters), KERNEL[kernel], _p(out), _p(its)), "hamilton")
    keys = ("h", "hg", "hp", "s", "hf")
    return dict(zip(keys, out.tolist())), (int(its[0]), int(its[1]))


# ---- device-resident API --------------------------------------------------

def _stream(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.ptr)


def coef_dev(links64, theta_dev, dims, ibound=-1, stream=None):
    init()
    _check(lib().thr_coef_dev(_p(links64), _p(theta_dev), _dims(dims), int(ibound),
                              _stream(stream)), "coef_dev")


def links_to_f32_dev(links32, links64, dims, stream=None):
    init()
    _check(lib().thr_links_to_f32_dev(_p(links32), _p(links64), _dims(dims), _stream(stream)),
           "links_to_f32_dev")


def device_copy(dst, src, nbytes, stream=None):
    """Streaming device-to-device copy (nbytes a multiple of 16)."""
    init()
    _check(lib().thr_device_copy(_p(dst), _p(src), int(nbytes), _stream(stream)), "device_copy")


def gaussian_dev(out, n, sigma, seed, prec="f64", stream=None):
    """Fills out with n reals sigma*N(0,1) (counter-based: element i depends
    only on (seed, i)); prec "f64" (double) or "f32" (float)."""
    init()
    _check(lib().thr_gaussian_dev(_p(out), int(n), PREC[prec], float(sigma),
                                  int(seed) & 0xFFFFFFFFFFFFFFFF, _stream(stream)),
           "gaussian_dev")


def norms_dev(x, nvec, n, stream=None):
    """FP64 L2 norms of nvec complex128 vectors of n elements each."""
    init()
    out = np.zeros(int(nvec), dtype=np.float64)
    _check(lib().thr_norms_dev(_p(x), int(nvec), int(n), _p(out), _stream(stream)), "norms_dev")
    return out


def dslash_dev(out, inp, links, dims, am, imass=3, am3=1.0, dagger=False, prec="f64",
               in_norm=None, out_norm=None, nrhs=1, stream=None):
    init()
    ams = np.ascontiguousarray(np.broadcast_to(np.asarray(am, dtype=np.float32), (nrhs,)))
    _check(lib().thr_dslash_dev(_p(out), _p(out_norm), _p(inp), _p(in_norm), _p(links),
                                _dims(dims), int(nrhs), _p(ams), int(imass), am3,
                                int(bool(dagger)), PREC[prec], _stream(stream)), "dslash_dev")


SHAPE_XCD_REMAP, SHAPE_UNIFORM_LINKS, SHAPE_LANESHIFT, SHAPE_XCD_BOXES, SHAPE_LDS_X = 1, 2, 4, 8, 16
SHAPE_NT_STORES = 32


def dslash_autotune(links, dims, prec="f64", stream=None, nrhs=1):
    """Times the Dslash launch shapes (over `nrhs` right-hand sides per launch)
    and pins the fastest for these dims and precision; returns (tchunk, flags)."""
    init()
    tc, fl = ctypes.c_int(), ctypes.c_int()
    if nrhs == 1:
        rc = lib().thr_dslash_autotune(_p(links), _dims(dims), PREC[prec], ctypes.byref(tc),
                                       ctypes.byref(fl), _stream(stream))
    else:
        rc = lib().thr_dslash_autotune_nrhs(_p(links), _dims(dims), PREC[prec], int(nrhs),
                                            ctypes.byref(tc), ctypes.byref(fl), _stream(stream))
    _check(rc, "dslash_autotune")
    return tc.value, fl.value


def dslash_set_config(dims, prec, tchunk, flags):
    init()
    _check(lib().thr_dslash_set_config(_dims(dims), PREC[prec], int(tchunk), int(flags)),
           "dslash_set_config")


def dslash_get_config(dims, prec="f64", nrhs=1):
    """(tchunk, flags) used for launches of nrhs right-hand sides with these
    dims and precision (shapes are pinned per batch class: 1, or more)."""
    init()
    tc, fl = ctypes.c_int(), ctypes.c_int()
    _check(lib().thr_dslash_get_config_nrhs(_dims(dims), PREC[prec], int(nrhs), ctypes.byref(tc),
                                            ctypes.byref(fl)), "dslash_get_config")
    return tc.value, fl.value


def congrad_dev(phi, x, links64, dims, res, am, imass, am3=1.0, niterc=None, kernel="shamir",
                stream=None):
    init()
    lx, ly, lt, ls = dims
    if niterc is None:
        niterc = ls * lx * ly * lt
    it, nx = ctypes.c_int(), ctypes.c_int()
    _check(lib().thr_congrad_dev(_p(phi), _p(x), _p(links64), _dims(dims), res, am, int(imass),
                                 am3, int(niterc), KERNEL[kernel], ctypes.byref(it),
                                 ctypes.byref(nx), _stream(stream)), "congrad_dev")
    return it.value, nx.value


def derivs_dev(r, x2, anum, iflag, dims, dsdpi, am=0.0, imass=3, kernel="shamir", stream=None):
    """derivs on device fields (R, X2 spinors, dsdpi [mu][t][y][x] float64,
    accumulated in place)."""
    init()
    _check(lib().thr_derivs_dev(_p(r), _p(x2), float(anum), int(iflag), _dims(dims), am,
                                int(imass), KERNEL[kernel], _p(dsdpi), _stream(stream)),
           "derivs_dev")
    return dsdpi


def derivs_batch_dev(rs, x2s, anums, iflags, dims, dsdpi, stream=None):
    """dsdpi += the Shamir derivs of every (rs[k], x2s[k], anums[k], iflags[k])
    in one pass (Ls <= 64): the per-pole derivs of the force in one launch."""
    init()
    n = len(rs)
    if not (len(x2s) == len(anums) == len(iflags) == n):
        raise ValueError("rs, x2s, anums and iflags must have the same length")
    rp = (ctypes.c_void_p * n)(*[_p(a).value for a in rs])
    xp = (ctypes.c_void_p * n)(*[_p(a).value for a in x2s])
    an = np.ascontiguousarray(anums, dtype=np.float64)
    fl = np.ascontiguousarray(iflags, dtype=np.int32)
    _check(lib().thr_derivs_batch_dev(rp, xp, _p(an), _p(fl), int(n), _dims(dims), _p(dsdpi),
                                      _stream(stream)), "derivs_batch_dev")
    return dsdpi


def congrad_dev_noalloc(phi, x, links64, dims, res, am, imass, am3=1.0, niterc=None,
                        stream=None):
    """congrad_dev (Shamir kernel) with no allocation: scratch from the
    library's fixed device array (thr_noalloc_scratch_bytes())."""
    init()
    lx, ly, lt, ls = dims
    if niterc is None:
        niterc = ls * lx * ly * lt
    it, nx = ctypes.c_int(), ctypes.c_int()
    _check(lib().thr_congrad_dev_noalloc(_p(phi), _p(x), _p(links64), _dims(dims), res, am,
                                         int(imass), am3, int(niterc), ctypes.byref(it),
                                         ctypes.byref(nx), _stream(stream)), "congrad_dev_noalloc")
    return it.value, nx.value


def multishift_dev(links64, dims, am, imass, aden, anum, out, inp, res, maxcg, am3=1.0, sp=False,
                   kernel="shamir", stream=None):
    init()
    n = len(aden)
    aden = np.ascontiguousarray(aden, dtype=np.float64)
    anum = np.ascontiguousarray(anum, dtype=np.float64)
    cgr = ctypes.c_int()
    cgrs = np.zeros(n, dtype=np.int32)
    _check(lib().thr_multishift_dev(_p(links64), _dims(dims), am, int(imass), am3, n, _p(aden),
                                    _p(anum), _p(out), _p(inp), res, int(maxcg), int(bool(sp)),
                                    KERNEL[kernel], ctypes.byref(cgr), _p(cgrs), _stream(stream)),
           "multishift_dev")
    return cgr.value, cgrs


def multishift_plain_dev(links64, dims, am, imass, aden, inp, out, tol, maxcg, am3=1.0,
                         stream=None):
    """Plain FP64 path: reference recurrences, stop when |zeta_i||r| < tol |b|."""
    init()
    n = len(aden)
    aden = np.ascontiguousarray(aden, dtype=np.float64)
    cgr = ctypes.c_int()
    cgrs = np.zeros(n, dtype=np.int32)
    _check(lib().thr_multishift_plain_dev(_p(links64), _dims(dims), am, int(imass), am3, n,
                                          _p(aden), _p(inp), _p(out), tol, int(maxcg),
                                          ctypes.byref(cgr), _p(cgrs), _stream(stream)),
           "multishift_plain_dev")
    return cgr.value, cgrs


def multishift_fast_dev(links64, links32, dims, am, imass, aden, b, x, tol, maxcg, nsrc=1,
                        th32=0.0, th16=0.0, am3=1.0, true_residuals=True, corrections=None,
                        stats=None, stream=None):
    """Batched blocked-update multishift; returns (iterations, cg_return[nsrc], relres).

    corrections: rounds of FP64 defect correction (default: 4 when th32 or th16
    lowers the Krylov precision, else 0); stats (dict, optional) receives
    "corr_iterations", "switches" (iteration at which the Krylov storage
    entered FP64, FP32, FP16; -1 never) and "reliable" (reliable updates of
    the base residual taken, see set_reliable_updates).  Raises ThrNoConvergence when a true
    residual stays above tol (ThrNonFinite for NaN/Inf); the partial results
    are then still in x and in the exception's `result`."""
    init()
    n = len(aden)
    if corrections is None:
        corrections = 4 if (th32 > 0 or th16 > 0) else 0
    aden = np.ascontiguousarray(aden, dtype=np.float64)
    ams = np.ascontiguousarray(np.broadcast_to(np.asarray(am, dtype=np.float32), (nsrc,)))
    it, ci = ctypes.c_int(), ctypes.c_int()
    cgr = np.zeros(nsrc, dtype=np.int32)
    rel = np.zeros(nsrc * n, dtype=np.float64) if (true_residuals or corrections) else None
    sw = np.full(3, -1, dtype=np.int32)
    rc = lib().thr_multishift_fast_dev(_p(links64), _p(links32), _dims(dims), int(nsrc), _p(ams),
                                       int(imass), am3, n, _p(aden), _p(b), _p(x), float(tol),
                                       int(maxcg), float(th32), float(th16), int(corrections),
                                       ctypes.byref(it), _p(cgr), ctypes.byref(ci), _p(rel),
                                       _p(sw), _stream(stream))
    if stats is not None:
        stats["corr_iterations"] = ci.value
        stats["switches"] = [int(v) for v in sw]
        stats["reliable"] = reliable_updates()
    result = (it.value, cgr, (rel.reshape(nsrc, n) if rel is not None else None))
    try:
        _check(rc, "multishift_fast_dev")
    except ThrError as e:
        e.result = result
        raise
    return result


def true_residuals_dev(links64, dims, am, imass, aden, b, x, nsrc=1, am3=1.0, stream=None):
    """|b - (M^+M + sigma_i) x_i| / |b| in FP64 for x laid out [nsrc][nshift][4 V5]."""
    init()
    n = len(aden)
    aden = np.ascontiguousarray(aden, dtype=np.float64)
    ams = np.ascontiguousarray(np.broadcast_to(np.asarray(am, dtype=np.float32), (nsrc,)))
    rel = np.zeros(nsrc * n, dtype=np.float64)
    _check(lib().thr_true_residuals_dev(_p(links64), _dims(dims), int(nsrc), _p(ams), int(imass),
                                        am3, n, _p(aden), _p(b), _p(x), _p(rel), _stream(stream)),
           "true_residuals_dev")
    return rel.reshape(nsrc, n)
