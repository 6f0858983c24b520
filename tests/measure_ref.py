"""Restatement of the reference's psibarpsi measurements for known-answer
tests: measure_shamir (measure_module.F90:170-423) and measure_wilson
(measure_module.F90:431-640), imass 1 or 3, on one process.  The operators
are passed in, so the same composition runs on the oracle and on the GPU
front end."""
import numpy as np

import refstate


def _wall_source(dims, s, spin, eta):
    lx, ly, lt, ls = dims
    xi = np.zeros((ls + 2, lx + 2, ly + 2, lt + 2, 4), dtype=np.complex128, order="F")
    xi[s, 1:-1, 1:-1, 1:-1, spin] = eta
    return refstate.halo5(xi)


def _noise(ps):
    # cmplx(ps(..,1), ps(..,2)) is single precision, promoted exactly
    return ps[..., 0].astype(np.complex128) + 1j * ps[..., 1]


def _finish(p1, p2, imass, kvol):
    if imass == 1:
        return (p1 / kvol + p2 / kvol).real
    return (-1j * p1 / kvol + 1j * p2 / kvol).real


def _average(pbp, iters, knoise):
    # psibarpsi = real(sum(pbp)) is default real; aviter = iter / (4 knoise)
    return float(np.float32(np.sum(pbp)) / knoise), iters / (4 * knoise)


def measure_shamir(dslashd, congrad, u, dims, res, am, imass, rng, knoise=10):
    lx, ly, lt, ls = dims
    pbp, iters = [], 0
    for _ in range(knoise):
        ps, pt = rng.gauss0(), rng.gauss0()
        p1 = p2 = 0j
        for ids in (0, 1):
            # source on the wall s = 1 (spin ids), read at s = Ls; then on s = Ls
            # (spin ids + 2), read at s = 1.  congrad starts from the source.
            for wall, noise, spin, read in ((1, ps, ids, ls), (ls, pt, ids + 2, 1)):
                eta = _noise(noise)
                xi = _wall_source(dims, wall, spin, eta)
                phi = refstate.halo5(dslashd(xi, u, dims, am, imass))
                it, _ = congrad(phi, xi, u, dims, res, am, imass)
                iters += it
                acc = np.sum(np.conj(eta) * xi[read, 1:-1, 1:-1, 1:-1, spin])
                if wall == 1:
                    p1 += acc
                else:
                    p2 += acc
        pbp.append(_finish(p1, p2, imass, lx * ly * lt))
    return _average(pbp, iters, knoise)


def measure_wilson(dslashd_wilson, congrad, dwilson_slice, u, dims, res, am, imass, rng,
                   knoise=10, am3=np.float32(1.0)):
    lx, ly, lt, ls = dims
    pbp, iters = [], 0
    for _ in range(knoise):
        ps, pt = rng.gauss0(), rng.gauss0()
        p1 = p2 = 0j
        for first in (True, False):
            for ids in range(4):
                eta = _noise(ps if first else pt)
                wall, read = (1, ls) if first else (ls, 1)
                xi = _wall_source(dims, wall, ids, eta)
                phi = refstate.halo5(dslashd_wilson(xi, u, dims, am, imass))
                it, _ = congrad(phi, xi, u, dims, res, am, imass)
                iters += it
                refstate.halo5(xi)
                # P+ (spins 1,2) at s = Ls, or P- (spins 3,4) at s = 1, through DWilson - 1
                isl = np.zeros((lx + 2, ly + 2, lt + 2, 4), dtype=np.complex128, order="F")
                sp = slice(0, 2) if first else slice(2, 4)
                isl[..., sp] = xi[read, :, :, :, sp]
                xi[read] = dwilson_slice(isl, u, dims, -am3) - isl
                acc = np.sum(np.conj(eta) * xi[read, 1:-1, 1:-1, 1:-1, ids])
                if first:
                    p1 += acc
                else:
                    p2 += acc
        pbp.append(_finish(p1, p2, imass, lx * ly * lt))
    return _average(pbp, iters, knoise)
