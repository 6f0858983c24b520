"""Restatement of the reference's Wilson-kernel psibarpsi measurement for its
known-answer test: measure_wilson (measure_module.F90:431-640), imass 1 or 3,
on one process.  The operators are passed in, so the same composition runs
on the oracle and on the GPU front end.  (measure_shamir is a library module:
thirring-rhmc_amd/measure.py.)"""
import numpy as np

import refstate
from measure import _average, _finish, _noise, _wall_source


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
