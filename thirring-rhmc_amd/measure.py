"""Noisy psibarpsi measurement with domain-wall wall sources
(measure_shamir, measure_module.F90:170-423) on one GPU.

Per noise vector the reference draws two Gaussian fields (gauss0) and solves
four systems M^dagger M x = M^dagger xi: for imass 1 and 3, sources on the
wall s = 1 (spins 1, 2), read at s = Ls, and on s = Ls (spins 3, 4), read at
s = 1; for imass 5 the sources sit on the chirality-flipped spins and are read
at their own wall.  psibarpsi is the average over `knoise` vectors of the
wall-to-wall contractions.

The Wilson-kernel measurement (measure_wilson) runs through the library in
the same way (`measure_wilson`).

Two ways to run measure_shamir:
  * `measure_shamir` — the reference's order: one congrad per source,
    starting from the source (so iteration counts, hence `aviter`, match),
    imass 1, 3, 5; on the GPU one library call per noise vector;
  * `measure_shamir_batched` (imass 1, 3) — all 4 knoise sources of the measurement
    solved together in multi-RHS launches of the batched solver to an FP64
    true residual, the fields staying on the device.  Same noise, same
    contractions; psibarpsi agrees with a converged `measure_shamir`.
"""
import numpy as np

import hiprt as hr
import thirring_hip as G


def _wall_source(dims, s, spin, eta):
    lx, ly, lt, ls = dims
    xi = np.zeros((ls + 2, lx + 2, ly + 2, lt + 2, 4), dtype=np.complex128, order="F")
    xi[s, 1:-1, 1:-1, 1:-1, spin] = eta
    return G.halo5(xi)


def _noise(ps):
    # cmplx(ps(..,1), ps(..,2)) is single precision, promoted exactly
    return ps[..., 0].astype(np.complex128) + 1j * ps[..., 1]


def _finish(p1, p2, imass, kvol):
    """psibarpsi of one noise vector (measure_module.F90:394-405)."""
    if imass == 1:
        return (p1 / kvol + p2 / kvol).real
    if imass == 3:
        return (-1j * p1 / kvol + 1j * p2 / kvol).real
    return (-1j * p1 / kvol - 1j * p2 / kvol).real


def _average(pbp, iters, knoise):
    # psibarpsi = real(sum(pbp)) is default real; aviter = iter / (4 knoise)
    return float(np.float32(np.sum(pbp)) / knoise), iters / (4 * knoise)


def _walls(dims, imass):
    """The four wall solves of one noise vector in the reference's order: for
    ids in (1, 2), side 0 (source on s = 1 from ps, summed into psibarpsi1)
    then side 1 (source on s = Ls from pt, into psibarpsi2).  Yields
    (side, src_s, src_spin, sign, read_s, read_spin), s 1-based, spins
    0-based.  imass 1, 3: eta on spin ids (side 0) or ids + 2 (side 1), read
    on the same spin at the other wall; imass 5: -eta on the chirality-flipped
    spin, read at the same wall (measure_module.F90:224-348)."""
    ls = dims[3]
    for ids in (0, 1):
        if imass != 5:
            yield 0, 1, ids, 1.0, ls, ids
            yield 1, ls, ids + 2, 1.0, 1, ids + 2
        else:
            yield 0, 1, ids, -1.0, 1, ids + 2
            yield 1, ls, ids + 2, -1.0, ls, ids


def _haloed_noise(ps):
    """gauss0's field with its periodic halo, (Lx+2, Ly+2, Lt+2, 2) float32."""
    return np.asfortranarray(np.pad(ps, ((1, 1), (1, 1), (1, 1), (0, 0)), mode="wrap"))


def measure_shamir(u, dims, res, am, imass, rng, knoise=10, am3=1.0, ops=G, niterc=None):
    """measure_shamir (measure_module.F90:170-423) on one process: returns
    (psibarpsi, aviter).  u is the haloed link array, rng a
    site_random.SiteRandom (draws ps, pt = gauss0(), gauss0() per noise
    vector, as the reference does).  With `ops` the GPU library (the default)
    each noise vector's four solves run in one call of
    thr_measure_shamir_noise, fields staying on the device; any other `ops`
    supplies dslashd and congrad with the signatures of thirring_hip and the
    solves run one by one through it (the oracle, in the tests).  niterc is
    congrad's iteration cap (default Ls V4, params.F90:57)."""
    if imass not in (1, 3, 5):
        raise ValueError("imass must be 1, 3 or 5")
    lx, ly, lt, ls = dims
    pbp, iters = [], 0
    for _ in range(knoise):
        ps, pt = rng.gauss0(), rng.gauss0()
        if ops is G:
            p1, p2, it = G.measure_shamir_noise(_haloed_noise(ps), _haloed_noise(pt), u, dims,
                                                res, am, imass, am3=am3, niterc=niterc)
            iters += it
        else:
            p1 = p2 = 0j
            for side, src_s, src_spin, sign, read_s, read_spin in _walls(dims, imass):
                eta = _noise(ps if side == 0 else pt)
                xi = _wall_source(dims, src_s, src_spin, sign * eta)
                phi = G.halo5(ops.dslashd(xi, u, dims, am, imass, am3=am3))
                # x starts at xi
                it, _ = ops.congrad(phi, xi, u, dims, res, am, imass, am3=am3, niterc=niterc)
                iters += it
                acc = np.sum(np.conj(eta) * xi[read_s, 1:-1, 1:-1, 1:-1, read_spin])
                if side == 0:
                    p1 += acc
                else:
                    p2 += acc
        pbp.append(_finish(p1, p2, imass, lx * ly * lt))
    return _average(pbp, iters, knoise)


def measure_wilson(u, dims, res, am, imass, rng, knoise=10, am3=1.0, niterc=None,
                   kernel="shamir"):
    """measure_wilson (measure_module.F90:431-640), the Wilson-kernel
    psibarpsi, imass 1 or 3: returns (psibarpsi, aviter).  Per noise vector
    (ps, pt = gauss0(), gauss0()) one library call runs the eight wall-source
    solves on the device: dslashd_wilson, congrad with `kernel` (the
    generation kernel the caller's congrad uses) and the DWilson projection
    of the read slice."""
    if imass not in (1, 3):
        raise ValueError("measure_wilson supports imass 1 and 3")
    lx, ly, lt, ls = dims
    pbp, iters = [], 0
    for _ in range(knoise):
        ps, pt = rng.gauss0(), rng.gauss0()
        p1, p2, it = G.measure_wilson_noise(_haloed_noise(ps), _haloed_noise(pt), u, dims, res,
                                            am, imass, am3=am3, niterc=niterc, kernel=kernel)
        iters += it
        pbp.append(_finish(p1, p2, imass, lx * ly * lt))
    return _average(pbp, iters, knoise)


def _sources(rng, dims, knoise):
    """The 4 knoise wall sources (imass 1, 3) in the reference's draw order,
    for the batched form: yields (k, wall, spin, read, eta)."""
    for k in range(knoise):
        ps, pt = rng.gauss0(), rng.gauss0()
        for side, src_s, src_spin, _, read_s, _ in _walls(dims, 3):
            yield k, src_s, src_spin, read_s, _noise(ps if side == 0 else pt)


def measure_shamir_batched(links64, links32, dims, am, imass, rng, knoise=10, tol=1e-10,
                           maxcg=200000, am3=1.0, th32=1e-4, stream=None):
    """The same measurement with the 4 knoise solves batched (up to 32
    right-hand sides per launch): M^dagger xi for every source in one
    multi-RHS Dslash, then (M^dagger M) x = M^dagger xi by the batched
    solver to FP64 true residual `tol` (FP32 Krylov storage below `th32`,
    FP64 defect correction).  links64/links32 are device links.  Returns
    (psibarpsi, {"iterations", "max_true_residual", "sources"})."""
    if imass not in (1, 3):
        raise ValueError("measure_shamir supports imass 1 and 3")
    lx, ly, lt, ls = dims
    v5 = lx * ly * lt * ls
    srcs = list(_sources(rng, dims, knoise))
    p1 = np.zeros(knoise, dtype=np.complex128)
    p2 = np.zeros(knoise, dtype=np.complex128)
    iters, worst = 0, 0.0
    for c0 in range(0, len(srcs), G.MAXRHS):
        chunk = srcs[c0:c0 + G.MAXRHS]
        n = len(chunk)
        xi = np.zeros((n, 4, lt, ly, lx, ls), dtype=np.complex128)  # device order
        for j, (_, wall, spin, _, eta) in enumerate(chunk):
            xi[j, spin, :, :, :, wall - 1] = eta.transpose(2, 1, 0)
        xh = np.stack([xi.real, xi.imag], axis=-1).reshape(n, 4 * v5, 2)
        xd = hr.DeviceArray.from_host(xh)
        bd = hr.DeviceArray((n, 4 * v5, 2), np.float64)
        G.dslash_dev(bd, xd, links64, dims, np.full(n, am, np.float32), imass=imass, am3=am3,
                     dagger=True, nrhs=n, stream=stream)
        xd.free()
        x = hr.DeviceArray((n, 1, 4 * v5, 2), np.float64)
        it, _, rel = G.multishift_fast_dev(links64, links32, dims, am, imass, np.zeros(1), bd, x,
                                           tol=tol, maxcg=maxcg, nsrc=n, th32=th32, am3=am3,
                                           true_residuals=True, stream=stream)
        sol = x.to_host()
        bd.free()
        x.free()
        iters += int(it)
        worst = max(worst, float(np.max(rel)))
        sol = (sol[..., 0] + 1j * sol[..., 1]).reshape(n, 4, lt, ly, lx, ls)
        for j, (k, wall, spin, read, eta) in enumerate(chunk):
            acc = np.sum(np.conj(eta) * sol[j, spin, :, :, :, read - 1].transpose(2, 1, 0))
            if wall == 1:
                p1[k] += acc
            else:
                p2[k] += acc
    pbp = [_finish(p1[k], p2[k], imass, lx * ly * lt) for k in range(knoise)]
    psibarpsi, _ = _average(pbp, 0, knoise)
    return psibarpsi, {"iterations": iters, "max_true_residual": worst, "sources": len(srcs)}
