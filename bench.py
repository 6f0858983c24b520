"""Benchmark of the domain-wall hot path on one MI355X.

Headline (`value`): Shamir Dslash bandwidth, FP64, at 32^3 x Ls=64 (imass 3,
am3 = 1, am = 0.005) on synthetic U(1) links u = 1 + i theta with
theta ~ N(0, 1/sqrt(0.3)) and unit-variance complex Gaussian sources.  One
step = one application of M or M^dagger (alternating) to a fresh source;
inputs rotate over enough buffer sets to exceed the 256 MB Infinity Cache.

    value = algorithmic bytes per application / mean kernel duration
    algorithmic bytes = 2 * V5 * 4 * 16 + 3 * V4 * 16   (inputs read once,
                                                          output written once)

Kernel durations come from HIP events on the stream the kernels run on,
around K launches issued back to back (the time per application of a stream
of Dslash applications); the line also gives the mean of events recorded
around each launch, which adds the launch gaps.

Also reported in the same JSON line:
  * FP32 / FP16 operator bandwidth, and the FP64 operator over 4 and 16
    right-hand sides per launch;
  * the fermion force kernel derivs at 16^3 x Ls=48 and 32^3 x Ls=64;
  * multishift solves/s to FP64 true residual 1e-10 for all 25 poles of the
    x^(-1/4) approximation (remez4 block 2) at 32^3 x Ls=64: the plain FP64
    GPU path (the reference's recurrences) against the batched,
    blocked-update, mixed-precision path, with the true residuals of both and
    the per-pole difference of their solutions;
  * the batched multi-RHS workload (16 sources x 25 shifts at 24^3 x Ls=32,
    per-source masses): solves/s of the batched solver against the plain
    FP64 path;
  * the RHMC force at 16^3 x Ls=48: the FP64 solver against the
    mixed-precision solver with FP64 correction (time and difference);
  * mixed against FP64 at 32^3 x Ls=64: force and dH differences;
  * solve time per iteration at launch-bound lattices (8^3 x 8, 12^2 x 24 x
    16): the persistent cooperative solver against the streaming one with
    hipGraph replay and with eager launches;
  * serving: a sweep of Poisson arrival rates, each with p50/p99 latency and
    solves/s under continuous batching (thirring-rhmc_amd/serving.py);
  * the CPU baseline: the FP64 C oracle's Dslash on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("thirring-rhmc_amd", "oracle", "tests"):
    sys.path.insert(0, os.path.join(ROOT, sub))

HBM_PEAK_GBS = 8000.0


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def dslash_bytes(v4, v5, prec, nrhs=1):
    b_c = {"f64": 16, "f32": 8, "f16": 5}[prec]
    b_link = 16 if prec == "f64" else 8
    return nrhs * 2 * v5 * 4 * b_c + 3 * v4 * b_link


def remez_block(name, block):
    """read_remez_file (remez_common_subroutines.F90:20-35): a0, then (a_i, s_i)."""
    path = os.path.join(ROOT, "thirring-rhmc_amd", "data", name)
    rows = [l.replace(",", " ").split() for l in open(path) if l.strip()]
    n = (len(rows) - 2) // 2
    blk = rows[block * (n + 1):(block + 1) * (n + 1)]
    return (float(blk[0][0]), np.array([float(r[0]) for r in blk[1:]]),
            np.array([float(r[1]) for r in blk[1:]]))


def make_fields(G, DA, dims, beta, seed):
    lx, ly, lt, ls = dims
    rng = np.random.default_rng(seed)
    th = (rng.standard_normal((lx, ly, lt, 3)) / np.sqrt(beta)).astype(np.float32)
    theta = DA.from_host(th.reshape(-1, order="F"))
    l64 = DA((lx * ly * lt * 3, 2), np.float64)
    l32 = DA((lx * ly * lt * 3, 2), np.float32)
    G.coef_dev(l64, theta, dims)
    G.links_to_f32_dev(l32, l64, dims)
    return th, l64, l32


def time_dslash(G, hr, dims, links, prec, steps, warmup, nbuf, stream, nrhs=1, am=0.005):
    """Per-launch times on `stream`: of K launches back to back (HIP events
    around the whole stream), the mean of events around each launch, and the
    event-bracketed region's; `nrhs` spinors per launch, masses `am`."""
    lx, ly, lt, ls = dims
    v5 = lx * ly * lt * ls
    rng = np.random.default_rng(1)
    ins, outs, inn, outn = [], [], [], []
    for _ in range(nbuf):
        x = (rng.standard_normal((nrhs, 4, v5, 2)) / np.sqrt(2)).astype(np.float32)
        if prec == "f16":  # [site][spin] with a per-site scale
            xs = np.ascontiguousarray(x.transpose(0, 2, 1, 3))
            sc = np.abs(xs).max(axis=(2, 3)).astype(np.float32)
            ins.append(hr.DeviceArray.from_host((xs / sc[:, :, None, None]).astype(np.float16)))
            inn.append(hr.DeviceArray.from_host(sc))
            outn.append(hr.DeviceArray((nrhs, v5), np.float32))
            outs.append(hr.DeviceArray((nrhs, v5, 4, 2), np.float16))
        else:
            dt = np.float64 if prec == "f64" else np.float32
            ins.append(hr.DeviceArray.from_host(x.astype(dt)))
            outs.append(hr.DeviceArray((nrhs, 4, v5, 2), dt))
            inn.append(None)
            outn.append(None)

    def launch(k):
        j = k % nbuf
        G.dslash_dev(outs[j], ins[j], links, dims, am, imass=3, am3=1.0, dagger=bool(k & 1),
                     prec=prec, in_norm=inn[j], out_norm=outn[j], nrhs=nrhs, stream=stream)

    for k in range(warmup):
        launch(k)
    stream.synchronize()
    ev = [(hr.Event(), hr.Event()) for _ in range(steps)]
    t0, t1 = hr.Event(), hr.Event()
    t0.record(stream)
    for k in range(steps):
        ev[k][0].record(stream)
        launch(k)
        ev[k][1].record(stream)
    t1.record(stream)
    stream.synchronize()
    per = np.array([a.elapsed_ms(b) for a, b in ev])
    # the same launches back to back, events only around the whole stream:
    # the time per launch of a stream of applications (what a solver sees)
    b0, b1 = hr.Event(), hr.Event()
    b0.record(stream)
    for k in range(steps):
        launch(k)
    b1.record(stream)
    stream.synchronize()
    for a in ins + outs + inn + outn:
        if a is not None:
            a.free()
    return float(b0.elapsed_ms(b1) / steps), float(np.mean(per)), float(t0.elapsed_ms(t1) / steps)


def multi_rhs_legs(G, hr, steps, warmup, stream, pinned=None):
    """FP64 Dslash over several right-hand sides per launch (links read once
    per launch): 4 sources at 32^3 x Ls=64, and 16 sources with masses drawn
    from {0.005, 0.01, 0.02, 0.05} at 24^3 x Ls=32; each with the shape tuned
    for its batch (the library pins shapes per batch class)."""
    out = {}
    rng = np.random.default_rng(7)
    for dims, nrhs in (((32, 32, 32, 64), 4), ((24, 24, 24, 32), 16)):
        lx, ly, lt, ls = dims
        v4, v5 = lx * ly * lt, lx * ly * lt * ls
        _, l64, l32 = make_fields(G, hr.DeviceArray, dims, 0.3, 99)
        if pinned:
            G.dslash_set_config(dims, "f64", *pinned)
        else:
            G.dslash_autotune(l64, dims, "f64", stream=stream, nrhs=nrhs)
        tc, fl = G.dslash_get_config(dims, "f64", nrhs=nrhs)
        am = rng.choice([0.005, 0.01, 0.02, 0.05], size=nrhs).astype(np.float32)
        nb = max(2, int(np.ceil(4 * 256e6 / (2 * 16 * 4 * v5 * nrhs))))
        progress(f"dslash f64 {dims} nrhs={nrhs}")
        ms, _, _ = time_dslash(G, hr, dims, l64, "f64", steps, warmup, nb, stream, nrhs=nrhs,
                               am=am)
        gbs = dslash_bytes(v4, v5, "f64", nrhs) / (ms * 1e-3) / 1e9
        out["x".join(map(str, dims)) + f"_nrhs{nrhs}"] = {
            "ms": round(ms, 5), "GB_s": round(gbs, 1), "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4),
            "launch_shape": {"tchunk": tc, "flags": fl}}
        l64.free()
        l32.free()
    return out


def bench_multishift(G, hr, dims, l64, l32, am, nsrc, stream, tol=1e-10, maxcg=50000,
                     mixed_th32=1e-2, mixed_th16=1e-6,
                     legs_wanted=("plain_fp64", "blocked_fp64", "blocked_mixed")):
    """Plain FP64 path (one source at a time) vs the batched mixed path."""
    lx, ly, lt, ls = dims
    nsp = 4 * lx * ly * lt * ls
    a0, a, sig = remez_block("remez4", 1)
    n = len(sig)
    rng = np.random.default_rng(7)
    bh = rng.standard_normal((nsrc, nsp, 2)) / np.sqrt(2)
    b = hr.DeviceArray.from_host(bh)
    out = {}
    plain_x = None
    if "plain_fp64" in legs_wanted:
        leg, plain_x = _plain_leg(G, hr, dims, l64, am, bh, sig, n, nsp, tol, maxcg, stream)
        out.update(leg)
    # both batched legs verify every (source, shift) in FP64 inside the timed
    # region and correct the ones above tol (defect correction)
    legs = (("blocked_fp64", 0.0, 0.0, 2), ("blocked_mixed", mixed_th32, mixed_th16, 4))
    for label, th32, th16, corr in legs:
        if label not in legs_wanted:
            continue
        x = hr.DeviceArray((nsrc, n, nsp, 2), np.float64)
        progress(f"multishift {label} x{nsrc} sources (th32={th32}, th16={th16})")
        hr.synchronize()
        st = {}
        t = time.perf_counter()
        it, cgr, _ = G.multishift_fast_dev(l64, l32, dims, am, 3, sig, b, x, tol=tol, maxcg=maxcg,
                                           nsrc=nsrc, th32=th32, th16=th16, corrections=corr,
                                           true_residuals=False, stats=st, stream=stream)
        hr.synchronize()
        dt = time.perf_counter() - t
        rel = G.true_residuals_dev(l64, dims, am, 3, sig, b, x, nsrc=nsrc, stream=stream)
        out[label] = {"solves_per_s": round(nsrc / dt, 4), "seconds": round(dt, 3),
                      "sources": nsrc, "iterations": int(it),
                      "correction_iterations": int(st.get("corr_iterations", 0)),
                      "th32": th32, "th16": th16, "reliable_updates": st.get("reliable"),
                      "max_true_residual": float(rel.max())}
        if plain_x is not None:
            # source 0 against the plain FP64 solution, per pole
            x0 = x.head_to_host(1)[0]
            out[label]["max_pole_rel_diff_vs_plain_fp64"] = max(
                float(np.linalg.norm(x0[i] - plain_x[i]) / np.linalg.norm(plain_x[i]))
                for i in range(n))
        x.free()
    for label in ("blocked_fp64", "blocked_mixed"):
        if label in out and "plain_fp64" in out:
            out[f"speedup_{label}_vs_plain_fp64"] = round(out[label]["solves_per_s"] /
                                                          out["plain_fp64"]["solves_per_s"], 3)
    return out


def _plain_leg(G, hr, dims, l64, am, bh, sig, n, nsp, tol, maxcg, stream):
    # plain FP64: the reference's recurrences, one source
    b0 = hr.DeviceArray.from_host(bh[:1])
    xp = hr.DeviceArray((1, n, nsp, 2), np.float64)
    progress(f"multishift plain FP64 {dims} am={am}")
    hr.synchronize()
    t = time.perf_counter()
    # iterated stop at 3/4 of the target, as the fast path does, so that the
    # FP64 true residual ends below the target
    cg_plain, _ = G.multishift_plain_dev(l64, dims, am, 3, sig, b0, xp, tol=0.75 * tol, maxcg=maxcg,
                                         stream=stream)
    hr.synchronize()
    t_plain = time.perf_counter() - t
    rel_plain = G.true_residuals_dev(l64, dims, am, 3, sig, b0, xp, stream=stream)
    xh = xp.to_host()[0]
    xp.free()
    b0.free()
    return {"plain_fp64": {"solves_per_s": round(1.0 / t_plain, 4), "seconds": round(t_plain, 3),
                           "iterations": int(cg_plain),
                           "max_true_residual": float(rel_plain.max())}}, xh


def bench_derivs(G, hr, stream, steps=50, warmup=5):
    """The Shamir fermion force derivs (derivs_module.F90:99-196) on device
    fields at 16^3 x Ls=48 and 32^3 x Ls=64: time per launch from HIP events
    around `steps` launches issued back to back on the stream (and the mean
    of events around each launch, which adds the launch gaps), and GB/s of
    algorithmic bytes (R and X2 read once, dSdpi read and written once:
    2 V5 64 + 2 3 V4 8); and the force's use of it, the derivs of 25 poles
    in one launch (thr_derivs_batch_dev: 25 2 V5 64 + 2 3 V4 8 bytes)."""
    out = {}
    rng = np.random.default_rng(11)
    for dims in ((16, 16, 16, 48), (32, 32, 32, 64)):
        lx, ly, lt, ls = dims
        v4, v5 = lx * ly * lt, lx * ly * lt * ls
        nb = max(2, int(np.ceil(4 * 256e6 / (2 * 64 * v5))))  # rotate past the Infinity Cache
        rs = [hr.DeviceArray.from_host(rng.standard_normal((4 * v5, 2))) for _ in range(nb)]
        xs = [hr.DeviceArray.from_host(rng.standard_normal((4 * v5, 2))) for _ in range(nb)]
        ds = hr.DeviceArray((3 * v4,), np.float64)
        ds.zero()
        progress(f"derivs {dims}")
        for k in range(warmup):
            G.derivs_dev(rs[k % nb], xs[k % nb], 0.37, 0, dims, ds, stream=stream)
        stream.synchronize()
        ev = [(hr.Event(), hr.Event()) for _ in range(steps)]
        for k in range(steps):
            ev[k][0].record(stream)
            G.derivs_dev(rs[k % nb], xs[k % nb], 0.37, k & 1, dims, ds, stream=stream)
            ev[k][1].record(stream)
        stream.synchronize()
        per_launch = float(np.mean([a.elapsed_ms(b) for a, b in ev]))
        e0, e1 = hr.Event(), hr.Event()
        e0.record(stream)
        for k in range(steps):
            G.derivs_dev(rs[k % nb], xs[k % nb], 0.37, k & 1, dims, ds, stream=stream)
        e1.record(stream)
        stream.synchronize()
        ms = e0.elapsed_ms(e1) / steps
        nbytes = 2 * v5 * 64 + 2 * 3 * v4 * 8
        gbs = nbytes / (ms * 1e-3) / 1e9
        row = {"ms": round(ms, 5), "GB_s": round(gbs, 1),
               "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4),
               "per_launch_event_ms": round(per_launch, 5), "algorithmic_bytes": nbytes}
        for a in rs + xs:
            a.free()
        # the force's use: the derivs of 25 poles (R_k, X2_k) in one launch
        npair = 25
        pr, px = [], []
        for k in range(npair):
            for lst, seed in ((pr, 1000 + k), (px, 2000 + k)):
                a = hr.DeviceArray((4 * v5, 2), np.float64)
                G.gaussian_dev(a, 8 * v5, 1.0, seed, stream=stream)
                lst.append(a)
        an = 0.1 + rng.random(npair)
        fl = [0] * npair
        G.derivs_batch_dev(pr, px, an, fl, dims, ds, stream=stream)
        stream.synchronize()
        e0, e1 = hr.Event(), hr.Event()
        nrep = 10
        e0.record(stream)
        for k in range(nrep):
            G.derivs_batch_dev(pr, px, an, fl, dims, ds, stream=stream)
        e1.record(stream)
        stream.synchronize()
        bms = e0.elapsed_ms(e1) / nrep
        bbytes = npair * 2 * v5 * 64 + 2 * 3 * v4 * 8
        bgbs = bbytes / (bms * 1e-3) / 1e9
        row["batched_25_poles"] = {"ms": round(bms, 4), "GB_s": round(bgbs, 1),
                                   "frac_of_peak": round(bgbs / HBM_PEAK_GBS, 4),
                                   "algorithmic_bytes": bbytes}
        out["x".join(map(str, dims))] = row
        for a in pr + px + [ds]:
            a.free()
    return out


def bench_batched_config4(G, hr, stream, dims=(24, 24, 24, 32), nsrc=16, tol=1e-10,
                          maxcg=50000, seed=24):
    """The batched multi-RHS workload: 16 sources x 25 shifts of x^(-1/4)
    (remez4 block 2) at 24^3 x Ls=32, one link field, per-source masses drawn
    from {0.005, 0.01, 0.02, 0.05}: the batched solver (FP32 Krylov storage
    from the start, or after |r|/|b| < 1e-2 or 1e-4; reliable updates,
    hipGraph-captured blocks, FP64 defect correction) against
    the plain FP64 path (the reference's recurrences, one source at a time,
    each with its own mass); solves/s to FP64 true residual `tol` for both."""
    lx, ly, lt, ls = dims
    nsp = 4 * lx * ly * lt * ls
    _, l64, l32 = make_fields(G, hr.DeviceArray, dims, 0.3, seed)
    a0, a, sig = remez_block("remez4", 1)
    n = len(sig)
    rng = np.random.default_rng(seed)
    masses = rng.choice([0.005, 0.01, 0.02, 0.05], size=nsrc).astype(np.float32)
    bh = rng.standard_normal((nsrc, nsp, 2)) / np.sqrt(2)
    b = hr.DeviceArray.from_host(bh)
    out = {"lattice": list(dims), "sources": nsrc, "shifts": n, "masses": masses.tolist()}
    # precision schedules of the Krylov storage: FP32 from the first block,
    # and FP64 until |r|/|b| < 1e-2 or 1e-4 then FP32; all with reliable
    # updates of the base residual (the library default) and the FP64 defect
    # correction of every (source, shift) above tol
    scheds = (("batched_fp32_from_start", 1.0), ("batched_fp64_then_fp32", 1e-2),
              ("batched_fp64_to_1e-4_then_fp32", 1e-4))
    for label, th32 in scheds:
        x = hr.DeviceArray((nsrc, n, nsp, 2), np.float64)
        progress(f"config-4 batched solve {dims}: {nsrc} sources x {n} shifts, th32={th32}")
        hr.synchronize()
        st = {}
        t = time.perf_counter()
        it, _, _ = G.multishift_fast_dev(l64, l32, dims, masses, 3, sig, b, x, tol=tol,
                                         maxcg=maxcg, nsrc=nsrc, th32=th32,
                                         true_residuals=False, corrections=4, stats=st,
                                         stream=stream)
        hr.synchronize()
        dt = time.perf_counter() - t
        rel = G.true_residuals_dev(l64, dims, masses, 3, sig, b, x, nsrc=nsrc, stream=stream)
        out[label] = {"solves_per_s": round(nsrc / dt, 4), "seconds": round(dt, 3),
                      "th32": th32, "iterations": int(it),
                      "correction_iterations": int(st.get("corr_iterations", 0)),
                      "reliable_updates": st.get("reliable"),
                      "max_true_residual": float(rel.max())}
        x.free()
    progress(f"config-4 plain FP64, {nsrc} sources one at a time")
    xp = hr.DeviceArray((1, n, nsp, 2), np.float64)
    worst, total, its = 0.0, 0.0, []
    for k in range(nsrc):
        bk = hr.DeviceArray.from_host(bh[k:k + 1])
        hr.synchronize()
        t = time.perf_counter()
        cg, _ = G.multishift_plain_dev(l64, dims, masses[k], 3, sig, bk, xp, tol=0.75 * tol,
                                       maxcg=maxcg, stream=stream)
        hr.synchronize()
        total += time.perf_counter() - t
        its.append(int(cg))
        worst = max(worst, float(G.true_residuals_dev(l64, dims, masses[k], 3, sig, bk, xp,
                                                      stream=stream).max()))
        bk.free()
    xp.free()
    out["plain_fp64"] = {"solves_per_s": round(nsrc / total, 4), "seconds": round(total, 3),
                         "iterations": its, "max_true_residual": worst}
    for label, _ in scheds:
        out[f"speedup_{label}_vs_plain_fp64"] = round(out[label]["solves_per_s"]
                                                      / out["plain_fp64"]["solves_per_s"], 3)
    b.free()
    l64.free()
    l32.free()
    return out


def bench_small_lattices(G, hr, stream, nsrc=4, tol=1e-10):
    """Launch-bound sizes of the serving mix: the batched solve of
    (M^+M) x = b (the serving request's system, 4 sources, am 0.05) at
    8^3 x Ls=8 and 12^2 x 24 x Ls=16 by the persistent cooperative kernel
    (THR_PERSIST=1, an option) and by the streaming solver (the default) with
    its steady-state blocks replayed from hipGraphs and launched eagerly:
    wall time per solve and per iteration."""
    out = {}
    keys = ("THR_GRAPHS", "THR_PERSIST")
    saved = {k: os.environ.get(k) for k in keys}
    rng = np.random.default_rng(5)
    modes = [("persistent", {"THR_PERSIST": "1"}),
             ("graphs", {"THR_PERSIST": "0", "THR_GRAPHS": "1"}),
             ("eager", {"THR_PERSIST": "0", "THR_GRAPHS": "0"})]
    if under_profiler():  # no graph launches under the profiler
        modes = [m for m in modes if m[0] != "graphs"]
    try:
        for dims in ((8, 8, 8, 8), (12, 12, 24, 16)):
            lx, ly, lt, ls = dims
            nsp = 4 * lx * ly * lt * ls
            _, l64, l32 = make_fields(G, hr.DeviceArray, dims, 0.3, 17)
            b = hr.DeviceArray.from_host(rng.standard_normal((nsrc, nsp, 2)) / np.sqrt(2))
            x = hr.DeviceArray((nsrc, 1, nsp, 2), np.float64)
            row = {}
            for label, env in modes:
                os.environ.update(env)
                progress(f"small lattice {dims} {label}")
                G.multishift_fast_dev(l64, l32, dims, 0.05, 3, np.zeros(1), b, x, tol=tol,
                                      maxcg=100000, nsrc=nsrc, th32=1e-4, true_residuals=False,
                                      stream=stream)  # warm: workspace, graph capture
                hr.synchronize()
                t = time.perf_counter()
                _, cgr, _ = G.multishift_fast_dev(l64, l32, dims, 0.05, 3, np.zeros(1), b, x,
                                                  tol=tol, maxcg=100000, nsrc=nsrc, th32=1e-4,
                                                  true_residuals=False, stream=stream)
                hr.synchronize()
                dt = time.perf_counter() - t
                it = int(max(cgr))
                rel = G.true_residuals_dev(l64, dims, 0.05, 3, np.zeros(1), b, x, nsrc=nsrc,
                                           stream=stream)
                row[label] = {"seconds": round(dt, 5), "iterations": it,
                              "us_per_iteration": round(1e6 * dt / max(1, it), 2),
                              "max_true_residual": float(rel.max())}
            if "graphs" in row:
                row["speedup_persistent_vs_graphs"] = round(row["graphs"]["seconds"]
                                                            / row["persistent"]["seconds"], 3)
            out["x".join(map(str, dims))] = row
            for a in (b, x, l64, l32):
                a.free()
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return out


def bench_config3_errors(G, dims=(32, 32, 32, 64), am=0.005, beta=0.3, seed=303):
    """Mixed precision against FP64 at 32^3 x Ls=64 (am 0.005) for the RHMC
    force (force_module.F90:9-113, res1 = 1e-8) and the Hamiltonian
    (dwf3d_lib.F90:511-616, res2 = 1e-10) at two gauge fields a small MD
    step apart: relative L2 difference of dSdpi, relative difference of the
    fermion action hf and the absolute difference of dH, with times."""
    import md
    lx, ly, lt, ls = dims
    rng = np.random.default_rng(seed)
    th0 = np.asfortranarray((rng.standard_normal((lx, ly, lt, 3)) / np.sqrt(beta))
                            .astype(np.float32))
    phi = np.zeros((ls + 2, lx + 2, ly + 2, lt + 2, 4), dtype=np.complex128, order="F")
    phi[1:-1, 1:-1, 1:-1, 1:-1] = (rng.standard_normal((ls, lx, ly, lt, 4))
                                   + 1j * rng.standard_normal((ls, lx, ly, lt, 4))) / np.sqrt(2)
    G.halo5(phi)
    pp = np.asfortranarray(rng.standard_normal(th0.shape).astype(np.float32))
    th1 = np.asfortranarray((th0 + np.float32(0.02) * pp).astype(np.float32))
    u0, u1 = G.coef(th0, dims), G.coef(th1, dims)
    cs = md.coefficient_sets()
    out = {"lattice": list(dims), "am": am, "res1": 1e-8, "res2": 1e-10}
    forces, ham = {}, {}
    for name, sp in (("fp64", False), ("mixed", "mixed")):
        progress(f"config-3 force and hamilton, {name}, {dims}")
        t = time.perf_counter()
        forces[name], _ = G.force(phi, 1e-8, am, 3, beta, th0, u0, dims, cs["force"], use_sp=sp)
        tf = time.perf_counter() - t
        t = time.perf_counter()
        ham[name] = [G.hamilton(phi, 1e-10, am, 3, beta, th, pp, u, dims, cs["hamilton"],
                                use_sp=sp)[0] for th, u in ((th0, u0), (th1, u1))]
        out[name] = {"force_seconds": round(tf, 3),
                     "hamilton_seconds": round((time.perf_counter() - t) / 2, 3)}
    f64 = forces["fp64"]
    out["force_mixed_vs_fp64_rel_l2"] = float(np.linalg.norm(forces["mixed"] - f64)
                                              / np.linalg.norm(f64))
    out["hf_mixed_vs_fp64_rel"] = max(abs(ham["mixed"][k]["hf"] / ham["fp64"][k]["hf"] - 1)
                                      for k in (0, 1))
    dh = {k: v[0]["h"] - v[1]["h"] for k, v in ham.items()}
    out["dH_fp64"] = dh["fp64"]
    out["dH_mixed"] = dh["mixed"]
    out["dH_abs_diff"] = abs(dh["mixed"] - dh["fp64"])
    return out


def bench_force(G, dims=(16, 16, 16, 48), am=0.01, beta=0.3, res1=1e-8, seed=77):
    """force_module.F90 at the production lattice (16^3 x Ls=48, am 0.01,
    beta 0.3, 12-pole guidance approximations): the reference's FP64 solver
    against the mixed-precision solver with FP64 correction (use_sp =
    "mixed"), both on this GPU through the host-array entry point; reports
    seconds per force and the relative L2 difference of dSdpi."""
    import confio
    lx, ly, lt, ls = dims
    rng = np.random.default_rng(seed)
    th = np.asfortranarray((rng.standard_normal((lx, ly, lt, 3)) / np.sqrt(beta))
                           .astype(np.float32))
    u = G.coef(th, dims)
    phi = np.zeros((ls + 2, lx + 2, ly + 2, lt + 2, 4), dtype=np.complex128, order="F")
    phi[1:-1, 1:-1, 1:-1, 1:-1] = (rng.standard_normal((ls, lx, ly, lt, 4))
                                   + 1j * rng.standard_normal((ls, lx, ly, lt, 4))) / np.sqrt(2)
    G.halo5(phi)
    (a4g, d4g), _ = confio.remez("remez4g")
    (a2g, d2g), (b2g, e2g) = confio.remez("remez2g")
    coeffs = (a4g, d4g, b2g, e2g, a2g, d2g)
    out = {"lattice": list(dims), "am": am, "res1": res1}
    forces = {}
    for name, sp in (("fp64", False), ("mixed", "mixed")):
        progress(f"force {name} {dims}")
        t0 = time.perf_counter()
        f, its = G.force(phi, res1, am, 3, beta, th, u, dims, coeffs, use_sp=sp)
        out[name] = {"seconds": round(time.perf_counter() - t0, 3), "itercg": [int(i) for i in its]}
        forces[name] = f
    out["mixed_vs_fp64_rel_l2"] = float(np.linalg.norm(forces["mixed"] - forces["fp64"])
                                        / np.linalg.norm(forces["fp64"]))
    out["speedup"] = round(out["fp64"]["seconds"] / out["mixed"]["seconds"], 3)
    return out


def bench_serving(rates, seconds, slots):
    """Poisson load of measurement solves (L in {8,12,16,24}, T in {L,2L}, Ls in
    {8,16,32,48}, beta in [0.2,0.5], am in {0.005..0.05}, 1-12 sources, FP64
    true residual 1e-10), continuous batching over `slots` contexts, at each
    rate of the sweep: p50/p99 latency, solves/s, and the requests shed when
    the arrival window closed; the highest sustainable rate is the largest one
    that shed and failed nothing."""
    import serving as S
    return S.rate_sweep(rates, seconds, seed=3, slots=slots, progress=progress)


KERNEL_SOURCES = ("thirring-rhmc_amd/csrc/dslash.hip", "thirring-rhmc_amd/csrc/dslash_site.h",
                  "thirring-rhmc_amd/csrc/kernels.h", "thirring-rhmc_amd/csrc/common.h")


def kernel_source_digest():
    """sha256 over the Dslash kernel's sources (what the counters measured)."""
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def measured_traffic():
    """HBM bytes per FP64 Dslash launch from the committed rocprofv3 counter
    summary, only if it was taken on the current kernel sources (else null:
    a changed kernel needs new counters)."""
    tpath = os.path.join(ROOT, "profiles", "dslash_f64_32x32x32x64_traffic.json")
    if not os.path.exists(tpath):
        return None, "no counter summary"
    with open(tpath) as fh:
        rec = json.load(fh)
    if rec.get("kernel_source_sha256") != kernel_source_digest():
        return None, "counter summary predates the current Dslash sources"
    return rec.get("hbm_bytes_per_launch"), rec.get("source")


def traffic_shape():
    """(tchunk, flags) of the FP64 launch shape the counter summary profiled."""
    tpath = os.path.join(ROOT, "profiles", "dslash_f64_32x32x32x64_traffic.json")
    if not os.path.exists(tpath):
        return None
    with open(tpath) as fh:
        kernel = json.load(fh).get("kernel", "")
    import re
    m = re.search(r"f64:(\d+):(\d+)", kernel)
    return (int(m.group(1)), int(m.group(2))) if m else None


def copy_rate(G, hr, stream, nbytes=1 << 30, reps=20):
    """Streaming device-to-device copy rate (bytes read + written per second)
    on this GPU, with the library's 16 B/lane non-temporal copy kernel: the
    practical HBM ceiling the Dslash is compared with."""
    a = hr.DeviceArray((nbytes // 8,), np.float64)
    a.zero()
    b = hr.DeviceArray((nbytes // 8,), np.float64)
    for _ in range(3):
        G.device_copy(b, a, nbytes, stream=stream)
    e0, e1 = hr.Event(), hr.Event()
    e0.record(stream)
    for _ in range(reps):
        G.device_copy(b, a, nbytes, stream=stream)
    e1.record(stream)
    ms = e0.elapsed_ms(e1) / reps
    a.free()
    b.free()
    return 2 * nbytes / (ms * 1e-3) / 1e9


def cpu_baseline(dims, seconds=12.0):
    """Oracle (own FP64 C restatement, OpenMP) Dslash on the host cores."""
    import dense as D
    import oracle as O
    ncore = min(os.cpu_count() or 1, 16)
    O.set_threads(ncore)
    lx, ly, lt, ls = dims
    rng = np.random.default_rng(2)
    th = (rng.standard_normal((lx, ly, lt, 3)) / np.sqrt(0.3)).astype(np.float32)
    u = O.coef(th, dims)
    r = D.rand_spinor(dims, rng)
    O.dslash(r, u, dims, 0.005, 3)  # warm
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        O.dslash(r, u, dims, 0.005, 3) if n % 2 == 0 else O.dslashd(r, u, dims, 0.005, 3)
        n += 1
    dt = (time.perf_counter() - t0) / n
    v4 = lx * ly * lt
    gbs = dslash_bytes(v4, v4 * ls, "f64") / dt / 1e9
    return {"value": round(gbs, 3), "unit": "GB/s", "cores": ncore, "kind": "port",
            "sample": f"{n} oracle Dslash applications (M and M^dagger alternating) at "
                      f"{lx}x{ly}x{lt}x{ls}, {dt * 1e3:.1f} ms each"}


def under_profiler():
    """True when a rocprofv3 tool library is preloaded into this process."""
    pre = os.environ.get("LD_PRELOAD", "")
    return "rocprof" in pre or any(k.startswith("ROCPROF") for k in os.environ)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--lattice", default="32,32,32,64")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solver", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-multi-rhs", action="store_true")
    ap.add_argument("--no-force", action="store_true")
    ap.add_argument("--no-config4", action="store_true")
    ap.add_argument("--no-derivs", action="store_true")
    ap.add_argument("--no-small", action="store_true")
    ap.add_argument("--no-config3-errors", action="store_true")
    ap.add_argument("--serve-rates", default="1,2,5,10,20,50,100,200",
                    help="Poisson arrival rates (per s) of the serving sweep")
    ap.add_argument("--serve-seconds", type=float, default=6.0,
                    help="arrival window per rate (requests still queued at its end are shed)")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--solver-lattice", default="32,32,32,64")
    ap.add_argument("--solver-am", type=float, default=0.005)
    ap.add_argument("--solver-sources", type=int, default=4)
    ap.add_argument("--mixed-th32", type=float, default=1e-2)
    ap.add_argument("--mixed-th16", type=float, default=1e-6)
    ap.add_argument("--solver-legs", default="plain_fp64,blocked_fp64,blocked_mixed",
                    help="comma-separated subset of the multishift legs to run")
    ap.add_argument("--tchunk", type=int, default=0,
                    help="pin the Dslash t-chunk (with --shape-flags) instead of autotuning")
    ap.add_argument("--shape-flags", type=int, default=7,
                    help="with --tchunk: 1 XCD remap | 2 uniform links | 4 lane-shift s hops | "
                         "8 XCD boxes | 16 LDS x exchange | 32 non-temporal stores | "
                         "k << 6 at most k + 1 resident blocks per CU")
    ap.add_argument("--shapes", default="",
                    help="pin per-precision shapes instead of autotuning: "
                         "prec:tchunk:flags,... (e.g. f64:16:55,f32:2:55,f16:2:39)")
    ap.add_argument("--solver-shapes", default="",
                    help="pin the solver lattice's shapes (prec:tchunk:flags,...) instead of "
                         "autotuning them for the solver's batch")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run N device copies of 1 GiB (scales the HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); it measures the "
              "single-GPU kernels at their real shape (independent configurations and sources "
              "shard one process per GPU)", file=sys.stderr)
        sys.exit(2)

    if under_profiler():
        # graph replays are never profiled (the profiler's interception of
        # hipGraph launches once aborted the host process): launch eagerly
        os.environ["THR_GRAPHS"] = "0"
    import hiprt as hr
    import thirring_hip as G
    G.init(0)
    if args.copy_probe:
        a = hr.DeviceArray((1 << 27,), np.float64)
        a.zero()
        b = hr.DeviceArray((1 << 27,), np.float64)
        for _ in range(args.copy_probe):
            b.copy_from(a)
        hr.synchronize()
        print(json.dumps({"copy_probe_bytes_read": a.nbytes, "copies": args.copy_probe}))
        return
    stream = hr.Stream()
    dims = tuple(int(v) for v in args.lattice.split(","))
    lx, ly, lt, ls = dims
    v4, v5 = lx * ly * lt, lx * ly * lt * ls
    th, l64, l32 = make_fields(G, hr.DeviceArray, dims, 0.3, 1234)

    t_wall0 = time.time()
    shapes = {}
    pins = {}
    for item in filter(None, args.shapes.split(",")):
        p, tc, fl = item.split(":")
        pins[p] = (int(tc), int(fl))
    for prec, lk in (("f64", l64), ("f32", l32), ("f16", l32)):
        if prec in pins:
            G.dslash_set_config(dims, prec, *pins[prec])
        elif args.tchunk > 0:
            G.dslash_set_config(dims, prec, args.tchunk, args.shape_flags)
        else:
            progress(f"autotune dslash {prec}")
            G.dslash_autotune(lk, dims, prec, stream=stream)
        tc, fl = G.dslash_get_config(dims, prec)
        shapes[prec] = {"tchunk": tc, "flags": fl}
    nbuf64 = max(2, int(np.ceil(4 * 256e6 / (2 * 16 * 4 * v5))))
    progress(f"dslash f64 {dims} x{args.steps}")
    ms64, ev64, region64 = time_dslash(G, hr, dims, l64, "f64", args.steps, args.warmup, nbuf64,
                                       stream)
    gbs64 = dslash_bytes(v4, v5, "f64") / (ms64 * 1e-3) / 1e9
    extra = {}
    for prec in ("f32", "f16"):
        nb = max(2, int(np.ceil(4 * 256e6 / (2 * {"f32": 8, "f16": 5}[prec] * 4 * v5))))
        progress(f"dslash {prec}")
        ms, _, _ = time_dslash(G, hr, dims, l32, prec, args.steps, args.warmup, nb, stream)
        extra[prec] = {"ms": round(ms, 5),
                       "GB_s": round(dslash_bytes(v4, v5, prec) / (ms * 1e-3) / 1e9, 1)}
    pinned = (args.tchunk, args.shape_flags) if args.tchunk > 0 else None
    multi = None if args.no_multi_rhs else multi_rhs_legs(G, hr, max(20, args.steps // 4),
                                                           args.warmup, stream, pinned)
    progress("device copy rate")
    copy_gbs = copy_rate(G, hr, stream)
    solver = None
    if not args.no_solver:
        sd = tuple(int(v) for v in args.solver_lattice.split(","))
        _, s64, s32 = make_fields(G, hr.DeviceArray, sd, 0.3, 4321)
        # the solver launches every pass over all its sources: pin the shapes
        # tuned for that batch
        batch_shapes = {}
        spins = {}
        for item in filter(None, args.solver_shapes.split(",")):
            p, tc, fl = item.split(":")
            spins[p] = (int(tc), int(fl))
        for prec, lk in (("f64", s64), ("f32", s32), ("f16", s32)):
            if prec in spins:
                G.dslash_set_config(sd, prec, *spins[prec])
            elif prec in pins:
                G.dslash_set_config(sd, prec, *pins[prec])
            elif args.tchunk > 0:
                G.dslash_set_config(sd, prec, args.tchunk, args.shape_flags)
            elif args.solver_sources > 1:
                progress(f"autotune dslash {prec} nrhs={args.solver_sources}")
                G.dslash_autotune(lk, sd, prec, stream=stream, nrhs=args.solver_sources)
            tc, fl = G.dslash_get_config(sd, prec, nrhs=args.solver_sources)
            batch_shapes[prec] = {"tchunk": tc, "flags": fl}
        solver = bench_multishift(G, hr, sd, s64, s32, args.solver_am, args.solver_sources, stream,
                                  mixed_th32=args.mixed_th32, mixed_th16=args.mixed_th16,
                                  legs_wanted=tuple(args.solver_legs.split(",")))
        solver["lattice"] = list(sd)
        solver["launch_shapes"] = batch_shapes
        solver["am"] = args.solver_am
    derivs = None if args.no_derivs else bench_derivs(G, hr, stream)
    config4 = None if args.no_config4 else bench_batched_config4(G, hr, stream)
    small = None if args.no_small else bench_small_lattices(G, hr, stream)
    force = None if args.no_force else bench_force(G)
    config3 = None if args.no_config3_errors else bench_config3_errors(G)
    serving = None
    if not args.no_serving:
        serving = bench_serving([float(r) for r in args.serve_rates.split(",")],
                                args.serve_seconds, args.serve_slots)
    wall = time.time() - t_wall0

    traffic, traffic_note = measured_traffic()
    # the counters were taken at one shape; the timed shape is the autotuner's
    # pick of this run, which may differ (bytes moved depend on the shape)
    tshape = traffic_shape()
    timed_shape = (shapes["f64"]["tchunk"], shapes["f64"]["flags"])

    progress("cpu baseline")
    cpu = None if args.no_cpu else cpu_baseline(dims, args.cpu_seconds)

    line = {
        "metric": "Dslash HBM GB/s (FP64 Shamir DWF, 32^3 x Ls=64)",
        "value": round(gbs64, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms64, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: theta ~ N(0, 1/sqrt(0.3)) float32, u = 1 + i theta (antiperiodic t), "
                "unit complex Gaussian sources",
        "config": {"workload": "dslash_shamir 32x32x32 Ls=64 imass=3 am=0.005 am3=1 (M and "
                               "M^dagger alternating), rotating buffers beyond the Infinity Cache",
                   "lattice": list(dims), "rotating_buffer_sets": nbuf64,
                   "launch_shape": shapes,
                   "per_launch_event_ms": round(ev64, 5),
                   "event_bracketed_region_ms_per_step": round(region64, 5)},
        "roofline": {"bound": "hbm", "achieved": round(gbs64, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs64 / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_note,
                     "traffic_shape": ({"tchunk": tshape[0], "flags": tshape[1],
                                        "same_as_timed": tshape == timed_shape}
                                       if tshape else None)},
        "hbm_copy_GB_s": round(copy_gbs, 1),
        "frac_of_copy": round(gbs64 / copy_gbs, 4),
        "cpu_baseline": cpu,
        "dslash_other_precisions": extra,
        "dslash_multi_rhs": multi,
        "derivs": derivs,
        "multishift": solver,
        "batched_multi_rhs_config4": config4,
        "small_lattice_solves": small,
        "force": force,
        "mixed_vs_fp64_32x32x32x64": config3,
        "serving": serving,
        "bench_wall_s": round(wall, 2),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
