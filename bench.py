"""Benchmark of the domain-wall hot path on one MI355X.

Headline workload: the Shamir Dslash (imass 3, am3 = 1) at 32^3 x Ls=64,
FP64 (the mixed-precision production shape), on synthetic U(1) links
u = 1 + i theta, theta ~ N(0, 1/sqrt(beta)) with beta = 0.3, and unit-variance
complex Gaussian sources.  One step = one application of M (alternating with
M^dagger) to a fresh source; inputs rotate over enough buffer sets to exceed
the 256 MB Infinity Cache, so the figure is HBM bandwidth, not cache.

value = algorithmic bytes per application / average kernel duration, with
algorithmic bytes = 2 * V5 * 4 * 16 + 3 * V4 * 16 (each input read once, the
output written once).  Kernel durations come from HIP events recorded on the
stream the kernels are launched on.

Also reported: FP32 and FP16 operator bandwidth, the multishift solve rates
(plain FP64 vs the mixed-precision path) and the oracle CPU baseline.

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


def dslash_bytes(v4, v5, prec, nrhs=1):
    b_c = {"f64": 16, "f32": 8, "f16": 5}[prec]
    b_link = 16 if prec == "f64" else 8
    return nrhs * 2 * v5 * 4 * b_c + 3 * v4 * b_link


def make_fields(dims, beta, seed, torch, dev):
    import thirring_hip as G
    lx, ly, lt, ls = dims
    rng = np.random.default_rng(seed)
    th = (rng.standard_normal((lx, ly, lt, 3)) / np.sqrt(beta)).astype(np.float32)
    theta = torch.from_numpy(th.reshape(-1, order="F")).to(dev)
    l64 = torch.empty(lx * ly * lt * 3, 2, dtype=torch.float64, device=dev)
    l32 = torch.empty(lx * ly * lt * 3, 2, dtype=torch.float32, device=dev)
    G.coef_dev(l64, theta, dims)
    G.links_to_f32_dev(l32, l64, dims)
    return th, l64, l32


def time_dslash(G, torch, dims, links, prec, steps, warmup, nbuf, stream):
    lx, ly, lt, ls = dims
    v5 = lx * ly * lt * ls
    nsp = 4 * v5
    gen = torch.Generator(device="cuda").manual_seed(1)
    if prec == "f64":
        mk = lambda: torch.randn(nsp, 2, dtype=torch.float64, device="cuda", generator=gen) / np.sqrt(2)
    else:
        mk = lambda: (torch.randn(nsp, 2, dtype=torch.float32, device="cuda", generator=gen) / np.sqrt(2))
    ins, outs, inn, outn = [], [], [], []
    for _ in range(nbuf):
        x = mk()
        if prec == "f16":
            sc = x.view(4, v5, 2).abs().amax(dim=(0, 2)).contiguous()
            x = (x.view(4, v5, 2) / sc[None, :, None]).half().contiguous().view(nsp, 2)
            inn.append(sc)
            outn.append(torch.empty_like(sc))
        else:
            inn.append(None)
            outn.append(None)
        ins.append(x)
        outs.append(torch.empty_like(x))

    def launch(k):
        j = k % nbuf
        G.dslash_dev(outs[j], ins[j], links, dims, 0.005, imass=3, am3=1.0, dagger=bool(k & 1),
                     prec=prec, in_norm=inn[j], out_norm=outn[j], stream=stream)

    for k in range(warmup):
        launch(k)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    t_all0 = torch.cuda.Event(enable_timing=True)
    t_all1 = torch.cuda.Event(enable_timing=True)
    t_all0.record(stream)
    for k in range(steps):
        ev[k][0].record(stream)
        launch(k)
        ev[k][1].record(stream)
    t_all1.record(stream)
    torch.cuda.synchronize()
    per = np.array([a.elapsed_time(b) for a, b in ev])  # ms
    region = t_all0.elapsed_time(t_all1) / steps
    return float(np.mean(per)), float(region)


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--lattice", default="32,32,32,64")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run N device copies of 1 GiB (scales the HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this build measures "
              "the single-GPU kernel at its real shape", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_probe:
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(args.copy_probe):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": a.numel() * 8, "copies": args.copy_probe}))
        return
    import thirring_hip as G
    G.init(0)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    dims = tuple(int(v) for v in args.lattice.split(","))
    lx, ly, lt, ls = dims
    v4, v5 = lx * ly * lt, lx * ly * lt * ls
    th, l64, l32 = make_fields(dims, 0.3, 1234, torch, dev)

    t_wall0 = time.time()
    # enough rotating buffer sets to exceed the 256 MB Infinity Cache
    nbuf64 = max(2, int(np.ceil(4 * 256e6 / (2 * 16 * 4 * v5))))
    ms64, region64 = time_dslash(G, torch, dims, l64, "f64", args.steps, args.warmup, nbuf64, stream)
    b64 = dslash_bytes(v4, v5, "f64")
    gbs64 = b64 / (ms64 * 1e-3) / 1e9
    extra = {}
    for prec, links in (("f32", l32), ("f16", l32)):
        nb = max(2, int(np.ceil(4 * 256e6 / (2 * {"f32": 8, "f16": 5}[prec] * 4 * v5))))
        ms, reg = time_dslash(G, torch, dims, links, prec, args.steps, args.warmup, nb, stream)
        extra[prec] = {"ms": round(ms, 5), "GB_s": round(dslash_bytes(v4, v5, prec) / (ms * 1e-3) / 1e9, 1)}
    wall = time.time() - t_wall0

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "dslash_f64_32x32x32x64_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")

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
                   "region_ms_per_step": round(region64, 5)},
        "roofline": {"bound": "hbm", "achieved": round(gbs64, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs64 / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
        "dslash_other_precisions": extra,
        "bench_wall_s": round(wall, 2),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
