"""Writes tests/golden/config2_congrad.npz: the oracle's congrad (the C
restatement of measure_module.F90:9-140) at the production lattice
16^3 x Ls=48, where one oracle solve to res = 1e-10 takes tens of thousands of
iterations (about a quarter of an hour on 8 host cores) -- too long to run
inside a GPU test, so the test compares against this fixture instead.

Inputs are re-created by the test from the same seeds (numpy's default
generator, identical on every machine with this numpy):
  theta(16,16,16,3) float32 ~ N(0, 1/sqrt(beta)), beta = 0.3, seed 202;
  u = coef(theta) with ibound = -1; Phi = unit-variance complex Gaussian
  (tests/dense.py rand_spinor) from the same generator; am = 0.01, imass = 3,
  am3 = 1, res = 1e-10, initial guess X = 0, niterc = Ls V4 (params.F90:57).

Stored: itercg and iterations, the host threads used (the oracle's OpenMP
sums depend on them), |X|^2, X at 4096 fixed interior entries, and the true
residual |Phi - M^+M X| / |Phi|.

Run: OMP_NUM_THREADS=8 python tests/golden/make_config2_congrad.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)

import dense as D  # noqa: E402
import oracle as O  # noqa: E402

DIMS = (16, 16, 16, 48)
SEED = 202
AM, IMASS, RES = 0.01, 3, 1e-10
NSAMPLE = 4096


def inputs():
    lx, ly, lt, ls = DIMS
    rng = np.random.default_rng(SEED)
    th = (rng.standard_normal((lx, ly, lt, 3)) / np.sqrt(0.3)).astype(np.float32)
    phi = D.rand_spinor(DIMS, rng)
    return th, phi


def sample_index():
    """Flat indices into the interior vector (Fortran order of (Ls,Lx,Ly,Lt,4))."""
    lx, ly, lt, ls = DIMS
    n = ls * lx * ly * lt * 4
    return np.sort(np.random.default_rng(99).choice(n, NSAMPLE, replace=False))


def main():
    th, phi = inputs()
    u = O.coef(th, DIMS)
    x = np.zeros_like(phi)
    t0 = time.time()
    itercg, iterations = O.congrad(phi, x, u, DIMS, res=RES, am=AM, imass=IMASS)
    secs = time.time() - t0
    xv = D.interior_vec(x)
    r = D.interior_vec(phi) - D.interior_vec(O.dirac_op_shifted(x, u, DIMS, AM, IMASS))
    relres = float(np.linalg.norm(r) / np.linalg.norm(D.interior_vec(phi)))
    idx = sample_index()
    out = os.path.join(HERE, "config2_congrad.npz")
    np.savez_compressed(out, itercg=itercg, iterations=iterations, threads=O.num_threads(),
                        xnorm2=float(np.vdot(xv, xv).real), index=idx, x=xv[idx],
                        true_residual=relres, seconds=secs)
    print(f"itercg {itercg} iterations {iterations} true residual {relres:.3e} "
          f"({secs:.0f} s on {O.num_threads()} threads) -> {out}")


if __name__ == "__main__":
    main()
