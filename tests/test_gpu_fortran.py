"""The Fortran drop-in boundary end to end: reference-style Fortran programs
(tests/fortran/*.F90, built against thirring-rhmc_amd/fortran/shims.F90)
call dslash/dslashd, congrad, qmrherm and derivs with the reference's
subroutine signatures and haloed arrays; their outputs are compared with the
oracle on the same closed-form inputs (tests/test_utils.F90:35-151).

Tolerances: dslash relative L2 <= 1e-13; congrad itercg == 27 and X relative
L2 <= 1e-10; qmrherm itercg == 2 (max_qmr_iters = 2) and x, Phi0, dSdpi
relative L2 <= 1e-10; derivs <= 1e-12.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import refstate

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "thirring-rhmc_amd", "fortran", "_build")
DIMS = (12, 12, 12, 24)
TAU = 8.0 * np.arctan(1.0)


def run(name, outdir):
    exe = os.path.join(BIN, name)
    assert os.path.exists(exe), f"{exe} not built (run __graft_entry__.build())"
    p = subprocess.run([exe, str(outdir)], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    return p.stdout


def load(path, shape, dtype=np.complex128):
    return np.fromfile(path, dtype=dtype).reshape(shape, order="F")


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def check(a, b, tol):
    e = rel(a, b)
    if not e <= tol:
        raise AssertionError(f"relative L2 error {e:.3e} > {tol:.1e}")


def test_fortran_dslash(tmp_path):
    run("test_dslash_drop_in", tmp_path)
    st = refstate.starting_state(DIMS, want=("r", "u"))
    shape = (24, 12, 12, 12, 4)
    for imass in (1, 3, 5):
        g = load(tmp_path / f"dslash_M_{imass}.bin", shape)
        o = O.interior(O.dslash(st["r"], st["u"], DIMS, 0.05, imass))
        check(g, o, 1e-13)
        g = load(tmp_path / f"dslash_Md_{imass}.bin", shape)
        o = O.interior(O.dslashd(st["r"], st["u"], DIMS, 0.05, imass))
        check(g, o, 1e-13)
    # inverter_utils (dirac_op_shifted with the default-real shift 0.37, and
    # dirac_operator) through the Fortran module
    for name, shift in (("dirac_op_shifted", float(np.float32(0.37))), ("dirac_operator", 0.0)):
        g = load(tmp_path / f"{name}.bin", shape)
        o = O.interior(O.dirac_op_shifted(st["r"], st["u"], DIMS, 0.05, 3, shift=shift))
        check(g, o, 1e-13)
def test_fortran_congrad(tmp_path):
    out = run("test_congrad_drop_in", tmp_path)
    assert "PASS itercg" in out and "27" in out, out
    # the oracle on the Fortran program's own inputs (its exp() may differ from
    # numpy's in the last bit, which CG amplifies)
    phi = load(tmp_path / "congrad_phi.bin", (26, 14, 14, 14, 4))
    x = load(tmp_path / "congrad_x0.bin", (26, 14, 14, 14, 4))
    u = load(tmp_path / "congrad_u.bin", (14, 14, 14, 3))
    it, _ = O.congrad(phi, x, u, DIMS, res=0.1, am=0.05, imass=3)
    assert it == 27
    g = load(tmp_path / "congrad_x.bin", (24, 12, 12, 12, 4))
    check(g, O.interior(x), 1e-10)
def test_fortran_qmrherm_and_derivs(tmp_path):
    out = run("test_qmrherm_drop_in", tmp_path)
    assert out.count("PASS itercg") == 4, out
    st = refstate.starting_state(DIMS, want=("phi", "r", "x", "u", "phi0", "dsdpi"))
    i = np.arange(1, 26)
    anum = np.concatenate([[0.5], np.cos(i * TAU / 25)])
    aden = np.cos(0.5 * i * TAU / 25)
    # Phi0 as the Fortran program built it (its 1.7*exp(1.0) may be folded in
    # higher precision than the reference's default real)
    pin = load(tmp_path / "qmrherm_phi0_in.bin", (24, 12, 12, 12, 4, 25))
    assert rel(pin, st["phi0"][1:-1, 1:-1, 1:-1, 1:-1]) <= 1e-6
    phi0 = np.zeros_like(st["phi0"])
    for l in range(25):
        phi0[1:-1, 1:-1, 1:-1, 1:-1, :, l] = pin[..., l]
        phi0[..., l] = refstate.halo5(np.asfortranarray(phi0[..., l]))
    for iflag in range(4):
        p0 = phi0.copy(order="F")
        ds = np.asfortranarray(st["dsdpi"].copy())
        x, it, x1, p0, ds, _ = O.qmrherm(st["phi"], 0.1, 0.05, 3, anum, aden, iflag, st["u"], DIMS,
                                         phi0=p0, dsdpi=ds, max_qmr_iters=2)
        assert it == 2
        gx = load(tmp_path / f"qmrherm_x_{iflag}.bin", (24, 12, 12, 12, 4))
        check(gx, O.interior(x), 1e-10)
        gp = load(tmp_path / f"qmrherm_phi0_{iflag}.bin", (24, 12, 12, 12, 4, 25))
        check(gp, p0[1:-1, 1:-1, 1:-1, 1:-1], 1e-10)
        gd = load(tmp_path / f"qmrherm_dsdpi_{iflag}.bin", (12, 12, 12, 3), np.float64)
        check(gd, ds, 1e-10)
    # derivs on the Fortran program's own R and X
    R = load(tmp_path / "derivs_R.bin", (26, 14, 14, 14, 4))
    X = load(tmp_path / "derivs_X.bin", (26, 14, 14, 14, 4))
    assert rel(R, st["r"]) <= 1e-6 and rel(X, st["x"]) <= 1e-6
    ds = np.asfortranarray(st["dsdpi"].copy())
    O.derivs(R, X, TAU, 0, DIMS, ds)
    gd = load(tmp_path / "derivs_shamir.bin", (12, 12, 12, 3), np.float64)
    check(gd - st["dsdpi"], ds - st["dsdpi"], 1e-12)

def test_fortran_measure(tmp_path):
    """measure_shamir and measure (measure_module.F90:142-423) through the
    drop-in measure_module: the reference's known answer psibarpsi =
    2.504295E-04, aviter = 5 (tests/test_measure_shamir.F90:45-46) with the
    reference's noise (SITE_RANDOM streams, tests/fortran/test_random.F90),
    the same result through `measure` with isweep_total (dwf3d_lib.F90:355),
    the reference's output records (knoise lines in fort.100 per
    measurement, one in fort.200), and measure_wilson's known answer
    |psibarpsi| = 5.4226685E-02, aviter = 10.9 (tests/test_measure_wilson.F90:
    46-47; the reference's negative value predates its imass = 3 phase swap)."""
    exe = os.path.join(BIN, "test_measure_drop_in")
    assert os.path.exists(exe), f"{exe} not built (run __graft_entry__.build())"
    print("case: Fortran measure_shamir 12^3 x 24 imass 3", flush=True)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    print(p.stdout, flush=True)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "PASS measure_shamir" in p.stdout
    f100 = (tmp_path / "fort.100").read_text().splitlines()
    f200 = (tmp_path / "fort.200").read_text().splitlines()
    assert len(f100) == 30 and len(f200) == 3
    assert "Wilson measure" in p.stdout

    def fields(line, width, lead):  # fixed-width records: optional I5, then E fields
        body = line[5:] if lead else line
        return [float(body[k:k + width]) for k in range(0, len(body), width)]
    # '(4E17.9E3)' then '(I5,4E17.9E3)' with isweep_total; '(2E15.7E3)', '(I5,2E15.7E3)'
    assert len(f100[0]) == 68 and len(fields(f100[0], 17, False)) == 4
    assert len(f100[10]) == 73 and int(f100[10][:5]) == 7 and len(fields(f100[10], 17, True)) == 4
    assert len(f200[0]) == 30 and int(f200[1][:5]) == 7
    pbp0, pbp1 = fields(f200[0], 15, False)[0], fields(f200[1], 15, True)[0]
    assert abs(pbp0 - 2.504295e-04) <= 1e-3 * 2.504295e-04 and pbp1 == pbp0
