"""On-disk formats: con (both layouts), midout, remez; and the reference's
production configuration as a realistic input."""
import json
import os

import numpy as np

import confio

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_read_reference_configuration():
    th, seed, layout = confio.read_con(os.path.join(GOLD, "con_1_16x16x16_mpi"), 16, 16, 16)
    assert layout == "mpi"
    assert th.shape == (16, 16, 16, 3) and th.dtype == np.float32
    assert seed == 92213954478080.0  # the 48-bit LCG seed, stored as real(dp)
    # an equilibrated beta = 0.3 ensemble: <theta> ~ 0, spread O(1/sqrt(beta))
    assert abs(float(th.mean())) < 0.05 and 0.2 < float(th.std()) < 2.0


def test_con_round_trip_both_layouts(tmp_path):
    rng = np.random.default_rng(0)
    th = np.asfortranarray(rng.standard_normal((4, 6, 8, 3)).astype(np.float32))
    for layout in ("serial", "mpi"):
        p = str(tmp_path / f"con_{layout}")
        confio.write_con(p, th, 123456789.0, layout)
        assert os.path.getsize(p) == confio.con_nbytes(4, 6, 8)
        t2, s2, l2 = confio.read_con(p, 4, 6, 8)
        assert l2 == layout and s2 == 123456789.0 and np.array_equal(t2, th)
    # the reference quirk: reading a serial file as MPI shifts theta by one element
    t3, _, _ = confio.read_con(str(tmp_path / "con_serial"), 4, 6, 8, layout="mpi")
    assert np.array_equal(t3.reshape(-1, order="F")[1:], th.reshape(-1, order="F")[:-1])


def test_midout_and_remez():
    m = confio.read_midout(os.path.join(GOLD, "midout_intruns"))
    assert m["beta"] == 0.3 and m["am"] == 0.005 and m["imass"] == 3
    (a4, d4), (b4, e4) = confio.remez("remez4")
    assert len(a4) == 26 and len(d4) == 25
    x = np.logspace(-4, np.log10(64), 2001)
    r = b4[0] + np.sum(b4[1:][None, :] / (x[:, None] + e4[None, :]), axis=1)
    ka = json.load(open(os.path.join(GOLD, "known_answers.json")))
    assert np.max(np.abs(r * x ** 0.25 - 1)) < 1.01 * ka["remez4_block2_max_rel_error"]["value"]


def test_congrad_on_reference_configuration():
    """CG on the production configuration.  The reference's run logged 209.9
    iterations on average at res 1e-8 (intruns/output:31), but from warm
    starts (vector::X carries over between calls) on sources M^+(1) Phi; a cold
    Gaussian source needs ~335 here.  Only the order of magnitude is checked."""
    import dense as D
    import oracle as O
    th, _, _ = confio.read_con(os.path.join(GOLD, "con_1_16x16x16_mpi"), 16, 16, 16)
    dims = (16, 16, 16, 8)
    u = O.coef(th, dims, ibound=-1)
    rng = np.random.default_rng(1)
    phi = D.rand_spinor(dims, rng)
    x = np.zeros_like(phi)
    it, _ = O.congrad(phi, x, u, dims, res=1e-8, am=0.005, imass=3)
    ka = json.load(open(os.path.join(GOLD, "known_answers.json")))["intruns_acceptance_congrad_avg"]
    assert 0.5 * ka["value"] < it < 2.5 * ka["value"], it
