"""Host logic of bench.py (no GPU): the counter summary it quotes for
roofline.traffic and the shape that summary was taken at, and the
algorithmic byte counts the roofline divides by."""
import json
import os

import bench


def test_traffic_summary_names_its_shape():
    """The committed FP64 traffic summary names the launch shape it was
    profiled at, and the bench reads it back as (tchunk, flags)."""
    tpath = os.path.join(bench.ROOT, "profiles", "dslash_f64_32x32x32x64_traffic.json")
    rec = json.load(open(tpath))
    shape = bench.traffic_shape()
    assert shape is not None
    assert f"f64:{shape[0]}:{shape[1]}" in rec["kernel"]
    assert 1 <= shape[0] <= 64 and 0 <= shape[1] < 512


def test_traffic_is_quoted_only_for_the_profiled_sources():
    """roofline.traffic is the summary's bytes when its source digest is the
    current Dslash sources' digest, and null with the reason otherwise."""
    bytes_, note = bench.measured_traffic()
    tpath = os.path.join(bench.ROOT, "profiles", "dslash_f64_32x32x32x64_traffic.json")
    rec = json.load(open(tpath))
    if rec["kernel_source_sha256"] == bench.kernel_source_digest():
        assert bytes_ == rec["hbm_bytes_per_launch"]
        # counted bytes at least the algorithmic input and output
        assert bytes_ >= rec["algorithmic_bytes_per_launch"]
    else:
        assert bytes_ is None and "predates" in note


def test_dslash_algorithmic_bytes():
    """2 V5 4 b_c per right-hand side plus 3 V4 b_link: 270.0 MB for one FP64
    application at 32^3 x Ls=64, 135.0 MB in FP32, 84.7 MB in FP16."""
    v4, v5 = 32 ** 3, 32 ** 3 * 64
    assert bench.dslash_bytes(v4, v5, "f64") == 2 * v5 * 64 + 3 * v4 * 16 == 270008320
    assert bench.dslash_bytes(v4, v5, "f32") == 2 * v5 * 32 + 3 * v4 * 8
    assert bench.dslash_bytes(v4, v5, "f16") == 2 * v5 * 20 + 3 * v4 * 8
    assert bench.dslash_bytes(v4, v5, "f64", 4) == 4 * 2 * v5 * 64 + 3 * v4 * 16
