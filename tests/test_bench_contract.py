"""bench.py's roofline inputs that come from the committed profiles: the
algorithmic bytes of the five passes and the PMC traffic of the x pass for
every kernel variant the autotuner may pick at 128^3 (complex64, diagonal
chi), so that the bench line's ``roofline.traffic`` is measured whichever
variant the timed run uses."""
from __future__ import annotations

import os

import bench

M = 128 ** 3
X_VARIANTS = 4    # x-pass variants of the 256-point complex64 plan (fast_variant_count: 4, the small tile apart)


def test_pass_bytes_closed_form():
    b = dict(zip(bench.PASS_NAMES, bench.pass_bytes(128, 8, 3)))
    assert b["x_fwd_khat_inv"] == (24 * M + 6 * 129 ** 3) * 8 == 505_694_256
    assert sum(b.values()) == 1_512_327_216
    assert b["y_fwd"] == b["y_inv"] == 18 * M * 8


def test_traffic_summary_covers_every_x_variant():
    assert os.path.exists(bench.TRAFFIC_SUMMARY)
    alg = bench.pass_bytes(128, 8, 3)[bench.PASS_NAMES.index("x_fwd_khat_inv")]
    for v in range(X_VARIANTS):
        t = bench.measured_traffic("x_fwd_khat_inv", 128, 3, 8, variant=v)
        assert t is not None, v
        # HBM bytes cannot fall below the algorithmic bytes by more than the
        # counters' rounding, and a variant reading every row twice stays below 2.2x
        assert 0.98 * alg < t < 2.2 * alg, (v, t)


def test_traffic_is_none_for_other_workloads():
    assert bench.measured_traffic("x_fwd_khat_inv", 96, 3, 8) is None
    assert bench.measured_traffic("x_fwd_khat_inv", 128, 3, 16) is None
    assert bench.measured_traffic("x_fwd_khat_inv", 128, 1, 8) is None
    assert bench.measured_traffic("x_fwd_khat_inv", 128, 3, 8, variant=9) is None
