"""The committed rocprofv3 summaries that bench.py folds into its JSON line cover the
launch shapes it reports: the traffic table is keyed by the same kernel labels the bench
builds, and the counter table names the kernels the tuner picks at 2048^2 x 1800."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_traffic_table_covers_every_forward_shape_and_the_tuned_back():
    for f in range(3):
        t = bench.measured_traffic(bench.forward_half_name(f))
        assert t is not None and t > 0, bench.forward_half_name(f)
    assert bench.measured_traffic(bench.back_half_name(0)) > 0


def test_counter_table_names_the_tuned_kernels():
    with open(os.path.join(ROOT, "profiles", "counters_dd_sirt_2048x1800.json")) as f:
        table = json.load(f)["kernels"]
    assert table["forward"]["kernels"] == ["dd_fwd_kernel<0, 0, false, 64, 1>",
                                           "dd_fwd_kernel<0, 0, false, 64, 2>"]
    assert table["back"]["kernels"] == ["dd_back_kernel<0, 3, false, 64, 4>"]
    for half in ("forward", "back"):
        assert 0.0 < table[half]["valu_issue_util"] < 1.0
    assert bench.measured_limiter("forward") == table["forward"]
