"""Host-side measurement tools (CPU): the growth-rate fit of the validation
script, the seeded ensemble member draw, and the rocprofv3 CSV summaries."""
import csv
import os

import numpy as np

from unstaggered_pic_amd import ensemble_bench, profile_summary, validate


def test_growth_rate_recovers_exponential_phase():
    """Noise plateau, exponential growth of the field amplitude at rate 0.08
    (energy at 0.16), saturation: the fit returns 0.08 and a window inside the
    exponential phase."""
    t = np.arange(0, 100.0, 0.05)
    gamma = 0.08
    wb = 1e-2 + 1e-4 * np.exp(2 * gamma * t)
    wb = np.minimum(wb, 5.0)
    wb[0] = 0.0
    rate, window = validate.growth_rate(t, wb)
    assert abs(rate - gamma) / gamma < 2e-2
    assert window[0] > 10.0 and window[1] < 60.0


def test_growth_rate_none_without_growth():
    t = np.arange(0, 10.0, 0.05)
    assert validate.growth_rate(t, np.full(t.size, 1e-2))[0] is None


def test_ensemble_members_are_seeded_and_in_range():
    a, b = ensemble_bench.member_specs(16), ensemble_bench.member_specs(16)
    assert a == b
    for s in a:
        assert s["side"] in (64, 128, 256) and s["ppc"] in (16, 64, 256)
        assert 0.02 <= s["vth"] <= 0.2 and 1e-3 <= s["amp"] <= 1e-2
    assert {s["kind"] for s in a} == {"landau", "weibel"}


def test_trace_stats_after_marker(tmp_path):
    rows = [("k_mlp_adam", 0, 10), ("k_a", 20, 30), ("k_b", 30, 50), ("k_a", 60, 65)]
    path = tmp_path / "run_kernel_trace.csv"
    with open(path, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        for r in rows:
            w.writerow(r)
    agg, span, work = profile_summary.trace_stats(str(tmp_path), "k_mlp")
    assert agg == {"k_a": [2, 15], "k_b": [1, 20]}
    assert span == 45
    assert work == {"k_a": [10, 5], "k_b": [20]}


def test_gated_launches_left_out_of_means(tmp_path):
    """A launch whose gate was closed returns at once: below 5% of the
    kernel's median it does not enter the means."""
    path = tmp_path / "run_counter_collection.csv"
    with open(path, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Kernel_Name", "Counter_Name", "Counter_Value"])
        for v in (100.0, 0.5, 102.0):
            w.writerow(["k", "SQ_INSTS_VALU", v])
    d = profile_summary.counters_by_name(str(tmp_path))
    assert d["k"]["SQ_INSTS_VALU"] == (101.0, 2)
    assert profile_summary.counters(str(tmp_path))["k"] == [100.0, 102.0]


def test_counters_by_name_averages_last_dispatches(tmp_path):
    path = tmp_path / "run_counter_collection.csv"
    with open(path, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Kernel_Name", "Counter_Name", "Counter_Value"])
        for v in (1, 2, 3, 4, 5, 6):
            w.writerow(["k", "SQ_WAVES", v])
    d = profile_summary.counters_by_name(str(tmp_path))
    assert d["k"]["SQ_WAVES"] == (4.5, 6)
    assert os.path.exists(path)
