"""CPU: bench.py's rocprof tooling on synthetic rocprofv3 CSV files (no GPU, no profiler)."""
import csv
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

K1 = "void vt::k_local_move<false, 2>(vt::MoveArgs, int const*, vt::SweepPlan)"
K2 = "vt::(anonymous namespace)::k_cmp_count(long const*, long const*, long)"


def _write(path, header, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


@pytest.fixture
def dirs(tmp_path):
    kt, fe, wr = (str(tmp_path / d) for d in ("kt", "f", "w"))
    _write(os.path.join(kt, "x", "kt_kernel_stats.csv"), ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage"],
           [[K1, 4, 8_000_000, 2_000_000, 80.0], [K2, 2, 2_000_000, 1_000_000, 20.0]])
    cc = ["Kernel_Name", "Counter_Name", "Counter_Value"]
    _write(os.path.join(fe, "y", "f_counter_collection.csv"), cc,
           [[K1, "FETCH_SIZE", 1000.0], [K1, "FETCH_SIZE", 3000.0], [K2, "FETCH_SIZE", 500.0]])
    _write(os.path.join(wr, "z", "w_counter_collection.csv"), cc,
           [[K1, "WRITE_SIZE", 100.0], [K1, "WRITE_SIZE", 300.0], [K2, "WRITE_SIZE", 24.0]])
    return kt, fe, wr


def test_hbm_table_joins_duration_and_counters(dirs):
    out = bench.hbm_table(*dirs).splitlines()
    assert out[0].startswith("## kernel")
    row = [x.strip() for x in out[1].split("|")]
    # K1: mean FETCH 2000 KiB x 2 + mean WRITE 200 KiB = 4200 KiB per dispatch over 2 ms
    b = 1024.0 * (2000.0 * bench.FETCH_SCALE + 200.0 * bench.WRITE_SCALE)
    assert row[0] == "void vt::k_local_move<false, 2>" and row[1] == "4"
    assert float(row[3]) == pytest.approx(b / 1e6, abs=0.01)
    assert float(row[4]) == pytest.approx(b / 2e6, abs=0.1)
    assert float(row[5]) == pytest.approx(100 * b / 2e6 / bench.HBM_PEAK_GBS, abs=0.1)
    # anonymous-namespace kernels keep their name
    assert out[2].split("|")[0].strip() == "vt::k_cmp_count"


def test_summarize_profile_lists_stats_and_counters(dirs):
    text = bench.summarize_profile(os.path.dirname(dirs[0]))
    assert "void vt::k_local_move<false, 2> | 4 | 2.0000 | 8.000 | 80.00" in text
    assert "vt::k_cmp_count | FETCH_SIZE | 1 | 500" in text


def test_measured_traffic_reads_named_kernel(tmp_path, monkeypatch):
    prof = tmp_path / "profiles"
    prof.mkdir()
    (prof / "local_move_rgg_n8_fetch_write.txt").write_text(
        "# header\nvoid vt::k_local_move<false, 1> | FETCH_SIZE | 5 | 10\n"
        "void vt::k_local_move<false, 2> | FETCH_SIZE | 5 | 100\n"
        "void vt::k_local_move<false, 2> | WRITE_SIZE | 5 | 7\n")
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    t, src = bench.measured_traffic("local_move", "rgg_n8", "k_local_move<false, 2>")
    assert t == 1024.0 * (100 * bench.FETCH_SCALE + 7 * bench.WRITE_SCALE)
    assert src == os.path.join("profiles", "local_move_rgg_n8_fetch_write.txt")
    assert bench.measured_traffic("local_move", "rgg_n9", "k_local_move<false, 2>") == (None, None)


def test_degree_buckets_follow_the_bounds():
    deg = np.array([0, 3, 7, 8, 15, 16, 47, 48, 63, 64, 1000, 5000])
    b = bench.degree_buckets(deg, [7, 15, 47, 63, 191, 1023])
    assert b.tolist() == [0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 5, 6]
    # an empty bucket (-1) passes its rows to the next one
    b = bench.degree_buckets(deg, [7, 15, -1, 63, 191, 1023])
    assert b.tolist() == [0, 0, 0, 1, 1, 3, 3, 3, 3, 4, 5, 6]


def test_summarize_profile_lists_last_local_move_dispatches(tmp_path):
    """Kernel traces: the local-move launches one by one in start order, with the mean of the
    last five (a profiling run's timed sweeps); counters: the same kernels over their last
    five dispatches, listed after the all-dispatch means."""
    d = str(tmp_path / "p")
    K3 = "void vt::k_local_move<false, 2, false>(vt::MoveArgs, int const*, vt::SweepPlan)"
    rows = [[K3, 1000 * i, 1000 * i + ns] for i, ns in enumerate([9_000_000, 8_000_000] + [2_000_000] * 5)]
    rows.insert(3, [K2, 10, 20])
    _write(os.path.join(d, "t", "kt_kernel_trace.csv"), ["Kernel_Name", "Start_Timestamp", "End_Timestamp"],
           list(reversed(rows)))
    cc = [[K3, "FETCH_SIZE", v, i] for i, v in enumerate([90.0, 80.0, 10.0, 10.0, 10.0, 10.0, 10.0])]
    _write(os.path.join(d, "c", "f_counter_collection.csv"), ["Kernel_Name", "Counter_Name", "Counter_Value",
                                                              "Dispatch_Id"], cc)
    text = bench.summarize_profile(d)
    assert "void vt::k_local_move<false, 2, false> | 7 | 2.0000 2.0000 2.0000 2.0000 2.0000 | 2.0000" in text
    lines = text.splitlines()
    mean_all = lines.index("void vt::k_local_move<false, 2, false> | FETCH_SIZE | 7 | 31.4286")
    assert lines.index("void vt::k_local_move<false, 2, false> | FETCH_SIZE | 5 | 10") > mean_all
