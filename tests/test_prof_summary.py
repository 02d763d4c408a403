"""The rocprofv3 summarizer (mujoco_contact_surfaces_amd.prof_summary) on
synthetic CSV files shaped like rocprofv3's: kernel statistics, dispatch
shapes, HBM bytes scaled by a copy of known size, and the derived busy and
occupancy figures."""
import csv
import json
import os

from mujoco_contact_surfaces_amd import prof_summary as P


def _write(path, header, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def _counters(d, kernel, values, disp=7):
    hdr = ["Correlation_Id", "Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    rows = [[disp, disp, kernel, k, v] for k, v in values.items()]
    rows += [[1, 1, kernel, k, 1.0] for k in values]  # an earlier dispatch: ignored
    _write(os.path.join(d, "host", "123_counter_collection.csv"), hdr, rows)


def test_kernel_tables(tmp_path, capsys):
    d = str(tmp_path / "trace")
    _write(os.path.join(d, "h", "1_kernel_stats.csv"),
           ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs", "StdDev"],
           [["void hcs::k_cand<0, 0>(hcs::QueryArgs)", 2, 3000000, 1500000, 60.0, 1400000, 1600000, 1]])
    _write(os.path.join(d, "h", "1_kernel_trace.csv"),
           ["Kernel_Name", "Grid_Size_X", "Start_Timestamp", "End_Timestamp"],
           [["void hcs::k_cand<0, 0>(hcs::QueryArgs)", 196608, 0, 1400000],
            ["void hcs::k_cand<0, 0>(hcs::QueryArgs)", 196608, 2000000, 3600000]])
    P.main(["kernels", d, "--command", "rocprofv3 ..."])
    out = capsys.readouterr().out
    assert "| 2 | 3.000 | 1500.0 | 1400.0 | 1600.0 | 60.00 |" in out
    assert "| 196608 | 2 | 1500.0 | 1400.0 | 1600.0 |" in out


def test_counter_scaling_and_derived_figures(tmp_path, capsys):
    k = "void hcs::k_cand<0, 0>(hcs::QueryArgs)"
    copy = "__amd_rocclr_copyBuffer"
    _counters(str(tmp_path / "cf"), copy, {"FETCH_SIZE": 524288.0})      # reads half of 1 GiB
    _counters(str(tmp_path / "cw"), copy, {"WRITE_SIZE": 1048576.0})
    _counters(str(tmp_path / "f"), k, {"FETCH_SIZE": 1000.0})
    _counters(str(tmp_path / "w"), k, {"WRITE_SIZE": 10.0})
    _counters(str(tmp_path / "p1"), k, {"SQ_WAVE_CYCLES": 3.0 * 256 * 100, "GRBM_GUI_ACTIVE": 100.0, "SQ_WAVES": 3072.0})
    _counters(str(tmp_path / "p2"), k, {"TD_TD_BUSY_sum": 0.5 * 256 * 100, "TA_TA_BUSY_sum": 0.25 * 256 * 100,
                                         "GRBM_GUI_ACTIVE": 100.0, "TCP_TCC_READ_REQ_LATENCY_sum": 8000.0,
                                         "TCP_TCC_READ_REQ_sum": 10.0})
    js = str(tmp_path / "out.json")
    P.main(["counters", "--kernels", "k_cand", "--fetch", str(tmp_path / "f"), "--write", str(tmp_path / "w"),
            "--calib-fetch", str(tmp_path / "cf"), "--calib-write", str(tmp_path / "cw"), "--calib-mib", "1024",
            "--pass", str(tmp_path / "p1"), "--pass", str(tmp_path / "p2"), "--json", js])
    capsys.readouterr()
    d = json.load(open(js))["per_kernel"]["k_cand"]
    assert d["hbm_read_bytes"] == 1000.0 * 1024 * 2.0
    assert d["hbm_write_bytes"] == 10.0 * 1024
    assert abs(d["mean_waves_per_cu"] - 12.0) < 1e-12  # SQ_WAVE_CYCLES counts 4-cycle units
    assert abs(d["td_busy_pct"] - 50.0) < 1e-12 and abs(d["ta_busy_pct"] - 25.0) < 1e-12
    assert d["l2_read_latency_cycles"] == 800.0
    assert json.load(open(js))["hbm_bytes_per_launch"]["k_cand"] == 1000.0 * 1024 * 2.0 + 10.0 * 1024


def test_xcd_count_and_valu_figures(tmp_path, capsys):
    """GRBM_GUI_ACTIVE is summed over the XCDs that agent_info.csv names;
    VALU issue use, active lanes and waiting share."""
    k = "void hcs::k_clip<double, 0>(hcs::QueryArgs)"
    p = str(tmp_path / "p")
    _counters(p, k, {"GRBM_GUI_ACTIVE": 8 * 1000.0, "SQ_INSTS_VALU": 0.5 * 1000 * 256,
                     "SQ_THREAD_CYCLES_VALU": 0.25 * 64 * 0.5 * 1000 * 256, "SQ_WAVE_CYCLES": 400.0,
                     "SQ_WAIT_ANY": 100.0, "TD_TD_BUSY_sum": 0.5 * 256 * 1000})
    _write(os.path.join(p, "host", "123_agent_info.csv"), ["Node_Id", "Agent_Type", "Num_Xcc"],
           [[0, "CPU", 0], [1, "GPU", 8]])
    js = str(tmp_path / "out.json")
    P.main(["counters", "--kernels", "k_clip", "--pass", p, "--json", js])
    out = capsys.readouterr().out
    assert "/ 8 XCDs" in out
    d = json.load(open(js))["per_kernel"]["k_clip"]
    assert d["gpu_clocks"] == 1000.0
    assert abs(d["td_busy_pct"] - 50.0) < 1e-12
    assert abs(d["valu_issue_pct"] - 50.0) < 1e-12
    assert abs(d["valu_lanes_pct"] - 25.0) < 1e-12
    assert abs(d["wait_pct"] - 25.0) < 1e-12
