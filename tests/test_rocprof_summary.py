"""The rocprofv3 summary tool on small synthetic CSVs: per-dispatch counter averages and the matrix-core rate."""
import csv

from hexed_amd import rocprof_summary


def write_counters(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for r in rows:
            w.writerow(r)


def write_stats(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Name", "Calls", "AverageNs", "MinNs", "MaxNs", "Percentage"])
        w.writeheader()
        for r in rows:
            w.writerow(r)


def test_counters_average_and_mfma_rate(tmp_path, capsys):
    k = "void hx::k_local_ns<3, 6, true, hx::Fp32_mfma_form, true>(hx::Dev_mesh)"
    # two dispatches; the counter of one dispatch comes in two instances that are summed
    write_counters(tmp_path / "c.csv", [
        {"Dispatch_Id": 1, "Kernel_Name": k, "Counter_Name": "SQ_INSTS_MFMA", "Counter_Value": 6e7},
        {"Dispatch_Id": 1, "Kernel_Name": k, "Counter_Name": "SQ_INSTS_MFMA", "Counter_Value": 4e7},
        {"Dispatch_Id": 2, "Kernel_Name": k, "Counter_Name": "SQ_INSTS_MFMA", "Counter_Value": 1.2e8},
        {"Dispatch_Id": 1, "Kernel_Name": k, "Counter_Name": "SQ_WAVES", "Counter_Value": 100},
        {"Dispatch_Id": 2, "Kernel_Name": k, "Counter_Name": "SQ_WAVES", "Counter_Value": 300},
    ])
    write_stats(tmp_path / "s.csv", [{"Name": k, "Calls": 2, "AverageNs": 2e6, "MinNs": 2e6, "MaxNs": 2e6,
                                      "Percentage": 100.0}])
    rocprof_summary.main(["counters", str(tmp_path / "c.csv"), "--stats", str(tmp_path / "s.csv"), "--title", "t"])
    out = capsys.readouterr().out
    row = [ln for ln in out.splitlines() if "k_local_ns" in ln][0].split("|")
    cells = [c.strip() for c in row[1:-1]]
    assert cells[1] == "2000.0"                       # average us
    assert float(cells[2]) == 1.1e8                   # SQ_INSTS_MFMA: (1e8 + 1.2e8) / 2
    assert float(cells[3]) == 200                     # SQ_WAVES
    tf = 1.1e8 * 512 / 2e-3 / 1e12                    # 28.16 TFLOP/s
    assert abs(float(cells[4]) - tf) < 0.01
    assert abs(float(cells[5]) - tf / 157.3) < 1e-3
