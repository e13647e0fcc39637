"""The rocprofv3 summary helpers behind profiles/ and bench.py's roofline traffic:
kernel-name shortening, the copy calibration of FETCH_SIZE / WRITE_SIZE, and the
per-kernel averages, on small synthetic counter files."""
import csv
import json
import os

from cluestering_alpaka3_amd import profiling

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _write(path, rows, fields):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        for r in rows:
            w.writerow(r)


def test_short_name():
    assert profiling.short_name("void clue_hip::k_density<2, 0, 1, true, true>(clue_hip::PassArgs)") == "k_density"
    assert profiling.short_name("clue_hip::k_gather(clue_hip::Coords, float const*)") == "k_gather"


def test_traffic_calibration_and_averages(tmp_path):
    fields = ["Kernel_Name", "Counter_Name", "Counter_Value", "Grid_Size"]
    copy = "__amd_rocclr_copyBuffer"
    # the 64 MiB calibration copy reads as half its size on FETCH_SIZE and exactly on WRITE_SIZE
    fetch = [{"Kernel_Name": copy, "Counter_Name": "FETCH_SIZE", "Counter_Value": 32768.0, "Grid_Size": 1 << 20},
             {"Kernel_Name": copy, "Counter_Name": "FETCH_SIZE", "Counter_Value": 4.0, "Grid_Size": 64},
             {"Kernel_Name": "void clue_hip::k_density<2>(a)", "Counter_Name": "FETCH_SIZE",
              "Counter_Value": 100.0, "Grid_Size": 4096},
             {"Kernel_Name": "void clue_hip::k_density<2>(a)", "Counter_Name": "FETCH_SIZE",
              "Counter_Value": 300.0, "Grid_Size": 4096}]
    write = [{"Kernel_Name": copy, "Counter_Name": "WRITE_SIZE", "Counter_Value": 65536.0, "Grid_Size": 1 << 20},
             {"Kernel_Name": "void clue_hip::k_density<2>(a)", "Counter_Name": "WRITE_SIZE",
              "Counter_Value": 10.0, "Grid_Size": 4096},
             {"Kernel_Name": "void clue_hip::k_density<2>(a)", "Counter_Name": "WRITE_SIZE",
              "Counter_Value": 30.0, "Grid_Size": 4096}]
    fp, wp = tmp_path / "f.csv", tmp_path / "w.csv"
    _write(fp, fetch, fields)
    _write(wp, write, fields)
    res = profiling.traffic(str(fp), str(wp), copy_mib=64)
    c = res["calibration"]
    assert c["fetch_factor"] == 2.0 and c["write_factor"] == 1.0  # the largest copy dispatch is the calibration
    k = res["kernels"]["k_density"]
    assert k["dispatches"] == 2
    assert k["fetch_bytes"] == 200 * 1024 * 2 and k["write_bytes"] == 20 * 1024
    assert k["traffic_bytes"] == k["fetch_bytes"] + k["write_bytes"]
    assert copy not in res["kernels"]


def test_kernel_stats_sorted_by_share(tmp_path):
    p = tmp_path / "s.csv"
    _write(p, [{"Name": "a", "Calls": 2, "AverageNs": 10.0, "Percentage": 20.0},
               {"Name": "b", "Calls": 1, "AverageNs": 80.0, "Percentage": 80.0}],
           ["Name", "Calls", "AverageNs", "Percentage"])
    rows = profiling.kernel_stats(str(p))
    assert [r[0] for r in rows] == ["b", "a"]


def test_committed_traffic_matches_its_summary():
    # bench.py reads the density kernel's traffic from the committed JSON
    with open(os.path.join(ROOT, "profiles", "traffic_2d_1m_gauss.json")) as f:
        data = json.load(f)
    d = data["kernels"]["k_density"]
    assert d["traffic_bytes"] == d["fetch_bytes"] + d["write_bytes"] > 0
    with open(os.path.join(ROOT, "profiles", "traffic_2d_1m_gauss.txt")) as f:
        txt = f.read()
    assert f"{d['traffic_bytes'] / 1e6:10.2f}".strip() in txt
