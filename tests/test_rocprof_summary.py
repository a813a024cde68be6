"""The profile summarizer reads rocprofv3's CSV and SQLite outputs alike and
derives calibrated traffic and matrix-core rates correctly (synthetic
result files)."""
import csv
import json
import os
import sqlite3
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
K = "void tp::k_demo<3>(tp::Args)"
COPY = "__amd_rocclr_copyBuffer(void*)"
CTR_COLS = ["Correlation_Id", "Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]


def write_csv(path, cols, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f, quoting=csv.QUOTE_NONNUMERIC)
        w.writerow(cols)
        w.writerows(rows)


def ctr_rows(kernel, counter, per_dispatch, dispatches, instances=2):
    rows = []
    for d in range(dispatches):
        for _ in range(instances):  # instances of one dispatch are summed
            rows.append([d, d, kernel, counter, per_dispatch / instances])
    return rows


def test_csv_and_db_summaries(tmp_path):
    t = str(tmp_path)
    write_csv(f"{t}/kt/h/kt_kernel_stats.csv",
              ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs", "StdDev"],
              [[K, 4, 400000, 100000, 90.0, 1, 1, 0], [COPY, 2, 20000, 10000, 10.0, 1, 1, 0]])
    # copy: FETCH counts half the bytes (factor 2), WRITE all of them
    write_csv(f"{t}/cf/h/cf_counter_collection.csv", CTR_COLS, ctr_rows(COPY, "FETCH_SIZE", 524288, 3))
    write_csv(f"{t}/cw/h/cw_counter_collection.csv", CTR_COLS, ctr_rows(COPY, "WRITE_SIZE", 1048576, 3))
    write_csv(f"{t}/f/h/f_counter_collection.csv", CTR_COLS, ctr_rows(K, "FETCH_SIZE", 1000, 4))
    write_csv(f"{t}/p/h/p_counter_collection.csv", CTR_COLS,
              ctr_rows(K, "SQ_INSTS_VALU_MFMA_MOPS_BF16", 1e6, 4) + ctr_rows(K, "SQ_WAVES", 64, 4))
    # the WRITE pass as a SQLite database
    os.makedirs(f"{t}/w/h")
    con = sqlite3.connect(f"{t}/w/h/w_results.db")
    con.execute("create table counters_collection (kernel_name, counter_name, dispatch_id, value)")
    con.executemany("insert into counters_collection values (?, ?, ?, ?)",
                    [(K, "WRITE_SIZE", d, 1500.0) for d in range(4)])
    con.commit()
    con.close()
    md, js = f"{t}/s.md", f"{t}/s.json"
    subprocess.check_call([sys.executable, "-m", "tile_processor_gpu_amd.rocprof_summary", "--kt", f"{t}/kt",
                           "--fetch", f"{t}/f", "--write", f"{t}/w", "--copy-fetch", f"{t}/cf",
                           "--copy-write", f"{t}/cw", "--pmc-dir", f"{t}/p", "--kernels", "void tp::k_demo",
                           "--cmd", "rocprofv3 --kernel-trace --stats -- python3 bench.py", "--title", "demo",
                           "-o", md, "--json", js], cwd=ROOT, stdout=subprocess.DEVNULL)
    r = json.load(open(js))["void tp::k_demo<3>"]
    assert r["avg_us"] == 100.0
    assert r["fetch_factor"] == 2.0 and r["write_factor"] == 1.0
    assert r["fetch_bytes"] == 1000 * 1024 * 2 and r["write_bytes"] == 1500 * 1024
    assert r["mfma_flop_bf16"] == 512e6
    assert abs(r["mfma_tflops_bf16"] - 512e6 / 100e-6 / 1e12) < 1e-9
    text = open(md).read()
    assert "rocprofv3 --kernel-trace --stats -- python3 bench.py" in text
    assert "| `SQ_WAVES` | 4 | 64 |" in text
    assert "0.2% of the 2500 TFLOP/s dense peak" in text
