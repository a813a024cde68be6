"""Summaries of rocprofv3 results for the profiles folder.

Reads what ``rocprofv3 -d DIR -o NAME`` writes, in either output format:
CSV (``--output-format csv``: ``NAME_kernel_stats.csv``,
``NAME_counter_collection.csv``) or the SQLite database
(``NAME_results.db``, tables ``top_kernels`` and ``counters_collection``).

* a ``--kernel-trace --stats`` run gives per-kernel call counts and average
  durations;
* ``--pmc FETCH_SIZE`` and ``--pmc WRITE_SIZE`` runs give per-dispatch KiB;
  the same two counters over the 1 GiB device copy of ``bench.py --mode
  copy`` calibrate them: factor = known bytes / counted bytes;
* further ``--pmc`` runs (``--pmc-dir``) give any other counters per
  dispatch; ``SQ_INSTS_VALU_MFMA_MOPS_<type>`` counts matrix-core work in
  units of 512 FLOP, from which the matrix-core rate and its fraction of the
  dense peak follow (with the kernel-trace duration).

Writes a Markdown summary (the figures used, the commands that made them)
and a JSON file ``{kernel: {"fetch_bytes", "write_bytes", "avg_us", ...}}``
that ``bench.py`` reports as the roofline's ``traffic``.

    python -m tile_processor_gpu_amd.rocprof_summary --kt DIR [--fetch DIR --write DIR \\
        --copy-fetch DIR --copy-write DIR] [--pmc-dir DIR ...] [--kernels PREFIX ...] \\
        --cmd "..." [--cmd "..."] --title "..." -o profiles/x.md [--json profiles/x.json]
"""
import argparse
import collections
import csv
import glob
import json
import os
import sqlite3

COPY_BYTES = 1 << 30
COPY_KERNEL = "__amd_rocclr_copyBuffer"
MOPS_FLOP = 512  # one SQ_INSTS_VALU_MFMA_MOPS_* count
# MI355X dense matrix-core peaks, TFLOP/s
MFMA_PEAK_TFLOPS = {"BF16": 2500.0, "F16": 2500.0, "F8": 5000.0, "F6F4": 10100.0, "I8": 5000.0,
                    "F32": 157.3, "XF32": 157.3, "F64": 78.6}


def _find(d, csv_suffix):
    """('csv', path) or ('db', path) for the first result file under d."""
    for kind, pat in (("csv", "*" + csv_suffix), ("db", "*_results.db")):
        files = sorted(glob.glob(os.path.join(d, "**", pat), recursive=True))
        if files:
            return kind, files[0]
    raise SystemExit(f"no *{csv_suffix} or *_results.db under {d}")


def _rows(db, sql):
    con = sqlite3.connect(db)
    try:
        return con.execute(sql).fetchall()
    finally:
        con.close()


def short(name):
    """Kernel name without the argument list."""
    depth = 0
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0 and i > 0:
            return name[:i]
    return name


def kernel_stats(d):
    """{kernel: (calls, total_us, avg_us, percent)} from a --stats run."""
    kind, path = _find(d, "_kernel_stats.csv")
    out = {}
    if kind == "db":  # the table holds microseconds
        for name, calls, total, avg, pct in _rows(
                path, "select name, total_calls, total_duration, average, percentage from top_kernels"):
            out[short(name)] = (int(calls), float(total), float(avg), float(pct))
        return out
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            out[short(r["Name"])] = (int(r["Calls"]), float(r["TotalDurationNs"]) / 1e3,
                                     float(r["AverageNs"]) / 1e3, float(r["Percentage"]))
    return out


def counters(d):
    """{kernel: {counter: (dispatches, mean value per dispatch)}}; a
    counter's instances (rows of one dispatch) are summed first."""
    kind, path = _find(d, "_counter_collection.csv")
    per = collections.defaultdict(float)  # (kernel, counter, dispatch) -> value
    if kind == "db":
        rows = _rows(path, "select kernel_name, counter_name, dispatch_id, value from counters_collection")
    else:
        with open(path, newline="") as f:
            rows = [(r["Kernel_Name"], r["Counter_Name"], r["Dispatch_Id"], r["Counter_Value"])
                    for r in csv.DictReader(f)]
    for name, ctr, disp, val in rows:
        per[(short(name), ctr, int(disp))] += float(val)
    acc = collections.defaultdict(list)
    for (k, c, _), v in per.items():
        acc[(k, c)].append(v)
    out = collections.defaultdict(dict)
    for (k, c), vs in acc.items():
        out[k][c] = (len(vs), sum(vs) / len(vs))
    return out


def _keep(names, prefixes):
    return [k for k in names if k.startswith(tuple(prefixes))] if prefixes else list(names)


def summarize(a):
    kt = kernel_stats(a.kt)
    lines = [f"# {a.title}", "", "Commands (from the repository root, on one MI355X):", ""]
    lines += [f"    {c}" for c in a.cmd]
    lines += ["", "## Kernel trace (`--kernel-trace --stats`)", "",
              "| kernel | calls | avg us | share % |", "|---|---|---|---|"]
    for k in sorted(kt, key=lambda k: -kt[k][1]):
        c, _, avg, pct = kt[k]
        lines.append(f"| `{k}` | {c} | {avg:.2f} | {pct:.1f} |")
    result = {}
    if a.fetch and a.write and a.copy_fetch and a.copy_write:
        fetch = {k: v["FETCH_SIZE"] for k, v in counters(a.fetch).items() if "FETCH_SIZE" in v}
        write = {k: v["WRITE_SIZE"] for k, v in counters(a.write).items() if "WRITE_SIZE" in v}
        cf = counters(a.copy_fetch)[COPY_KERNEL]["FETCH_SIZE"][1]
        cw = counters(a.copy_write)[COPY_KERNEL]["WRITE_SIZE"][1]
        f_fac = COPY_BYTES / (cf * 1024.0)
        w_fac = COPY_BYTES / (cw * 1024.0)
        lines += ["", "## HBM traffic (`--pmc FETCH_SIZE`, `--pmc WRITE_SIZE`)", "",
                  f"Calibration on a {COPY_BYTES >> 20} MiB device copy (`{COPY_KERNEL}`): "
                  f"FETCH_SIZE reads {cf:.0f} KiB -> factor {f_fac:.3f}; "
                  f"WRITE_SIZE reads {cw:.0f} KiB -> factor {w_fac:.3f}.", "",
                  "| kernel | FETCH KiB | WRITE KiB | read MB (calibrated) | written MB | "
                  "avg us | effective GB/s |", "|---|---|---|---|---|---|---|"]
        for k in _keep(kt, a.kernels):
            if k not in fetch or k not in write:
                continue
            fb = fetch[k][1] * 1024.0 * f_fac
            wb = write[k][1] * 1024.0 * w_fac
            avg = kt[k][2]
            result[k] = {"fetch_bytes": fb, "write_bytes": wb, "avg_us": avg,
                         "fetch_factor": f_fac, "write_factor": w_fac}
            lines.append(f"| `{k}` | {fetch[k][1]:.0f} | {write[k][1]:.0f} | {fb / 1e6:.1f} | "
                         f"{wb / 1e6:.1f} | {avg:.2f} | {(fb + wb) / avg / 1e3:.1f} |")
        try:
            cavg = kernel_stats(a.copy_fetch).get(COPY_KERNEL, (0, 0, None, 0))[2]
        except SystemExit:
            cavg = None
        if cavg:
            lines += ["", f"Copy kernel under the FETCH pass: {cavg:.1f} us for "
                          f"{2 * COPY_BYTES / 1e9:.2f} GB moved = "
                          f"{2 * COPY_BYTES / cavg / 1e3:.0f} GB/s."]
    if a.pmc_dir:
        merged = collections.defaultdict(dict)
        for d in a.pmc_dir:
            for k, v in counters(d).items():
                merged[k].update(v)
        for k in _keep(merged, a.kernels):
            ctr = merged[k]
            lines += ["", f"## Counters of `{k}` (mean per dispatch)", "",
                      "| counter | dispatches | value |", "|---|---|---|"]
            for c in sorted(ctr):
                lines.append(f"| `{c}` | {ctr[c][0]} | {ctr[c][1]:.6g} |")
            avg = kt.get(k, (0, 0, None, 0))[2]
            for c in sorted(ctr):
                if not c.startswith("SQ_INSTS_VALU_MFMA_MOPS_"):
                    continue
                typ = c[len("SQ_INSTS_VALU_MFMA_MOPS_"):]
                flop = ctr[c][1] * MOPS_FLOP
                rec = result.setdefault(k, {"avg_us": avg})
                rec[f"mfma_flop_{typ.lower()}"] = flop
                if avg and flop > 0 and typ in MFMA_PEAK_TFLOPS:
                    tf = flop / (avg * 1e-6) / 1e12
                    rec[f"mfma_tflops_{typ.lower()}"] = tf
                    lines += ["", f"Matrix-core work ({typ}): {ctr[c][1]:.6g} x {MOPS_FLOP} = "
                                  f"{flop / 1e9:.2f} GFLOP per dispatch; over the traced {avg:.2f} us: "
                                  f"{tf:.1f} TFLOP/s = {100 * tf / MFMA_PEAK_TFLOPS[typ]:.1f}% of the "
                                  f"{MFMA_PEAK_TFLOPS[typ]:g} TFLOP/s dense peak."]
    if a.note:
        lines += ["", a.note]
    with open(a.o, "w") as f:
        f.write("\n".join(lines) + "\n")
    if a.json:
        with open(a.json, "w") as f:
            json.dump(result, f, indent=1)
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--kt", required=True, help="result folder of a --kernel-trace --stats run")
    for k in ("fetch", "write", "copy_fetch", "copy_write"):
        ap.add_argument("--" + k.replace("_", "-"), dest=k)
    ap.add_argument("--pmc-dir", nargs="*", default=[], help="result folders of further --pmc runs")
    ap.add_argument("--kernels", nargs="*", default=[], help="name prefixes to keep")
    ap.add_argument("--cmd", action="append", default=[], help="a command that made the results")
    ap.add_argument("--title", default="rocprofv3 summary")
    ap.add_argument("--note", default="")
    ap.add_argument("-o", required=True)
    ap.add_argument("--json")
    a = ap.parse_args()
    t = summarize(a)
    print(json.dumps(t, indent=1))


if __name__ == "__main__":
    main()
