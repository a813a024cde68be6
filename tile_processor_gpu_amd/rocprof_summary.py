"""Summaries of rocprofv3 result databases for the profiles folder.

Reads the SQLite files that ``rocprofv3 -d DIR -o NAME`` writes
(``DIR/**/NAME_results.db``):

* a ``--kernel-trace --stats`` run gives per-kernel call counts and average
  durations (table ``top_kernels``);
* ``--pmc FETCH_SIZE`` and ``--pmc WRITE_SIZE`` runs give per-dispatch KiB
  (table ``counters_collection``);
* the same two counters over the 1 GiB device copy of ``bench.py --mode copy``
  calibrate them: factor = known bytes / counted bytes.

Writes a Markdown summary (the figures used, the commands that made them)
and a JSON file ``{kernel: {"fetch_bytes", "write_bytes", "avg_us"}}`` that
``bench.py`` reports as the roofline's ``traffic``.

    python -m tile_processor_gpu_amd.rocprof_summary --kt DIR --fetch DIR --write DIR \
        --copy-fetch DIR --copy-write DIR --cmd-kt "..." --cmd-pmc "..." \
        --title "..." -o profiles/x.md --json profiles/x.json
"""
import argparse
import glob
import json
import os
import sqlite3

COPY_BYTES = 1 << 30
COPY_KERNEL = "__amd_rocclr_copyBuffer"


def _db(d):
    files = sorted(glob.glob(os.path.join(d, "**", "*_results.db"), recursive=True))
    if not files:
        raise SystemExit(f"no *_results.db under {d}")
    return files[0]


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
    """{kernel: (calls, total_us, avg_us, percent)} from a --stats run (the
    table holds microseconds)."""
    out = {}
    for name, calls, total, avg, pct in _rows(
            _db(d), "select name, total_calls, total_duration, average, percentage from top_kernels"):
        out[short(name)] = (int(calls), float(total), float(avg), float(pct))
    return out


def counter_kib(d, counter):
    """{kernel: (dispatches, mean KiB per dispatch)} for one PMC counter."""
    out = {}
    for name, mean, n in _rows(
            _db(d), "select kernel_name, avg(value), count(*) from counters_collection "
                    f"where counter_name = '{counter}' group by kernel_name"):
        out[short(name)] = (int(n), float(mean))
    return out


def summarize(a):
    kt = kernel_stats(a.kt)
    fetch = counter_kib(a.fetch, "FETCH_SIZE")
    write = counter_kib(a.write, "WRITE_SIZE")
    cf = counter_kib(a.copy_fetch, "FETCH_SIZE")[COPY_KERNEL][1]
    cw = counter_kib(a.copy_write, "WRITE_SIZE")[COPY_KERNEL][1]
    f_fac = COPY_BYTES / (cf * 1024.0)
    w_fac = COPY_BYTES / (cw * 1024.0)
    names = [k for k in kt if k.startswith(tuple(a.kernels))] if a.kernels else list(kt)
    traffic = {}
    lines = [f"# {a.title}", "", "Commands (from the repository root, on one MI355X):", ""]
    lines += [f"    {c}" for c in (a.cmd_kt, a.cmd_pmc) if c]
    lines += ["", "## Kernel trace (`--kernel-trace --stats`)", "",
              "| kernel | calls | avg us | share % |", "|---|---|---|---|"]
    for k in sorted(kt, key=lambda k: -kt[k][1]):
        c, _, avg, pct = kt[k]
        lines.append(f"| `{k}` | {c} | {avg:.2f} | {pct:.1f} |")
    lines += ["", "## HBM traffic (`--pmc FETCH_SIZE`, `--pmc WRITE_SIZE`)", "",
              f"Calibration on a {COPY_BYTES >> 20} MiB device copy (`{COPY_KERNEL}`): "
              f"FETCH_SIZE reads {cf:.0f} KiB -> factor {f_fac:.3f}; "
              f"WRITE_SIZE reads {cw:.0f} KiB -> factor {w_fac:.3f}.", "",
              "| kernel | FETCH KiB | WRITE KiB | read MB (calibrated) | written MB | "
              "avg us | effective GB/s |", "|---|---|---|---|---|---|---|"]
    for k in names:
        if k not in fetch or k not in write:
            continue
        fb = fetch[k][1] * 1024.0 * f_fac
        wb = write[k][1] * 1024.0 * w_fac
        avg = kt[k][2]
        traffic[k] = {"fetch_bytes": fb, "write_bytes": wb, "avg_us": avg,
                      "fetch_factor": f_fac, "write_factor": w_fac}
        lines.append(f"| `{k}` | {fetch[k][1]:.0f} | {write[k][1]:.0f} | {fb / 1e6:.1f} | "
                     f"{wb / 1e6:.1f} | {avg:.2f} | {(fb + wb) / avg / 1e3:.1f} |")
    cavg = None
    try:
        cavg = kernel_stats(a.copy_fetch).get(COPY_KERNEL, (0, 0, None, 0))[2]
    except SystemExit:
        pass
    if cavg:
        lines += ["", f"Copy kernel under the FETCH pass: {cavg:.1f} us for "
                      f"{2 * COPY_BYTES / 1e9:.2f} GB moved = "
                      f"{2 * COPY_BYTES / cavg / 1e3:.0f} GB/s."]
    if a.note:
        lines += ["", a.note]
    with open(a.o, "w") as f:
        f.write("\n".join(lines) + "\n")
    if a.json:
        with open(a.json, "w") as f:
            json.dump(traffic, f, indent=1)
    return traffic


def main():
    ap = argparse.ArgumentParser()
    for k in ("kt", "fetch", "write", "copy_fetch", "copy_write"):
        ap.add_argument("--" + k.replace("_", "-"), dest=k, required=True)
    ap.add_argument("--kernels", nargs="*", default=[], help="name prefixes to keep")
    ap.add_argument("--cmd-kt", default="")
    ap.add_argument("--cmd-pmc", default="")
    ap.add_argument("--title", default="rocprofv3 summary")
    ap.add_argument("--note", default="")
    ap.add_argument("-o", required=True)
    ap.add_argument("--json")
    a = ap.parse_args()
    t = summarize(a)
    print(json.dumps(t, indent=1))


if __name__ == "__main__":
    main()
