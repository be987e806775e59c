"""Summaries of rocprofv3 output, and the HBM traffic they record.

    python -m wadjet_amd.profiling stats <rocprof dir> --command "<cmd>"
        per-kernel dispatch count / total / average / min / max duration of a
        `rocprofv3 --kernel-trace --stats` run (its *_results.db, or the
        *_kernel_stats.csv of --output-format csv)
    python -m wadjet_amd.profiling counters <dir>... [--kernel K]... [--pick largest] --command "<cmd>"
        raw counters per kernel from several `rocprofv3 --pmc` runs (LDS bank
        conflicts, L2 hits and atomics, HBM read / write request counts and
        the bytes they move), with derived ratios
    python -m wadjet_amd.profiling pmc <FETCH_SIZE dir> <WRITE_SIZE dir>
            --copy-bytes B --steps K --command "<cmd>" --command "<cmd>"
        per-kernel FETCH_SIZE / WRITE_SIZE (KiB per dispatch) from two
        `rocprofv3 --pmc` runs, each scaled by the factor that makes the same
        counter read exactly B bytes for copy_words_kernel (a B-byte
        device-to-device copy run in the same process)

Both print a short text summary meant for profiles/. `load_traffic` reads the
`traffic` lines of a pmc summary back (bench.py uses it for roofline.traffic).
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import sqlite3
from collections import defaultdict

COPY_KERNEL = "copy_words_kernel"


def _files(d: str, suffix: str, required: bool = True):
    out = sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True))
    if not out and required:
        raise SystemExit(f"no *{suffix} under {d}")
    return out


def short_name(name: str) -> str:
    """Kernel name without the argument list (template arguments kept)."""
    depth = 0
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return name[:i]
    return name


def kernel_stats(d: str):
    rows = []
    dbs = _files(d, "_results.db", required=False)
    if dbs:
        per = defaultdict(list)
        for path in dbs:
            con = sqlite3.connect(path)
            for name, dur in con.execute("select name, duration from kernels"):
                per[short_name(name)].append(float(dur))
            con.close()
        total = sum(sum(v) for v in per.values()) or 1.0
        for k, v in per.items():
            rows.append(dict(name=k, calls=len(v), total_ns=sum(v), avg_ns=sum(v) / len(v), min_ns=min(v),
                             max_ns=max(v), pct=100.0 * sum(v) / total))
    else:
        for path in _files(d, "kernel_stats.csv"):
            with open(path, newline="") as f:
                for r in csv.DictReader(f):
                    rows.append(dict(name=short_name(r["Name"]), calls=int(r["Calls"]),
                                     total_ns=float(r["TotalDurationNs"]), avg_ns=float(r["AverageNs"]),
                                     min_ns=float(r["MinNs"]), max_ns=float(r["MaxNs"]),
                                     pct=float(r["Percentage"])))
    rows.sort(key=lambda r: -r["total_ns"])
    return rows


def format_stats(rows, commands):
    out = ["# rocprofv3 --kernel-trace --stats summary", "# command(s), run from the repository root:"]
    out += [f"#   {c}" for c in commands]
    out.append(f"{'kernel':64s} {'calls':>6s} {'total ms':>10s} {'avg us':>10s} {'min us':>10s} {'max us':>10s} {'%':>6s}")
    for r in rows:
        out.append(f"{r['name'][:64]:64s} {r['calls']:6d} {r['total_ns'] / 1e6:10.3f} {r['avg_ns'] / 1e3:10.1f} "
                   f"{r['min_ns'] / 1e3:10.1f} {r['max_ns'] / 1e3:10.1f} {r['pct']:6.2f}")
    return "\n".join(out) + "\n"


def pmc_per_kernel(d: str, counter: str):
    """{kernel: [value per dispatch, ...]} for one counter (values in KiB)."""
    vals = defaultdict(dict)
    dbs = _files(d, "_results.db", required=False)
    for path in dbs:
        con = sqlite3.connect(path)
        for name, disp, v in con.execute("select name, dispatch_id, counter_value from pmc_events "
                                         "where counter_name = ?", (counter,)):
            k = short_name(name)
            vals[k][(path, disp)] = vals[k].get((path, disp), 0.0) + float(v)
        con.close()
    if dbs:
        return {k: list(v.values()) for k, v in vals.items()}
    for path in _files(d, "counter_collection.csv"):
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter:
                    continue
                key = (path, r.get("Dispatch_Id") or r.get("Correlation_Id"))
                vals[short_name(r["Kernel_Name"])][key] = vals[short_name(r["Kernel_Name"])].get(key, 0.0) + \
                    float(r["Counter_Value"])
    return {k: list(v.values()) for k, v in vals.items()}


def pmc_summary(fetch_dir, write_dir, copy_bytes, steps, commands):
    fetch = pmc_per_kernel(fetch_dir, "FETCH_SIZE")
    write = pmc_per_kernel(write_dir, "WRITE_SIZE")
    cf = [k for k in fetch if COPY_KERNEL in k]
    cw = [k for k in write if COPY_KERNEL in k]
    if not cf or not cw:
        raise SystemExit("the calibration copy kernel is missing from the counter runs")
    f_cal = copy_bytes / (1024.0 * max(fetch[cf[0]]))
    w_cal = copy_bytes / (1024.0 * max(write[cw[0]]))
    out = ["# HBM traffic per dispatch from two rocprofv3 --pmc runs (FETCH_SIZE, then WRITE_SIZE)",
           "# command(s), run from the repository root:"]
    out += [f"#   {c}" for c in commands]
    out.append(f"# calibration: {COPY_KERNEL} copying {copy_bytes} bytes reads FETCH_SIZE "
               f"{max(fetch[cf[0]]):.0f} KiB (factor {f_cal:.4f}) and WRITE_SIZE {max(write[cw[0]]):.0f} KiB "
               f"(factor {w_cal:.4f})")
    out.append(f"# steps profiled: {steps}")
    out.append(f"{'kernel':64s} {'disp':>5s} {'FETCH KiB/disp':>15s} {'WRITE KiB/disp':>15s} "
               f"{'fetch B/disp':>14s} {'write B/disp':>14s}")
    lines = []
    for k in sorted(set(fetch) | set(write)):
        fv, wv = fetch.get(k, []), write.get(k, [])
        n = max(len(fv), len(wv))
        fa = sum(fv) / len(fv) if fv else 0.0
        wa = sum(wv) / len(wv) if wv else 0.0
        out.append(f"{k[:64]:64s} {n:5d} {fa:15.1f} {wa:15.1f} {fa * 1024 * f_cal:14.0f} {wa * 1024 * w_cal:14.0f}")
        lines.append(f"traffic\t{k}\t{n}\t{fa * 1024 * f_cal:.0f}\t{wa * 1024 * w_cal:.0f}")
    return "\n".join(out + ["# machine-readable: traffic <kernel> <dispatches> <fetch bytes/disp> <write bytes/disp>"]
                     + lines) + "\n"


def load_traffic(path: str, match, steps_key="# steps profiled:"):
    """Calibrated HBM bytes per step (fetch + write) of the kernels whose name
    satisfies `match`, from a pmc summary; None when the file is absent."""
    if not os.path.exists(path):
        return None
    steps = 1
    total = 0.0
    found = False
    with open(path) as f:
        for line in f:
            if line.startswith(steps_key):
                steps = int(line.split(":")[1])
            elif line.startswith("traffic\t"):
                _, k, n, fb, wb = line.rstrip("\n").split("\t")
                if match(k):
                    total += int(n) * (float(fb) + float(wb))
                    found = True
    return total / steps if found else None


def pmc_all(dirs, pick: str = "last"):
    """{kernel: {counter: value}} over several --pmc runs (each counter from
    the run that collected it): per kernel its last dispatch, or with
    pick="largest" the dispatch whose counters sum highest in each run (the
    one big launch of a run that also makes small ones)."""
    out = defaultdict(dict)
    for d in dirs:
        for path in _files(d, "_results.db"):
            con = sqlite3.connect(path)
            per = defaultdict(lambda: defaultdict(float))
            for name, disp, c, v in con.execute("select name, dispatch_id, counter_name, counter_value "
                                                "from pmc_events"):
                per[(short_name(name), disp)][c] += float(v)
            con.close()
            last = {}
            for (k, disp), cv in per.items():
                key = sum(cv.values()) if pick == "largest" else disp
                if k not in last or key > last[k][0]:
                    last[k] = (key, cv)
            for k, (_, cv) in last.items():
                out[k].update(cv)
    return out


DERIVED = [
    ("LDS bank-conflict cycles / LDS active cycles", "SQ_LDS_BANK_CONFLICT", "SQ_LDS_IDX_ACTIVE"),
    ("LDS wait cycles / wave cycles", "SQ_WAIT_INST_LDS", "SQ_WAVE_CYCLES"),
    ("L2 hit rate (hits / requests)", "TCC_HIT_sum", "TCC_REQ_sum"),
]


def format_counters(data, kernels, commands, notes=(), pick: str = "last"):
    which = "last dispatch" if pick == "last" else "largest dispatch"
    out = [f"# rocprofv3 --pmc counters per kernel ({which} of each kernel in the run)",
           "# command(s), run from the repository root:"]
    out += [f"#   {c}" for c in commands]
    out += [f"# {n}" for n in notes]
    out.append("# bytes from requests: HBM read = TCC_EA0_RDREQ_128B x 128 + the other read requests x 64;"
               " HBM write = TCC_EA0_WRREQ_64B x 64 + the other write requests x 32")
    for k in sorted(data):
        if kernels and not any(p in k for p in kernels):
            continue
        cv = data[k]
        out.append(k[:100])
        for c in sorted(cv):
            out.append(f"   {c:34s} {cv[c]:18.0f}")
        if "TCC_EA0_RDREQ_sum" in cv and "TCC_EA0_RDREQ_128B_sum" in cv:
            rd = cv["TCC_EA0_RDREQ_128B_sum"] * 128 + (cv["TCC_EA0_RDREQ_sum"] - cv["TCC_EA0_RDREQ_128B_sum"]) * 64
            out.append(f"   {'= HBM read bytes':34s} {rd:18.0f}")
        if "TCC_EA0_WRREQ_sum" in cv and "TCC_EA0_WRREQ_64B_sum" in cv:
            wr = cv["TCC_EA0_WRREQ_64B_sum"] * 64 + (cv["TCC_EA0_WRREQ_sum"] - cv["TCC_EA0_WRREQ_64B_sum"]) * 32
            out.append(f"   {'= HBM write bytes':34s} {wr:18.0f}")
        for label, a, b in DERIVED:
            if a in cv and b in cv and cv[b]:
                out.append(f"   = {label}: {cv[a] / cv[b]:.3f}")
    return "\n".join(out) + "\n"


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--command", action="append", default=[])
    p = sub.add_parser("pmc")
    p.add_argument("fetch_dir")
    p.add_argument("write_dir")
    p.add_argument("--copy-bytes", type=int, required=True)
    p.add_argument("--steps", type=int, default=1)
    p.add_argument("--command", action="append", default=[])
    q = sub.add_parser("counters")
    q.add_argument("dirs", nargs="+")
    q.add_argument("--kernel", action="append", default=[])
    q.add_argument("--command", action="append", default=[])
    q.add_argument("--note", action="append", default=[])
    q.add_argument("--pick", choices=["last", "largest"], default="last")
    a = ap.parse_args()
    if a.cmd == "counters":
        print(format_counters(pmc_all(a.dirs, a.pick), a.kernel, a.command, a.note, a.pick), end="")
    elif a.cmd == "stats":
        print(format_stats(kernel_stats(a.dir), a.command), end="")
    else:
        print(pmc_summary(a.fetch_dir, a.write_dir, a.copy_bytes, a.steps, a.command), end="")


if __name__ == "__main__":
    main()
