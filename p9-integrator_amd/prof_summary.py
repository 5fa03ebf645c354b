"""Summarise rocprofv3 results as text.

    python prof_summary.py trace  <trace_results.db>            per-kernel stats (rocpd SQLite)
    python prof_summary.py counter <results.db> [kernel-substring]   per-kernel
        mean of each collected counter per dispatch (rocpd SQLite)
    python prof_summary.py csvtrace <kernel_trace.csv> [kernel-substring]
        per-kernel stats from a CSV kernel trace, and the duration of every
        dispatch of the kernels matching the substring
    python prof_summary.py csvcounter <counter_collection.csv> [kernel-substring]
        every dispatch's counter values (CSV), for the kernels matching
    python prof_summary.py csvcountermean <counter_collection.csv> [kernel-substring]
        per-kernel mean of each counter per dispatch (CSV)
    python prof_summary.py csvtimeline <kernel_trace.csv> <kernel-substring> [index] [pad-ms]
        the dispatches around one dispatch of the matching kernels (default
        the last), with start and end relative to it and their queues
"""
import csv
import sqlite3
import sys
from collections import defaultdict


def short(name):
    """Kernel name without the argument list (and without anonymous namespaces)."""
    return name.replace("(anonymous namespace)::", "").split("(")[0]


def trace(db):
    cur = sqlite3.connect(db).cursor()
    print("| kernel | calls | total ms | avg ms | % |")
    print("|---|---|---|---|---|")
    for name, n, tot, avg, pct in cur.execute(
            "select name,total_calls,total_duration,average,percentage from top_kernels"):
        print(f"| `{short(name)}` | {n} | {tot / 1e3:.3f} | {avg / 1e3:.4f} | {pct:.2f} |")
    for (name, vg, ag, sg, scr) in cur.execute(
            "select distinct name, vgpr_count, accum_vgpr_count, sgpr_count, scratch_size from kernels"):
        print(f"resources `{short(name)}`: vgpr {vg} agpr {ag} sgpr {sg} scratch {scr}")


def counter(db, filt=""):
    cur = sqlite3.connect(db).cursor()
    rows = cur.execute("select kernel_name, counter_name, avg(value), count(value), avg(duration) "
                       "from counters_collection group by kernel_name, counter_name").fetchall()
    print("| kernel | counter | mean per dispatch | dispatches |")
    print("|---|---|---|---|")
    for k, c, v, n, d in rows:
        if filt in k:
            print(f"| `{short(k)}` | {c} | {v:.6g} | {n} |")


def csvtrace(path, filt=""):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    agg = defaultdict(list)
    res = {}
    for r in rows:
        k = short(r["Kernel_Name"])
        agg[k].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6)
        res[k] = (r["VGPR_Count"], r["SGPR_Count"], r["Scratch_Size"], r["LDS_Block_Size"])
    tot = sum(sum(v) for v in agg.values())
    print("| kernel | calls | total ms | avg ms | % | vgpr / sgpr / scratch B / LDS B |")
    print("|---|---|---|---|---|---|")
    for k, v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        print(f"| `{k}` | {len(v)} | {sum(v):.3f} | {sum(v) / len(v):.4f} | {100 * sum(v) / tot:.2f} | "
              f"{' / '.join(res[k])} |")
    if filt:
        print()
        print(f"dispatches of kernels matching `{filt}` (ms, in order):")
        for k, v in agg.items():
            if filt in k:
                print(f"  `{k}`: " + ", ".join(f"{x:.4f}" for x in v))


def csvtimeline(path, anchor, which="-1", pad_ms="0.6"):
    """Every dispatch that starts within pad_ms of the which-th dispatch of the
    kernels matching anchor (default: the last one), in start order, with
    start and end relative to the anchor's start (ms) and its queue."""
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    hits = [r for r in rows if anchor in r["Kernel_Name"]]
    a = hits[int(which)]
    t0, t1 = int(a["Start_Timestamp"]), int(a["End_Timestamp"])
    pad = float(pad_ms) * 1e6
    print(f"| kernel | queue | start ms | end ms | ms |")
    print("|---|---|---|---|---|")
    for r in rows:
        s, e = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        if t0 - pad <= s <= t1 + pad:
            print(f"| `{short(r['Kernel_Name'])}` | {r.get('Queue_Id', r.get('Stream_Id', '?'))} | "
                  f"{(s - t0) / 1e6:.4f} | {(e - t0) / 1e6:.4f} | {(e - s) / 1e6:.4f} |")


def csvcounter(path, filt=""):
    rows = list(csv.DictReader(open(path)))
    per = defaultdict(dict)
    for r in rows:
        if filt in r["Kernel_Name"]:
            per[(int(r["Dispatch_Id"]), short(r["Kernel_Name"]))][r["Counter_Name"]] = float(r["Counter_Value"])
    for (d, k), cs in sorted(per.items()):
        print(f"dispatch {d} `{k}`: " + ", ".join(f"{c} {v:.6g}" for c, v in sorted(cs.items())))


def csvcountermean(path, filt=""):
    """Mean per dispatch of each counter, per kernel (a markdown table)."""
    rows = list(csv.DictReader(open(path)))
    per = defaultdict(lambda: defaultdict(dict))
    for r in rows:
        if filt in r["Kernel_Name"]:
            per[short(r["Kernel_Name"])][r["Counter_Name"]][int(r["Dispatch_Id"])] = float(r["Counter_Value"])
    print("| kernel | counter | mean per dispatch | dispatches |")
    print("|---|---|---|---|")
    for k in sorted(per):
        for c in sorted(per[k]):
            v = list(per[k][c].values())
            print(f"| `{k}` | {c} | {sum(v) / len(v):.6g} | {len(v)} |")


if __name__ == "__main__":
    cmd, path = sys.argv[1], sys.argv[2]
    filt = sys.argv[3] if len(sys.argv) > 3 else ""
    {"trace": lambda: trace(path), "counter": lambda: counter(path, filt), "csvtrace": lambda: csvtrace(path, filt),
     "csvcounter": lambda: csvcounter(path, filt), "csvcountermean": lambda: csvcountermean(path, filt),
     "csvtimeline": lambda: csvtimeline(path, filt, *sys.argv[4:])}[cmd]()
