"""Summarise rocprofv3 result databases (rocpd SQLite) as text.

    python prof_summary.py trace  <trace_results.db>            per-kernel stats
    python prof_summary.py counter <results.db> [kernel-substring]   per-kernel
        mean of each collected counter per dispatch
"""
import sqlite3
import sys


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


if __name__ == "__main__":
    if sys.argv[1] == "trace":
        trace(sys.argv[2])
    else:
        counter(sys.argv[2], sys.argv[3] if len(sys.argv) > 3 else "")
