"""Summaries of a rocprofv3 kernel trace (its SQLite results database) for the
profiles: top kernels by total time, and one kernel's dispatches inside the
closed-loop timed region that bench.py reports (roofline.timed_region_clock_ns,
host clock readings in ns at the region's start and end).

    python fnec-rust_amd/traceview.py RESULTS.db [--top 30]
    python fnec-rust_amd/traceview.py RESULTS.db --kernel k_lu_update_f16x2d --bench-log BENCH.log

The trace's timestamps are compared with each host clock the bench recorded;
the clock whose window contains dispatches of the kernel is used and named.
"""
import argparse
import json
import sqlite3


def top(db, n):
    rows = db.execute("select name, count(*), sum(end - start), avg(end - start) from kernels "
                      "group by name order by sum(end - start) desc limit ?", (n,)).fetchall()
    total = db.execute("select sum(end - start) from kernels").fetchone()[0] or 1
    out = ["| kernel | dispatches | total us | avg us | % of GPU time |", "|---|---:|---:|---:|---:|"]
    for name, cnt, tot, avg in rows:
        out.append(f"| `{short(name)}` | {cnt} | {tot / 1e3:.0f} | {avg / 1e3:.1f} | {100.0 * tot / total:.2f} |")
    return "\n".join(out)


def short(name):
    name = name.replace("fnec::", "")
    i = name.find("(")
    return name[:i] if i > 0 else name


def region(db, kernel, clocks):
    best = None
    for clk in ("monotonic", "boottime", "monotonic_raw"):
        t0, t1 = clocks["start"][clk], clocks["end"][clk]
        r = db.execute("select count(*), avg(end - start), min(start), max(end) from kernels "
                       "where name like ? and start >= ? and end <= ?", (f"%{kernel}(%", t0, t1)).fetchone()
        if r[0] and (best is None or r[0] > best[1][0]):
            best = (clk, r)
    if best is None:
        return None
    clk, (cnt, avg, first, last) = best
    return {"clock": clk, "dispatches": cnt, "avg_us": avg / 1e3, "first_ns": first, "last_ns": last}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--top", type=int, default=30)
    ap.add_argument("--kernel")
    ap.add_argument("--bench-log", help="bench.py output holding its JSON line")
    a = ap.parse_args()
    db = sqlite3.connect(a.db)
    if a.kernel:
        line = [x for x in open(a.bench_log) if x.startswith('{"metric"')][-1]
        rf = json.loads(line)["roofline"]
        r = region(db, a.kernel, rf["timed_region_clock_ns"])
        print(json.dumps({"kernel": a.kernel, "timed_region": r,
                          "bench_events_avg_us": rf["avg_launch_us"], "bench_launches": rf["launches"]}))
    else:
        print(top(db, a.top))


if __name__ == "__main__":
    main()
