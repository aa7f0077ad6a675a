"""Summarise rocprofv3 output databases (rocpd SQLite) as short text.

    python -m pyro_cmc_amd.rocprof_summary trace  <results.db> [--top N]
    python -m pyro_cmc_amd.rocprof_summary counter <results.db> <COUNTER>

`trace` prints per-kernel call count, total and average duration (us) from
a `rocprofv3 --kernel-trace --stats` run; `counter` prints the per-kernel
mean of one PMC counter per dispatch from a `rocprofv3 --pmc` run.
"""
from __future__ import annotations

import argparse
import re
import sqlite3


def _short(name: str) -> str:
    name = re.sub(r"pyro::\(anonymous namespace\)::", "", name)
    name = re.sub(r"\(pyro::.*$|\(float const\*.*$|\(double const\*.*$|\(unsigned.*$|\(int const\*.*$", "", name)
    return name.replace("void ", "")


def trace(db: str, top: int = 20) -> str:
    cur = sqlite3.connect(db).cursor()
    rows = list(cur.execute("select name, total_calls, total_duration, average, percentage "
                            "from top_kernels order by total_duration desc"))
    out = [f"{'kernel':<64} {'calls':>7} {'total_us':>12} {'avg_us':>10} {'pct':>6}"]
    for name, calls, tot, avg, pct in rows[:top]:
        out.append(f"{_short(name)[:64]:<64} {calls:>7} {tot:>12.1f} {avg:>10.3f} {pct:>6.2f}")
    return "\n".join(out)


def counter(db: str, cname: str) -> str:
    cur = sqlite3.connect(db).cursor()
    rows = list(cur.execute(
        "select name, count(*), avg(counter_value), min(counter_value), max(counter_value) "
        "from pmc_events where counter_name = ? group by name order by avg(counter_value) desc",
        (cname,)))
    out = [f"{'kernel':<64} {'dispatches':>10} {'mean':>14} {'min':>14} {'max':>14}   ({cname} per dispatch)"]
    for name, n, avg, mn, mx in rows:
        out.append(f"{_short(name)[:64]:<64} {n:>10} {avg:>14.1f} {mn:>14.1f} {mx:>14.1f}")
    return "\n".join(out)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["trace", "counter"])
    ap.add_argument("db")
    ap.add_argument("counter", nargs="?")
    ap.add_argument("--top", type=int, default=20)
    a = ap.parse_args()
    print(trace(a.db, a.top) if a.mode == "trace" else counter(a.db, a.counter))


if __name__ == "__main__":
    main()
