"""Summarise rocprofv3 output databases (rocpd SQLite) as short text.

    python -m pyro_cmc_amd.rocprof_summary trace  <results.db> [--top N]
    python -m pyro_cmc_amd.rocprof_summary counter <results.db> <COUNTER>
    python -m pyro_cmc_amd.rocprof_summary traffic <fetch.db> <write.db> --kernel REGEX
           --tag TAG [--fetch-scale F] [--write-scale W]
    python -m pyro_cmc_amd.rocprof_summary events <bench output>

`trace` prints per-kernel call count, total and average duration (us) from
a `rocprofv3 --kernel-trace --stats` run; `counter` prints the per-kernel
mean of one PMC counter per dispatch from a `rocprofv3 --pmc` run.
`traffic` combines a FETCH_SIZE run and a WRITE_SIZE run (KiB per dispatch)
into HBM bytes per launch of the kernels matching REGEX (dispatch-weighted
mean), each counter scaled by what it reads for a plain copy of known size,
and prints "traffic tag=TAG bytes_per_launch=B".
`events` reads the JSON line `bench.py` prints and tabulates its HIP-event
timing of the headline round: the sweep kernel's back-to-back launch sample
(the source of roofline.avg_launch_us and frac), the launches per round
against the measured round time, and the per-family event times of the
profiled rounds.
"""
from __future__ import annotations

import argparse
import json
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


def _mean_kib(db: str, cname: str, kernel: str) -> tuple[float, int]:
    cur = sqlite3.connect(db).cursor()
    rows = list(cur.execute("select name, counter_value from pmc_events where counter_name = ?", (cname,)))
    vals = [v for name, v in rows if re.search(kernel, _short(name))]
    if not vals:
        raise SystemExit(f"no {cname} dispatches of kernels matching {kernel!r} in {db}")
    return sum(vals) / len(vals), len(vals)


def traffic(fetch_db: str, write_db: str, kernel: str, tag: str, fetch_scale: float,
            write_scale: float) -> str:
    f, nf = _mean_kib(fetch_db, "FETCH_SIZE", kernel)
    w, nw = _mean_kib(write_db, "WRITE_SIZE", kernel)
    total = (f * fetch_scale + w * write_scale) * 1024.0
    return (f"FETCH_SIZE {f:.1f} KiB x {fetch_scale} ({nf} dispatches) + WRITE_SIZE {w:.1f} KiB x "
            f"{write_scale} ({nw} dispatches) per launch of /{kernel}/\n"
            f"traffic tag={tag} bytes_per_launch={total:.0f}")


def events(path: str) -> str:
    line = None
    for text in open(path):
        text = text.strip()
        if text.startswith("{") and '"metric"' in text:
            line = json.loads(text)
    if line is None:
        raise SystemExit(f"no bench.py result line in {path}")
    rf = line["roofline"]
    lt = rf["launch_timing"]
    avg_us = 1e3 * lt["elapsed_ms"] / lt["launches"]
    per_round = lt["launches_per_round"]
    # compulsory bytes per site-update x site updates per launch, and the
    # same bytes per site visited (a fused stage reads and writes a spin
    # once for two updates) x sites visited per launch
    b_unit = rf["bytes_per_update"]
    n_upd = rf["updates_per_launch"]
    n_vis = rf.get("sites_visited_per_launch", n_upd)
    gbs = b_unit * n_upd / (avg_us * 1e-6) / 1e9
    gbs_v = b_unit * n_vis / (avg_us * 1e-6) / 1e9
    out = [
        f"sweep kernel {rf['kernel']} in the timed rounds: {lt['launches']} launches over {lt['rounds']} rounds "
        f"in {lt['elapsed_ms']:.3f} ms (HIP events around each round's sweep launches, profiling mode) -> "
        f"{avg_us:.3f} us per launch",
        f"round: {per_round:g} sweep launches x {avg_us:.3f} us = {per_round * avg_us / 1e3:.4f} ms of the measured "
        f"{line['ms_per_step']:.4f} ms per round (wall clock over {line['steps']} rounds)",
        f"roofline: {b_unit} B x {n_upd:.0f} site updates / {avg_us:.3f} us = "
        f"{gbs:.1f} GB/s = {gbs / rf['peak']:.4f} of {rf['peak']:.0f} GB/s",
        f"  per site visited: {b_unit} B x {n_vis:.0f} sites / {avg_us:.3f} us = "
        f"{gbs_v:.1f} GB/s = {gbs_v / rf['peak']:.4f} of {rf['peak']:.0f} GB/s",
    ]
    bb = rf.get("back_to_back")
    if bb:
        out.append(f"back to back (plain variant, {bb['sweeps']} sweeps, no per-launch events): "
                   f"{bb['launches']} launches in {bb['elapsed_ms']:.3f} ms -> {bb['avg_launch_us']:.3f} us per launch")
    out += [
        "",
        f"per-family HIP-event times over {line.get('kernel_profiled_rounds', '?')} profiled rounds "
        "(events around each family's launches):",
        f"{'family':<12} {'launches':>10} {'total_ms':>12} {'avg_us':>10} {'us/round':>10}",
    ]
    n_rounds = line.get("kernel_profiled_rounds")
    for fam, ms in line["kernel_ms"].items():
        n = line["kernel_launches"][fam]
        if not n:
            continue
        per = f"{1e3 * ms / n_rounds:10.2f}" if n_rounds else f"{'':>10}"
        out.append(f"{fam:<12} {n:>10} {ms:>12.3f} {1e3 * ms / n:>10.3f} {per}")
    return "\n".join(out)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["trace", "counter", "traffic", "events"])
    ap.add_argument("db")
    ap.add_argument("arg", nargs="?", help="COUNTER (counter mode) or the WRITE_SIZE db (traffic mode)")
    ap.add_argument("--top", type=int, default=20)
    ap.add_argument("--kernel", default=".")
    ap.add_argument("--tag", default="kernel")
    ap.add_argument("--fetch-scale", type=float, default=2.0)
    ap.add_argument("--write-scale", type=float, default=1.0)
    a = ap.parse_args()
    if a.mode == "trace":
        print(trace(a.db, a.top))
    elif a.mode == "counter":
        print(counter(a.db, a.arg))
    elif a.mode == "events":
        print(events(a.db))
    else:
        print(traffic(a.db, a.arg, a.kernel, a.tag, a.fetch_scale, a.write_scale))


if __name__ == "__main__":
    main()
