"""Summaries of rocprofv3 output for the hot-path kernels.

    python -m wavefront-renderer_amd.profiling stats  <kernel_stats.csv> [--command "..."]
    python -m wavefront-renderer_amd.profiling traffic <fetch_dir> <write_dir> \\
        <copy_fetch_dir> <copy_write_dir> --copy-mib 1024 [--command "..."]
    python -m wavefront-renderer_amd.profiling pmc <run_dir> [--command "..."]

`stats` turns a `rocprofv3 --kernel-trace --stats --output-format csv` kernel table into a short
markdown summary. `traffic` turns two `--pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs of the bench
(per-dispatch KiB) into HBM bytes per launch for every kernel, each counter scaled by what it reads
for the library's streaming copy of a known size (copy reads and writes exactly copy-mib MiB).
The JSON it prints is what `bench.py` reports as `roofline.traffic`. `pmc` summarises any counter
run: the mean value per dispatch of every counter, per kernel (what `bench.py` reads for the
VALU-issue roofline of the traversal kernel).
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import sys
from pathlib import Path


def short_name(kernel: str) -> str:
    name = kernel.split("(")[0]
    for prefix in ("void ", "wf::"):
        name = name.replace(prefix, "")
    return name.strip()


def kernel_stats(path: Path, top: int = 25) -> list[dict]:
    rows = list(csv.DictReader(open(path)))
    out = []
    for r in rows[:top]:
        out.append({"kernel": short_name(r["Name"]), "calls": int(r["Calls"]),
                    "total_ms": float(r["TotalDurationNs"]) / 1e6,
                    "avg_us": float(r["AverageNs"]) / 1e3, "pct": float(r["Percentage"])})
    return out


def pmc_rows(run_dir: Path) -> list[tuple[str, str, int, float]]:
    """(kernel, counter, dispatch id, value) of a rocprofv3 --pmc run, from its csv files or its
    SQLite results database."""
    rows = []
    files = list(Path(run_dir).rglob("*counter_collection.csv"))
    for f in files:
        for r in csv.DictReader(open(f)):
            rows.append((short_name(r["Kernel_Name"]), r["Counter_Name"],
                         int(r.get("Dispatch_Id", 0) or 0), float(r["Counter_Value"])))
    if not files:
        import sqlite3
        dbs = list(Path(run_dir).rglob("*_results.db"))
        if not dbs:
            raise SystemExit(f"no counter results under {run_dir}")
        for db in dbs:
            con = sqlite3.connect(str(db))
            q = ("select kernel_name, counter_name, dispatch_id, sum(value) from "
                 "counters_collection group by kernel_name, counter_name, dispatch_id")
            rows += [(short_name(k), c, int(d), float(v)) for k, c, d, v in con.execute(q)]
    return rows


def pmc_per_kernel(run_dir: Path, counter: str | None = None) -> dict[str, tuple[int, float]]:
    """(dispatches, mean value per dispatch) per kernel of one counter (the run's only counter
    when None)."""
    acc: dict[str, list[float]] = collections.defaultdict(list)
    for k, c, _, v in pmc_rows(run_dir):
        if counter is None or c == counter:
            acc[k].append(v)
    return {k: (len(v), sum(v) / len(v)) for k, v in acc.items()}


def pmc_summary(run_dir: Path) -> dict:
    """Mean per dispatch of every counter, per kernel."""
    acc: dict[str, dict[str, list[float]]] = collections.defaultdict(lambda: collections.defaultdict(list))
    for k, c, _, v in pmc_rows(run_dir):
        acc[k][c].append(v)
    return {k: {"dispatches": max(len(v) for v in cs.values()),
                **{c: sum(v) / len(v) for c, v in sorted(cs.items())}}
            for k, cs in sorted(acc.items())}


def traffic(fetch: Path, write: Path, copy_fetch: Path, copy_write: Path, copy_mib: int) -> dict:
    copy_bytes = copy_mib * 1024 * 1024
    cf = pmc_per_kernel(copy_fetch)["k_stream_copy"][1] * 1024.0
    cw = pmc_per_kernel(copy_write)["k_stream_copy"][1] * 1024.0
    f_scale, w_scale = copy_bytes / cf, copy_bytes / cw
    fr, wr = pmc_per_kernel(fetch), pmc_per_kernel(write)
    kernels = {}
    for k in sorted(set(fr) & set(wr)):
        nb = fr[k][1] * 1024.0 * f_scale + wr[k][1] * 1024.0 * w_scale
        kernels[k] = {"dispatches": fr[k][0], "fetch_bytes": round(fr[k][1] * 1024.0 * f_scale),
                      "write_bytes": round(wr[k][1] * 1024.0 * w_scale),
                      "bytes_per_launch": round(nb)}
    return {"copy_bytes": copy_bytes, "fetch_scale": round(f_scale, 4),
            "write_scale": round(w_scale, 4), "kernels": kernels}


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("csv")
    a.add_argument("--command", default="")
    b = sub.add_parser("traffic")
    for k in ("fetch", "write", "copy_fetch", "copy_write"):
        b.add_argument(k)
    b.add_argument("--copy-mib", type=int, default=1024)
    b.add_argument("--command", default="")
    c = sub.add_parser("pmc")
    c.add_argument("run_dir")
    c.add_argument("--command", default="")
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        rows = kernel_stats(Path(args.csv))
        print(f"Command: `{args.command}`\n")
        print("| kernel | calls | total ms | avg us | % |\n|---|---|---|---|---|")
        for r in rows:
            print(f"| {r['kernel']} | {r['calls']} | {r['total_ms']:.2f} | {r['avg_us']:.2f} | "
                  f"{r['pct']:.2f} |")
    elif args.cmd == "pmc":
        json.dump({"command": args.command, "kernels": pmc_summary(Path(args.run_dir))},
                  sys.stdout, indent=1)
        print()
    else:
        t = traffic(Path(args.fetch), Path(args.write), Path(args.copy_fetch),
                    Path(args.copy_write), args.copy_mib)
        t["command"] = args.command
        json.dump(t, sys.stdout, indent=1)
        print()


if __name__ == "__main__":
    main()
