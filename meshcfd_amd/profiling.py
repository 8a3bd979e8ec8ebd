"""Summaries of rocprofv3 output (kernel stats and HBM byte counters).

Usage (on the output directories of rocprofv3 runs):

    python -m meshcfd_amd.profiling stats DIR
    python -m meshcfd_amd.profiling traffic FETCH_DIR WRITE_DIR COPY_BYTES [--json OUT]

`traffic` averages FETCH_SIZE and WRITE_SIZE (KiB per dispatch) per kernel and
scales each by the factor that makes the plain copy kernel `k_copy` of
COPY_BYTES read COPY_BYTES and write COPY_BYTES.
"""

from __future__ import annotations

import argparse
import csv
import json
import re
import sys
from collections import defaultdict
from pathlib import Path


def _short(name: str) -> str:
    m = re.search(r"(k_[a-z_0-9]+)", name)
    return m.group(1) if m else name[:60]


def _rows(directory: str, suffix: str):
    for path in sorted(Path(directory).rglob(f"*{suffix}")):
        with open(path, newline="") as fh:
            yield from csv.DictReader(fh)


def stats(directory: str) -> list[dict]:
    out = []
    for row in _rows(directory, "kernel_stats.csv"):
        out.append(
            {
                "kernel": _short(row.get("Name", "")),
                "calls": int(row.get("Calls", 0)),
                "total_ms": float(row.get("TotalDurationNs", 0)) / 1e6,
                "avg_ms": float(row.get("AverageNs", 0)) / 1e6,
                "percent": float(row.get("Percentage", 0)),
            }
        )
    return out


def counters(directory: str, counter: str) -> dict[str, float]:
    """Mean counter value per dispatch, by kernel."""
    acc = defaultdict(list)
    for row in _rows(directory, "counter_collection.csv"):
        if row.get("Counter_Name") != counter:
            continue
        acc[_short(row.get("Kernel_Name", ""))].append(float(row["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def all_counters(directory: str) -> dict[str, dict[str, float]]:
    """Mean value per dispatch of every collected counter, by kernel."""
    acc = defaultdict(lambda: defaultdict(list))
    for row in _rows(directory, "counter_collection.csv"):
        acc[_short(row.get("Kernel_Name", ""))][row.get("Counter_Name")].append(float(row["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in d.items()} for k, d in acc.items()}


def traffic(fetch_dir: str, write_dir: str, copy_bytes: int) -> dict:
    fetch = counters(fetch_dir, "FETCH_SIZE")
    write = counters(write_dir, "WRITE_SIZE")
    if "k_copy" not in fetch or "k_copy" not in write:
        raise SystemExit("k_copy dispatches not found: run bench.py --copy-probe under the same counters")
    f_scale = copy_bytes / (fetch["k_copy"] * 1024.0)
    w_scale = copy_bytes / (write["k_copy"] * 1024.0)
    kernels = {}
    for k in sorted(set(fetch) | set(write)):
        fb = fetch.get(k, 0.0) * 1024.0 * f_scale
        wb = write.get(k, 0.0) * 1024.0 * w_scale
        kernels[k.removeprefix("k_")] = round(fb + wb, 1)
    return {
        "unit": "bytes per dispatch (FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
        "fetch_scale": f_scale,
        "write_scale": w_scale,
        "kernels": kernels,
        "fetch_kib": fetch,
        "write_kib": write,
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("dir")
    c = sub.add_parser("counters")
    c.add_argument("dir")
    c.add_argument("--kernel", default=None)
    b = sub.add_parser("traffic")
    b.add_argument("fetch_dir")
    b.add_argument("write_dir")
    b.add_argument("copy_bytes", type=int)
    b.add_argument("--json")
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        rows = stats(args.dir)
        print(f"{'kernel':<22}{'calls':>8}{'total_ms':>12}{'avg_ms':>12}{'%':>8}")
        for r in rows:
            print(f"{r['kernel']:<22}{r['calls']:>8}{r['total_ms']:>12.3f}{r['avg_ms']:>12.5f}{r['percent']:>8.2f}")
        return 0
    if args.cmd == "counters":
        for k, d in sorted(all_counters(args.dir).items()):
            if args.kernel and args.kernel not in k:
                continue
            print(k, " ".join(f"{c}={v:.4g}" for c, v in sorted(d.items())))
        return 0
    t = traffic(args.fetch_dir, args.write_dir, args.copy_bytes)
    text = json.dumps(t, indent=1)
    if args.json:
        Path(args.json).write_text(text + "\n")
    print(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
