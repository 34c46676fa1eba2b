"""Summaries of rocprofv3 CSV output for the profiles/ folder.

  python3 -m msim.profsum stats DIR
      per-kernel table from DIR/*kernel_stats.csv (--kernel-trace --stats)
  python3 -m msim.profsum traffic --fetch DIR --write DIR --copy-fetch DIR --copy-write DIR \
      --copy-bytes N --agents N [--uniform] --workload TEXT --json OUT
      per-kernel HBM bytes per launch from two --pmc passes (FETCH_SIZE, WRITE_SIZE; KiB per
      dispatch), scaled by what the same counters read for movement-sim_amd/lib/copycal's
      copies of N bytes; writes a JSON table (read by bench.py) and prints a markdown table.

(Run with movement-sim_amd/ on sys.path.)
"""
import argparse
import csv
import glob
import json
import os
import re
import sys

from . import roofline as RL


def _one(pattern):
    hits = sorted(glob.glob(pattern))
    if not hits:
        raise SystemExit(f"no file matches {pattern}")
    return hits[0]


def short_name(name):
    """'void msim::k_collide<false, false>(...)' -> 'k_collide<false, false>'."""
    name = re.sub(r"^void ", "", name)
    name = name.split("(")[0]
    return name.replace("msim::", "")


def base_name(name):
    return short_name(name).split("<")[0]


def stats_table(d):
    rows = list(csv.DictReader(open(_one(os.path.join(d, "*kernel_stats.csv")))))
    out = ["| kernel | calls | avg us | min us | max us | % time |", "|---|---|---|---|---|---|"]
    for r in rows:
        out.append(f"| {short_name(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                   f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | {float(r['Percentage']):.2f} |")
    return "\n".join(out)


# Kernels whose first dispatch belongs to the handle's creation (the
# renumbering out of the caller's agent order) and is reported on its own.
CREATION_FIRST = ("k_reorder",)


def counter_means(d, counter):
    """{kernel base name: (mean KiB per dispatch, dispatches)} for one counter;
    the first dispatch of a CREATION_FIRST kernel is keyed '<name> (creation)'."""
    acc = {}
    rows = [r for r in csv.DictReader(open(_one(os.path.join(d, "*counter_collection.csv"))))
            if r["Counter_Name"] == counter]
    rows.sort(key=lambda r: int(r["Dispatch_Id"]))
    seen = set()
    for r in rows:
        k = base_name(r["Kernel_Name"])
        if k in CREATION_FIRST and k not in seen:
            seen.add(k)
            k += " (creation)"
        s, c = acc.get(k, (0.0, 0))
        acc[k] = (s + float(r["Counter_Value"]), c + 1)
    return {k: (s / c, c) for k, (s, c) in acc.items()}


def copy_scale(d, counter, copy_bytes):
    m = counter_means(d, counter)
    vals = [m[k][0] for k in ("copy8", "copy16") if k in m]
    if not vals:
        raise SystemExit(f"no copy8/copy16 dispatches in {d}")
    return copy_bytes / (1024.0 * sum(vals) / len(vals))


def traffic(a):
    fs = copy_scale(a.copy_fetch, "FETCH_SIZE", a.copy_bytes)
    ws = copy_scale(a.copy_write, "WRITE_SIZE", a.copy_bytes)
    f = counter_means(a.fetch, "FETCH_SIZE")
    w = counter_means(a.write, "WRITE_SIZE")
    kernels = {}
    for k in sorted(set(f) | set(w)):
        if not k.startswith("k_"):
            continue
        fb = f.get(k, (0.0, 0))[0] * 1024.0 * fs
        wb = w.get(k, (0.0, 0))[0] * 1024.0 * ws
        kernels[k] = {"fetch_bytes": fb, "write_bytes": wb, "bytes_per_launch": fb + wb,
                      "algorithmic_bytes": RL.kernel_bytes(k.split(" ")[0], a.uniform) * a.agents,
                      "dispatches": max(f.get(k, (0, 0))[1], w.get(k, (0, 0))[1])}
    out = {"workload": a.workload, "fetch_scale": fs, "write_scale": ws, "kernels": kernels}
    with open(a.json, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"FETCH_SIZE x {fs:.4f}, WRITE_SIZE x {ws:.4f}\n")
    print("| kernel | dispatches | fetch MB | write MB | traffic MB | algorithmic MB |")
    print("|---|---|---|---|---|---|")
    for k, v in kernels.items():
        print(f"| {k} | {v['dispatches']} | {v['fetch_bytes'] / 1e6:.1f} | {v['write_bytes'] / 1e6:.1f} | "
              f"{v['bytes_per_launch'] / 1e6:.1f} | {v['algorithmic_bytes'] / 1e6:.1f} |")


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    t = sub.add_parser("traffic")
    for k in ("fetch", "write", "copy_fetch", "copy_write", "workload", "json"):
        t.add_argument("--" + k.replace("_", "-"), required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--agents", type=int, required=True)
    t.add_argument("--uniform", action="store_true")
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        print(stats_table(a.dir))
    else:
        traffic(a)


if __name__ == "__main__":
    sys.exit(main())
