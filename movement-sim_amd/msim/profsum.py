"""Summaries of rocprofv3 CSV output for the profiles/ folder.

  python3 -m msim.profsum stats DIR
      per-kernel table from DIR/*kernel_stats.csv (--kernel-trace --stats)
  python3 -m msim.profsum traffic --rdreq DIR --write DIR --copy-rdreq DIR --copy-write DIR \
      --copy-bytes N --agents N [--uniform] --workload TEXT --json OUT
      per-kernel bytes per launch between the L2s and the fabric from two --pmc passes:
      TCC_EA0_RDREQ_{32B,64B,128B}_sum (read requests by size) and WRITE_SIZE (KiB), each
      checked against what the same counters read for movement-sim_amd/lib/copycal's copies of
      N bytes (and its random 8-byte gathers); writes a JSON table (read by bench.py) and prints
      a markdown table.
  python3 -m msim.profsum sq DIR [--kernels REGEX]
      per-kernel SQ summary from one --pmc pass of SQ_INSTS_VALU, SQ_WAVES,
      SQ_ACTIVE_INST_VALU, SQ_BUSY_CYCLES, SQ_INSTS_VMEM_RD, SQ_INSTS_LDS: instructions per
      wave and the VALU issue fraction (4 cycles per wave64 VALU instruction on each of the
      1024 SIMDs, over the kernel's busy cycles).

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


RDREQ = (("TCC_EA0_RDREQ_32B_sum", 32), ("TCC_EA0_RDREQ_64B_sum", 64), ("TCC_EA0_RDREQ_128B_sum", 128))


def read_bytes(d):
    """{kernel: (mean bytes read per dispatch, dispatches)} from the read-request counters."""
    acc = {}
    for counter, size in RDREQ:
        for k, (v, c) in counter_means(d, counter).items():
            s, _ = acc.get(k, (0.0, c))
            acc[k] = (s + v * size, c)
    return acc


def copy_scale(d, counter, copy_bytes):
    m = counter_means(d, counter)
    vals = [m[k][0] for k in ("copy8", "copy16") if k in m]
    if not vals:
        raise SystemExit(f"no copy8/copy16 dispatches in {d}")
    return copy_bytes / (1024.0 * sum(vals) / len(vals))


def traffic(a):
    cr = read_bytes(a.copy_rdreq)
    copies = [cr[k][0] for k in ("copy8", "copy16") if k in cr]
    if not copies:
        raise SystemExit(f"no copy8/copy16 dispatches in {a.copy_rdreq}")
    fs = a.copy_bytes / (sum(copies) / len(copies))
    ws = copy_scale(a.copy_write, "WRITE_SIZE", a.copy_bytes)
    gather = cr.get("gather8", (0.0, 0))[0] / (a.copy_bytes / 8) if "gather8" in cr else None
    f = read_bytes(a.rdreq)
    w = counter_means(a.write, "WRITE_SIZE")
    kernels = {}
    for k in sorted(set(f) | set(w)):
        if not k.startswith("k_"):
            continue
        fb = f.get(k, (0.0, 0))[0] * fs
        wb = w.get(k, (0.0, 0))[0] * 1024.0 * ws
        kernels[k] = {"read_bytes": fb, "write_bytes": wb, "bytes_per_launch": fb + wb,
                      "algorithmic_bytes": RL.kernel_bytes(k.split(" ")[0], a.uniform) * a.agents,
                      "dispatches": max(f.get(k, (0, 0))[1], w.get(k, (0, 0))[1])}
    out = {"workload": a.workload, "read_scale": fs, "write_scale": ws, "gather8_bytes_per_element": gather,
           "kernels": kernels}
    with open(a.json, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"read requests x {fs:.4f}, WRITE_SIZE x {ws:.4f}" +
          (f"; a random 8-byte gather reads {gather:.1f} bytes" if gather else "") + "\n")
    print("| kernel | dispatches | read MB | write MB | traffic MB | algorithmic MB |")
    print("|---|---|---|---|---|---|")
    for k, v in kernels.items():
        print(f"| {k} | {v['dispatches']} | {v['read_bytes'] / 1e6:.1f} | {v['write_bytes'] / 1e6:.1f} | "
              f"{v['bytes_per_launch'] / 1e6:.1f} | {v['algorithmic_bytes'] / 1e6:.1f} |")


SQ = ("SQ_INSTS_VALU", "SQ_WAVES", "SQ_ACTIVE_INST_VALU", "SQ_BUSY_CYCLES", "SQ_INSTS_VMEM_RD", "SQ_INSTS_LDS")
N_SIMD, N_SE = 1024, 32  # MI355X: 256 CUs x 4 SIMDs; 32 shader engines


def sq_table(d, pattern):
    m = {c: counter_means(d, c) for c in SQ}
    rx = re.compile(pattern)
    out = ["| kernel | dispatches | waves | VALU / wave | VMEM reads / wave | LDS / wave | busy cycles | "
           "VALU issue fraction |", "|---|---|---|---|---|---|---|---|"]
    for k in sorted(m["SQ_WAVES"]):
        if not k.startswith("k_") or not rx.search(k):
            continue
        waves, n = m["SQ_WAVES"][k]
        busy = m["SQ_BUSY_CYCLES"][k][0] / N_SE
        valu = m["SQ_INSTS_VALU"][k][0]
        frac = 4.0 * m["SQ_ACTIVE_INST_VALU"][k][0] / (N_SIMD * busy) if busy else 0.0
        out.append(f"| {k} | {n} | {waves:.0f} | {valu / waves:.0f} | {m['SQ_INSTS_VMEM_RD'][k][0] / waves:.1f} | "
                   f"{m['SQ_INSTS_LDS'][k][0] / waves:.1f} | {busy:.0f} | {frac:.2f} |")
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    t = sub.add_parser("traffic")
    for k in ("rdreq", "write", "copy_rdreq", "copy_write", "workload", "json"):
        t.add_argument("--" + k.replace("_", "-"), required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--agents", type=int, required=True)
    t.add_argument("--uniform", action="store_true")
    q = sub.add_parser("sq")
    q.add_argument("dir")
    q.add_argument("--kernels", default=".")
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        print(stats_table(a.dir))
    elif a.cmd == "sq":
        print(sq_table(a.dir, a.kernels))
    else:
        traffic(a)


if __name__ == "__main__":
    sys.exit(main())
