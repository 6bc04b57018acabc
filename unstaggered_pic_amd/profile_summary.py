"""Summarise rocprofv3 CSV output (--kernel-trace --stats, --pmc
FETCH_SIZE / WRITE_SIZE, --output-format csv) into the markdown tables kept
with the sources.

  python -m unstaggered_pic_amd.profile_summary --trace DIR [--fetch DIR]
         [--write DIR] [--copy-mib 1024] [--kernel SUBSTRING] --command "..."

HBM traffic per dispatch = FETCH_SIZE x (copy bytes / FETCH_SIZE of the
calibration copy) + WRITE_SIZE x (copy bytes / WRITE_SIZE of the copy): the
bench's --calibrate-copy-mib copy is a kernel that reads and writes a known
number of bytes, so the two factors convert the counters to bytes.
"""
import argparse
import collections
import csv
import glob
import os


def _one(dirname, pattern):
    fs = glob.glob(os.path.join(dirname, "**", pattern), recursive=True)
    if not fs:
        raise SystemExit(f"no {pattern} under {dirname}")
    return fs[0]


def short(name):
    for p in ("void ", "upic::(anonymous namespace)::", "upic::"):
        name = name.replace(p, "")
    return name.split("(")[0]


def counters(dirname):
    rows = list(csv.DictReader(open(_one(dirname, "*counter_collection.csv"))))
    agg = collections.defaultdict(list)
    for r in rows:
        agg[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return agg


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--kernel", default="k_p3_map")
    ap.add_argument("--command", default="")
    ap.add_argument("--top", type=int, default=12)
    a = ap.parse_args()
    rows = list(csv.DictReader(open(_one(a.trace, "*kernel_stats.csv"))))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    print("| kernel | calls | total ms | average us | % |")
    print("|---|---|---|---|---|")
    for r in rows[:a.top]:
        print(f"| `{short(r['Name'])}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.2f} | "
              f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} |")
    if a.fetch and a.write:
        copy = a.copy_mib * 1024.0  # KiB
        f, w = counters(a.fetch), counters(a.write)
        fc = max(v for k, vals in f.items() if "copyBuffer" in k for v in vals)
        wc = max(v for k, vals in w.items() if "copyBuffer" in k for v in vals)
        sf, sw = copy / fc, copy / wc
        print()
        print(f"Calibration copy of {a.copy_mib} MiB: FETCH_SIZE {fc:.0f} KiB (factor {sf:.4f}), "
              f"WRITE_SIZE {wc:.0f} KiB (factor {sw:.4f}).")
        print()
        print("| kernel | dispatches | fetch MB/dispatch | write MB/dispatch | HBM MB/dispatch |")
        print("|---|---|---|---|---|")
        for k in sorted(f):
            if a.kernel not in k or k not in w:
                continue
            fm = sum(f[k]) / len(f[k]) * sf * 1024 / 1e6
            wm = sum(w[k]) / len(w[k]) * sw * 1024 / 1e6
            print(f"| `{short(k)}` | {len(f[k])} | "
                  f"{fm:.1f} | {wm:.1f} | {fm + wm:.1f} |")
    if a.command:
        print()
        print("Command: `" + a.command + "`")


if __name__ == "__main__":
    main()
