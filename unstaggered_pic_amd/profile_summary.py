"""Summarise rocprofv3 CSV output (--kernel-trace --stats, --pmc
FETCH_SIZE / WRITE_SIZE, --output-format csv) into the markdown tables kept
with the sources.

  python -m unstaggered_pic_amd.profile_summary --trace DIR [--after NAME]
         [--nth-last N] [--steps K] [--fetch DIR] [--write DIR] [--copy-mib 1024]
         [--kernel SUBSTRING[,SUBSTRING]] [--json-out FILE] [--counters DIR,DIR]
         --command "..."

HBM traffic per dispatch = FETCH_SIZE x (copy bytes / FETCH_SIZE of the
calibration copy) + WRITE_SIZE x (copy bytes / WRITE_SIZE of the copy): the
bench's --calibrate-copy-mib copy is a kernel that reads and writes a known
number of bytes, so the two factors convert the counters to bytes.  With
a trace, each kernel's HBM rate is that traffic over its mean duration in
the summarised region.

FLOP rates come from the SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F{64,32} counters
(wavefront instructions: 64 lanes, 2 flops per FMA) over the GPU-active time
of the dispatch.
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


def _working(values):
    """The dispatches that did work: a launch of the device-driven step whose
    gate was closed returns at once (its workgroups still launch: about 1% of
    a 16.8M-particle sweep's duration) and reads below 5% of the kernel's
    median; it is left out of the means.  (The median, not the largest
    value: a kernel whose working launches differ in size, such as the retry
    sweep, keeps them all.)"""
    if not values:
        return values
    srt = sorted(values)
    med = srt[len(srt) // 2]
    return [v for v in values if v >= 0.05 * med] if med > 0.0 else values


def counters(dirname):
    rows = list(csv.DictReader(open(_one(dirname, "*counter_collection.csv"))))
    agg = collections.defaultdict(list)
    for r in rows:
        agg[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return {k: _working(v) for k, v in agg.items()}


def counters_by_name(dirname):
    """{kernel: {counter: (mean over the last 4 working dispatches, dispatches)}}."""
    rows = list(csv.DictReader(open(_one(dirname, "*counter_collection.csv"))))
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in rows:
        agg[r["Kernel_Name"]][r["Counter_Name"]].append(float(r["Counter_Value"]))
    out = {}
    for k, d in agg.items():
        out[k] = {}
        for c, v in d.items():
            w = _working(v)
            out[k][c] = (sum(w[-4:]) / len(w[-4:]), len(w))
    return out


def trace_stats(dirname, after, nth_last=1):
    """Per-kernel (calls, total ns) over the dispatches that start after the
    nth-last dispatch whose name contains `after`, from the kernel trace."""
    rows = list(csv.DictReader(open(_one(dirname, "*kernel_trace.csv"))))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    marks = [i for i, r in enumerate(rows) if after in r["Kernel_Name"]]
    tail = rows[marks[-min(nth_last, len(marks))] + 1:] if marks else rows
    agg = collections.defaultdict(lambda: [0, 0])
    durs = collections.defaultdict(list)
    for r in tail:
        d = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        agg[r["Kernel_Name"]][0] += 1
        agg[r["Kernel_Name"]][1] += d
        durs[r["Kernel_Name"]].append(d)
    span = int(tail[-1]["End_Timestamp"]) - int(tail[0]["Start_Timestamp"]) if tail else 0
    return agg, span, {k: _working(v) for k, v in durs.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True)
    ap.add_argument("--after", default=None,
                    help="summarise only the dispatches after the last one whose name contains this")
    ap.add_argument("--nth-last", type=int, default=1,
                    help="start after the nth-last dispatch matching --after instead of the last")
    ap.add_argument("--steps", type=int, default=0, help="time steps in the summarised region (per-step totals)")
    ap.add_argument("--json-out", default=None, help="write the per-dispatch HBM traffic of the kernels as JSON")
    ap.add_argument("--counters", default=None,
                    help="comma-separated --pmc output directories of the SQ/GRBM counter passes")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--kernel", default="k_p3_map")
    ap.add_argument("--command", default="")
    ap.add_argument("--top", type=int, default=12)
    a = ap.parse_args()
    mean_ns = {}
    if a.after:
        agg, span, work = trace_stats(a.trace, a.after, a.nth_last)
        # mean duration of the launches that did work (gated-off launches of
        # the device-driven step return at once)
        mean_ns = {short(k): sum(v) / len(v) for k, v in work.items()}
        busy = sum(v[1] for v in agg.values())
        where = "the last" if a.nth_last == 1 else f"the {a.nth_last}th-last"
        print(f"Dispatches after {where} `{a.after}`: GPU busy {busy / 1e6:.2f} ms over a span of {span / 1e6:.2f} ms"
              + (f" ({a.steps} steps: {busy / 1e6 / a.steps:.2f} ms busy per step)" if a.steps else "") + ".")
        print()
        print("| kernel | calls (gated off) | total ms | average us (working calls) | % of busy |"
              + (" ms per step |" if a.steps else ""))
        print("|---|---|---|---|---|" + ("---|" if a.steps else ""))
        for k, (n, t) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:a.top]:
            nw = len(work[k])
            calls = f"{n} ({n - nw})" if nw < n else f"{n}"
            line = f"| `{short(k)}` | {calls} | {t / 1e6:.2f} | {mean_ns[short(k)] / 1e3:.1f} | {100.0 * t / busy:.1f} |"
            if a.steps:
                line += f" {t / 1e6 / a.steps:.3f} |"
            print(line)
    else:
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
        print("| kernel | dispatches | fetch MB/dispatch | write MB/dispatch | HBM MB/dispatch |"
              + (" mean us (trace) | HBM GB/s |" if mean_ns else ""))
        print("|---|---|---|---|---|" + ("---|---|" if mean_ns else ""))
        out = {"copy_factors": {"fetch": sf, "write": sw}, "kernels": {}}
        for k in sorted(f):
            if not any(x in k for x in a.kernel.split(",")) or k not in w:
                continue
            fm = sum(f[k]) / len(f[k]) * sf * 1024 / 1e6
            wm = sum(w[k]) / len(w[k]) * sw * 1024 / 1e6
            line = f"| `{short(k)}` | {len(f[k])} | {fm:.1f} | {wm:.1f} | {fm + wm:.1f} |"
            if mean_ns:
                t = mean_ns.get(short(k))
                line += f" {t / 1e3:.1f} | {(fm + wm) * 1e6 / t:.0f} |" if t else " - | - |"
            print(line)
            out["kernels"][short(k)] = {"dispatches": len(f[k]), "fetch_bytes": fm * 1e6, "write_bytes": wm * 1e6}
        if a.json_out:
            import json
            with open(a.json_out, "w") as fh:
                json.dump(out, fh, indent=1)
    if a.counters:
        res = collections.defaultdict(dict)
        for d in a.counters.split(","):
            for k, vals in counters_by_name(d).items():
                if any(x in k for x in a.kernel.split(",")):
                    res[k].update(vals)
        print()
        print("| kernel | dispatches | VALU / wavefront | VMEM loads / wavefront | VMEM stores / wavefront | "
              "LDS / wavefront | VALU busy | GPU-active ms | FP64 TFLOP/s | FP32 TFLOP/s |")
        print("|---|---|---|---|---|---|---|---|---|---|")

        def flops(d, t):
            names = [f"SQ_INSTS_VALU_{op}_F{t}" for op in ("FMA", "MUL", "ADD", "TRANS")]
            if not all(n in d for n in names):
                return None
            return 64.0 * (2.0 * d[names[0]][0] + d[names[1]][0] + d[names[2]][0] + d[names[3]][0])

        for k in sorted(res):
            d = res[k]
            if "SQ_WAVES" not in d or "GRBM_GUI_ACTIVE" not in d:
                continue
            w = d["SQ_WAVES"][0]
            # GRBM_GUI_ACTIVE sums the 8 XCDs; a wave64 VALU op occupies a SIMD
            # (1024 per GPU) for 4 cycles; 2.4 GHz
            cyc = d["GRBM_GUI_ACTIVE"][0] / 8.0
            sec = cyc / 2.4e9

            def per_wave(n):
                return f"{d[n][0] / w:.0f}" if n in d else "-"

            busy = f"{d['SQ_INSTS_VALU'][0] * 4.0 / 1024.0 / cyc:.2f}" if "SQ_INSTS_VALU" in d else "-"
            f64, f32 = flops(d, 64), flops(d, 32)
            print(f"| `{short(k)}` | {d['SQ_WAVES'][1]} | {per_wave('SQ_INSTS_VALU')} | "
                  f"{per_wave('SQ_INSTS_VMEM_RD')} | {per_wave('SQ_INSTS_VMEM_WR')} | {per_wave('SQ_INSTS_LDS')} | "
                  f"{busy} | {sec * 1e3:.3f} | "
                  + (f"{f64 / sec / 1e12:.2f}" if f64 is not None else "-") + " | "
                  + (f"{f32 / sec / 1e12:.2f}" if f32 is not None else "-") + " |")
    if a.command:
        print()
        print("Command: `" + a.command + "`")


if __name__ == "__main__":
    main()
