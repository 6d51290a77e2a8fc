"""Condense rocprofv3 CSV output into the short text summaries kept with the
library, and read the HBM-traffic figures back.

  python -m wvamd.rocprof_summary trace STATS_CSV OUT_MD --title T --command C
  python -m wvamd.rocprof_summary pmc FETCH_CSV WRITE_CSV OUT_MD --title T --command C
  python -m wvamd.rocprof_summary counters COUNTER_CSV OUT_MD --title T --command C

`trace` tabulates a --kernel-trace --stats run (calls, average and total device
time per kernel).  `pmc` tabulates FETCH_SIZE and WRITE_SIZE (KiB per
dispatch, one --pmc pass each) per kernel and scales them by what the same
counters read for the calibration kernel: an elementwise multiply of
64 Mi doubles (torch.mul(a, 2.0, out=b)) that reads and writes exactly
536870912 bytes.  The scaled per-dispatch bytes are also written as a JSON
block, which `hbm_traffic()` reads.  `counters` tabulates the mean per
dispatch of every counter of one --pmc pass per kernel (JSON block read by
`kernel_counters()`).
"""
import argparse
import csv
import json
import re
from collections import OrderedDict, defaultdict

CALIBRATION_BYTES = 536870912
CALIBRATION_KEY = "AUnaryFunctor<double, double, double, at::native::binary_internal::MulFunctor<double>"


def short_name(name):
    """'void wvamd::k_proj<float, 4>(args...)' -> 'k_proj<float, 4>'."""
    n = name.strip()
    if n.startswith("void "):
        n = n[5:]
    depth = 0
    for i, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            n = n[:i]
            break
    return n.replace("wvamd::", "")


def _is_calibration(name):
    return CALIBRATION_KEY in name


def kernel_stats(path):
    rows = []
    for r in csv.DictReader(open(path)):
        rows.append({"kernel": short_name(r["Name"]), "calls": int(r["Calls"]),
                     "avg_ms": float(r["AverageNs"]) * 1e-6, "total_ms": float(r["TotalDurationNs"]) * 1e-6,
                     "pct": float(r["Percentage"])})
    return rows


def counter_means(path):
    """{(kernel, grid): [mean KiB per dispatch, dispatches]} and the calibration value."""
    acc = defaultdict(list)
    calib = []
    for r in csv.DictReader(open(path)):
        v = float(r["Counter_Value"])
        if _is_calibration(r["Kernel_Name"]):
            calib.append(v)
            continue
        acc[(short_name(r["Kernel_Name"]), int(r["Grid_Size"]))].append(v)
    out = OrderedDict((k, [sum(v) / len(v), len(v)]) for k, v in acc.items())
    return out, (sum(calib) / len(calib) if calib else None)


def write_trace(stats_csv, out_md, title, command):
    rows = kernel_stats(stats_csv)
    lines = [f"# {title}", "", "Command (from the repository root on the GPU host):", "", "```", command, "```", "",
             "| kernel | calls | avg ms | total ms | % of device time |", "|---|---:|---:|---:|---:|"]
    for r in rows:
        lines.append(f"| `{r['kernel']}` | {r['calls']} | {r['avg_ms']:.4f} | {r['total_ms']:.3f} | {r['pct']:.2f} |")
    open(out_md, "w").write("\n".join(lines) + "\n")


def write_pmc(fetch_csv, write_csv, out_md, title, command):
    fetch, fcal = counter_means(fetch_csv)
    write, wcal = counter_means(write_csv)
    if not fcal or not wcal:
        raise SystemExit("calibration kernel not found in the counter runs")
    ffac = CALIBRATION_BYTES / 1024.0 / fcal
    wfac = CALIBRATION_BYTES / 1024.0 / wcal
    lines = [f"# {title}", "", "Commands (from the repository root on the GPU host), one counter per pass:", "", "```",
             command, "```", "",
             f"Calibration: elementwise multiply of 64 Mi doubles, {CALIBRATION_BYTES} bytes read and written. "
             f"FETCH_SIZE read {fcal:.1f} KiB -> fetch factor {ffac:.4f}; WRITE_SIZE read {wcal:.1f} KiB -> "
             f"write factor {wfac:.4f}.", "",
             "| kernel | grid | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | read MB (scaled) | written MB (scaled) |"
             " traffic MB |", "|---|---:|---:|---:|---:|---:|---:|---:|"]
    table = OrderedDict()
    for key in fetch:
        f, n = fetch[key]
        w = write.get(key, [0.0, 0])[0]
        rb, wb = f * 1024 * ffac, w * 1024 * wfac
        table[f"{key[0]}|{key[1]}"] = {"read_bytes": rb, "write_bytes": wb, "traffic_bytes": rb + wb}
        lines.append(f"| `{key[0]}` | {key[1]} | {n} | {f:.1f} | {w:.1f} | {rb / 1e6:.1f} | {wb / 1e6:.1f} |"
                     f" {(rb + wb) / 1e6:.1f} |")
    lines += ["", "```json", json.dumps({"fetch_factor": ffac, "write_factor": wfac, "kernels": table}, indent=1),
              "```"]
    open(out_md, "w").write("\n".join(lines) + "\n")


def write_counters(counter_csv, out_md, title, command):
    acc = defaultdict(lambda: defaultdict(list))
    for r in csv.DictReader(open(counter_csv)):
        if _is_calibration(r["Kernel_Name"]):
            continue
        acc[(short_name(r["Kernel_Name"]), int(r["Grid_Size"]))][r["Counter_Name"]].append(float(r["Counter_Value"]))
    names = sorted({c for d in acc.values() for c in d})
    lines = [f"# {title}", "", "Command (from the repository root on the GPU host):", "", "```", command, "```", "",
             "Mean per dispatch of each counter, summed over the device.", "",
             "| kernel | grid | " + " | ".join(names) + " |", "|---|---:|" + "---:|" * len(names)]
    table = OrderedDict()
    for key, d in acc.items():
        means = {c: sum(v) / len(v) for c, v in d.items()}
        table[f"{key[0]}|{key[1]}"] = means
        lines.append(f"| `{key[0]}` | {key[1]} | " + " | ".join(f"{means.get(c, 0.0):.0f}" for c in names) + " |")
    lines += ["", "```json", json.dumps({"kernels": table}, indent=1), "```"]
    open(out_md, "w").write("\n".join(lines) + "\n")


def _lookup(md_path, kernel_prefix):
    """The JSON entry of the one kernel whose short name starts with
    `kernel_prefix` in a summary written here; None if absent or ambiguous."""
    try:
        text = open(md_path).read()
    except OSError:
        return None
    m = re.search(r"```json\n(.*?)\n```", text, re.S)
    if not m:
        return None
    kernels = json.loads(m.group(1))["kernels"]
    hits = [v for k, v in kernels.items() if k.startswith(kernel_prefix)]
    if len(hits) != 1:
        return None
    return hits[0]


def hbm_traffic(md_path, kernel_prefix):
    """Scaled HBM bytes per dispatch (read, write, traffic) of a kernel in a
    summary written by write_pmc."""
    return _lookup(md_path, kernel_prefix)


def kernel_counters(md_path, kernel_prefix):
    """Mean counter values per dispatch of a kernel in a summary written by
    write_counters."""
    return _lookup(md_path, kernel_prefix)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="mode", required=True)
    t = sub.add_parser("trace")
    t.add_argument("stats_csv")
    t.add_argument("out_md")
    p = sub.add_parser("pmc")
    p.add_argument("fetch_csv")
    p.add_argument("write_csv")
    p.add_argument("out_md")
    c = sub.add_parser("counters")
    c.add_argument("counter_csv")
    c.add_argument("out_md")
    for s in (t, p, c):
        s.add_argument("--title", required=True)
        s.add_argument("--command", required=True)
    a = ap.parse_args()
    if a.mode == "trace":
        write_trace(a.stats_csv, a.out_md, a.title, a.command)
    elif a.mode == "pmc":
        write_pmc(a.fetch_csv, a.write_csv, a.out_md, a.title, a.command)
    else:
        write_counters(a.counter_csv, a.out_md, a.title, a.command)


if __name__ == "__main__":
    main()
