"""Markdown summaries of rocprofv3 CSV output for this library's kernels.

    python -m cppwrf_amd.profile_summary trace <out_dir>
        per-kernel calls, total / average / min / max duration from a --kernel-trace run, FP64 and FP32-bulk apart
    python -m cppwrf_amd.profile_summary traffic <fetch_dir> <write_dir> --grid NXxNYxNZ
        HBM bytes per dispatch from one --pmc FETCH_SIZE run and one --pmc WRITE_SIZE run, each counter scaled by
        what it reads for k_copy_step, a copy of known size (u, v, w, t, ph and the rows of mu), and a
        "traffic:" line per kernel and precision (the most-dispatched launch shape first) that bench.py reads

The mode of a kernel is its namespace (mi_fp64 / mi_fp32); its shape the template arguments of the name.
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import re
import sys
from collections import defaultdict


def _rows(d: str, suffix: str):
    files = sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True))
    if not files:
        raise SystemExit(f"no *{suffix} under {d}")
    for f in files:
        with open(f, newline="") as fh:
            yield from csv.DictReader(fh)


def kernel_key(name: str):
    """(short name, mode, shape) of a demangled kernel name"""
    mode = "fp64" if "mi_fp64" in name else ("fp32" if "mi_fp32" in name else "-")
    m = re.search(r"k_([a-z0-9_]+)", name)
    short = m.group(1) if m else name.split("(")[0][-40:]
    t = re.search(r"k_[a-z0-9_]+<([^>]*)>", name)
    shape = t.group(1).replace(", ", "x").replace(",", "x") if t else "-"
    return short, mode, shape


def trace(d: str) -> str:
    stats = defaultdict(list)
    for r in _rows(d, "kernel_trace.csv"):
        name = r.get("Kernel_Name", "")
        dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3  # us
        stats[kernel_key(name)].append(dur)
    total = sum(sum(v) for v in stats.values())
    lines = ["| kernel | mode | calls | total ms | avg us | min us | max us | % |", "|---|---|---:|---:|---:|---:|---:|---:|"]
    for (k, mode, shape), v in sorted(stats.items(), key=lambda kv: -sum(kv[1])):
        lines.append(f"| `{k}<{shape}>` | {mode} | {len(v)} | {sum(v) / 1e3:.2f} | {sum(v) / len(v):.2f} | {min(v):.2f} | "
                     f"{max(v):.2f} | {100 * sum(v) / total:.1f} |")
    by_mode = defaultdict(float)
    for (k, mode, shape), v in stats.items():
        by_mode[mode] += sum(v)
    lines.append("")
    lines.append("GPU kernel time by mode: " + ", ".join(f"{m} {t / 1e3:.1f} ms" for m, t in sorted(by_mode.items())))
    return "\n".join(lines)


def _counter(d: str, counter: str):
    per = defaultdict(list)
    for r in _rows(d, "counter_collection.csv"):
        if r.get("Counter_Name") != counter:
            continue
        per[kernel_key(r.get("Kernel_Name", ""))].append(float(r["Counter_Value"]))
    return per


def traffic(fetch_dir: str, write_dir: str, grid: str) -> str:
    from .serving import alloc_elems
    nx, ny, nz = (int(v) for v in grid.lower().split("x"))
    n3 = alloc_elems(nx, ny, nz)
    f, w = _counter(fetch_dir, "FETCH_SIZE"), _counter(write_dir, "WRITE_SIZE")
    n2 = n3 // (nz + 1)  # the rows of a 2-D field (level 0 of each j block)
    out = [f"Calibration: `k_copy_step` copies u, v, w, t, ph ({n3:,} elements each, halos included) and the {n2:,} "
           f"doubles of mu's rows."]
    scale = {}
    for mode, esize in (("fp32", 4), ("fp64", 8)):
        known = n3 * 5 * esize + n2 * 8
        kf, kw = f.get(("copy_step", mode, "-")), w.get(("copy_step", mode, "-"))
        if kf and kw:
            sf = known / (sum(kf) / len(kf) * 1024.0)
            sw = known / (sum(kw) / len(kw) * 1024.0)
            scale[mode] = (sf, sw)
            out.append(f"{mode}: {known:,} B read and written; FETCH_SIZE {sum(kf) / len(kf):.1f} KiB, WRITE_SIZE "
                       f"{sum(kw) / len(kw):.1f} KiB per dispatch: scale factors fetch x{sf:.3f}, write x{sw:.3f}.")
    out += ["", "| kernel | mode | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes per dispatch (scaled) |",
            "|---|---|---:|---:|---:|---:|"]
    rows = []
    for key in sorted(set(f) & set(w)):
        k, mode, shape = key
        if mode not in scale:
            continue
        fv, wv = sum(f[key]) / len(f[key]), sum(w[key]) / len(w[key])
        b = (fv * scale[mode][0] + wv * scale[mode][1]) * 1024.0
        rows.append((len(f[key]), k, mode, shape, fv, wv, b))
    rows.sort(key=lambda r: -r[0] * r[6])
    for n, k, mode, shape, fv, wv, b in rows:
        out.append(f"| `{k}<{shape}>` | {mode} | {n} | {fv:.1f} | {wv:.1f} | {b:,.0f} |")
    out.append("")
    seen = set()
    for n, k, mode, shape, fv, wv, b in sorted(rows, key=lambda r: -r[0]):
        if (k, mode) in seen:
            continue
        seen.add((k, mode))
        out.append(f"    traffic: kernel={k} precision={mode} shape={shape} bytes={b:.0f}")
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("dir")
    c = sub.add_parser("traffic")
    c.add_argument("fetch_dir")
    c.add_argument("write_dir")
    c.add_argument("--grid", required=True)
    a = ap.parse_args(argv)
    print(trace(a.dir) if a.cmd == "trace" else traffic(a.fetch_dir, a.write_dir, a.grid))


if __name__ == "__main__":
    sys.exit(main())
