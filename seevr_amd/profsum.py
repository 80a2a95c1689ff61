"""Kernel-time and HBM-byte summaries of rocprofv3 runs of bench.py, over one
timed window of the run.

bench.py --window-file PATH writes, per section, the CLOCK_MONOTONIC
nanoseconds that bracket its timed frames (rocprofv3 stamps dispatches on
the same clock).  This module keeps the dispatches that start inside the
chosen window and writes a Markdown summary: kernel time from a
`--kernel-trace` run, FETCH_SIZE and WRITE_SIZE per dispatch from two
`--pmc` runs of the same command (each counter in its own run), and the
plain copy kernel's counters (k_copy_calib, run once by `bench.py --calib`
outside every window) as the calibration bench.py scales the counters by.

    python -m seevr_amd.profsum --window plain --title "..." \\
        --trace DIR --fetch DIR --write DIR --out profiles/NAME.md \\
        --command "rocprofv3 ..." [--command ...] [--note "..."]

Each DIR is a rocprofv3 -d directory holding run_kernel_trace.csv or
run_counter_collection.csv (--output-format csv) and the window.json the
bench wrote there.
"""
from __future__ import annotations

import argparse
import csv
import json
import re
from collections import defaultdict
from pathlib import Path

CALIB = "k_copy_calib"


def short_name(name: str) -> str:
    """'void sv::(anonymous namespace)::k_os_pass<512, 16>(unsigned int const*, ...)' -> 'k_os_pass<512, 16>'."""
    n = name.replace("(anonymous namespace)::", "").replace("(anonymous namespace)", "")
    n = re.sub(r"^void\s+", "", n)
    depth = 0
    for i, ch in enumerate(n):  # cut the argument list (the first '(' outside template brackets)
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            n = n[:i]
            break
    # the last '::' component outside template brackets ('k_refit<sv::Box3>' stays whole)
    depth, cut = 0, 0
    for i, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == ":" and depth == 0 and n[i:i + 2] == "::":
            cut = i + 2
    return n[cut:].strip()


def load_window(d: Path, window: str) -> tuple[int, int]:
    w = json.loads((d / "window.json").read_text())
    if window not in w:
        raise SystemExit(f"{d}/window.json has no window '{window}' (has {sorted(w)})")
    t0, t1 = w[window]
    return int(t0), int(t1)


def rows(d: Path, stem: str) -> list[dict]:
    files = sorted(d.rglob(f"*{stem}.csv"))
    if not files:
        raise SystemExit(f"no *{stem}.csv under {d}")
    out = []
    for f in files:
        with open(f, newline="") as fh:
            out.extend(csv.DictReader(fh))
    return out


def kernel_times(d: Path, window: str) -> dict[str, list[float]]:
    t0, t1 = load_window(d, window)
    per = defaultdict(list)
    for r in rows(d, "kernel_trace"):
        s, e = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        if t0 <= s <= t1:
            per[short_name(r["Kernel_Name"])].append((e - s) / 1000.0)
    return per


def counters(d: Path, window: str) -> tuple[dict[str, list[float]], list[float]]:
    """Per kernel the counter value (KiB) of each dispatch in the window, and
    the calibration kernel's values over the whole run."""
    t0, t1 = load_window(d, window)
    per, cal = defaultdict(list), []
    for r in rows(d, "counter_collection"):
        k = short_name(r["Kernel_Name"])
        v = float(r["Counter_Value"])
        if k == CALIB:
            cal.append(v)
        elif t0 <= int(r["Start_Timestamp"]) <= t1:
            per[k].append(v)
    return per, cal


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--window", required=True)
    ap.add_argument("--title", required=True)
    ap.add_argument("--trace", type=Path, required=True)
    ap.add_argument("--fetch", type=Path)
    ap.add_argument("--write", type=Path)
    ap.add_argument("--out", type=Path, required=True)
    ap.add_argument("--command", action="append", default=[])
    ap.add_argument("--note", action="append", default=[])
    ap.add_argument("--top", type=int, default=30)
    a = ap.parse_args()

    kt = kernel_times(a.trace, a.window)
    total = sum(sum(v) for v in kt.values())
    lines = [f"# {a.title}", "", "Commands (MI355X, one GPU; run from /tmp with TMPDIR=/tmp; R = repo root):", ""]
    lines += [f"    {c}" for c in a.command]
    lines += [""] + [n for n in a.note] + ([""] if a.note else [])
    lines += [f"Summarized by `python3 -m seevr_amd.profsum --window {a.window} --trace DIR/kt --fetch DIR/fetch "
              "--write DIR/write`.", "",
              f"Window `{a.window}`: dispatches that start inside the bench's timed frames of that section "
              "(CLOCK_MONOTONIC bounds the bench writes to window.json).", "",
              "## Kernel time (kernel trace, dispatches in the window)", "",
              "| kernel | calls | total us | avg us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
    for k, v in sorted(kt.items(), key=lambda kv: -sum(kv[1]))[: a.top]:
        lines.append(f"| {k} | {len(v)} | {sum(v):.0f} | {sum(v) / len(v):.2f} | {min(v):.2f} | {max(v):.2f} | "
                     f"{100 * sum(v) / total:.1f} |")
    if a.fetch and a.write:
        f, fc = counters(a.fetch, a.window)
        w, wc = counters(a.write, a.window)
        lines += ["", "## HBM bytes per dispatch (FETCH_SIZE and WRITE_SIZE runs, dispatches in the window)", "",
                  "| kernel | dispatches | FETCH_SIZE KiB/dispatch | WRITE_SIZE KiB/dispatch |", "|---|---|---|---|"]
        if fc and wc:
            lines.append(f"| {CALIB} | {len(fc)} | {sum(fc) / len(fc):.1f} | {sum(wc) / len(wc):.1f} |")
        for k in sorted(f, key=lambda k: -sum(kt.get(k, [0.0]))):
            if k in w and f[k] and w[k]:
                lines.append(f"| {k} | {len(f[k])} | {sum(f[k]) / len(f[k]):.1f} | {sum(w[k]) / len(w[k]):.1f} |")
        if fc and wc:
            true_kib = (256 << 20) / 1024.0
            lines += ["", f"Calibration: {CALIB} reads and writes 256 MiB ({true_kib:.0f} KiB each way); its counters "
                      f"read {sum(fc) / len(fc):.0f} and {sum(wc) / len(wc):.0f} KiB, so bytes = FETCH x "
                      f"{true_kib / (sum(fc) / len(fc)):.2f} + WRITE x {true_kib / (sum(wc) / len(wc)):.2f}."]
    a.out.write_text("\n".join(lines) + "\n")
    print(a.out)


if __name__ == "__main__":
    main()
