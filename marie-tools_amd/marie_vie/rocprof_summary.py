"""Text summaries of rocprofv3 SQLite output for the operator kernels.

  python3 -m marie_vie.rocprof_summary stats DIR
      per-kernel calls, average / total duration in microseconds (as the
      profiler's top_kernels view reports them; --kernel-trace --stats run)
  python3 -m marie_vie.rocprof_summary traffic FETCH_DIR WRITE_DIR --calib BYTES [--prec fp64|fp32]
      HBM bytes per dispatch of the five operator passes from one FETCH_SIZE
      and one WRITE_SIZE run (KiB per dispatch), each counter scaled by the
      ratio of the known size to what it reads for the largest plain device
      copy (__amd_rocclr_copyBuffer) of the same run; prints
      bytes_per_matvec = the sum over the passes of one operator application.
  python3 -m marie_vie.rocprof_summary window DIR --every REGEX [--skip K]
      per-iteration breakdown of a solver loop from a --kernel-trace run: the
      iterations are the launches of the kernel matching REGEX after the
      first K of them; prints per iteration the span, the summed kernel time,
      the idle time between kernels and each kernel's time.

Only reads the profiler's SQLite views (top_kernels, pmc_events, kernels).
"""
from __future__ import annotations

import argparse
import collections
import glob
import os
import re
import sqlite3
import sys

# the five passes of one VIE operator application, in launch order: z forward
# from the packed vector, y forward, fused x pass, y inverse, z inverse with the
# VIE epilogue (template arguments <T, IN, OUT, INV, P, TX> / <T, PWL, ISK, ADJ, P>).
# The axis passes are the Stockham kernel (axis_pass_ct, FP64) or the two-step
# register FFT (axis_pass_rf, FP32); the generic-length kernels are not matched.
PASS_PATTERNS = [
    ("z-fwd (packed in)", r"axis_pass_(ct|rf)<{t}, 2, 0, false,"),
    ("y-fwd", r"axis_pass_(ct|rf)<{t}, 0, 0, false,"),
    ("x fused (fft, product, ifft)", r"x_fused_rf<{t}, "),
    ("y-inv", r"axis_pass_(ct|rf)<{t}, 0, 0, true,"),
    ("z-inv (VIE epilogue, packed out)", r"axis_pass_(ct|rf)<{t}, 0, 3, true,"),
]


def _db(d: str) -> sqlite3.Connection:
    files = glob.glob(os.path.join(d, "**", "*.db"), recursive=True)
    if not files:
        raise SystemExit(f"no rocprofv3 database under {d}")
    return sqlite3.connect(files[0])


def _short(name: str) -> str:
    s = name.replace("HIP_vector_type<double, 2u>", "double2").replace("HIP_vector_type<float, 2u>", "float2")
    s = re.sub(r"vie::vec2<(\w+)>::type", r"\1_2", s)
    return s if len(s) <= 110 else s[:107] + "..."


def stats(d: str, top: int = 25) -> str:
    con = _db(d)
    rows = list(con.execute("select name, total_calls, total_duration, average, percentage from top_kernels"))
    out = [f"{'kernel':110s} {'calls':>6s} {'avg_us':>10s} {'total_us':>12s} {'pct':>6s}"]
    for n, c, t, a, p in rows[:top]:
        out.append(f"{_short(n):110s} {c:6d} {a:10.2f} {t:12.1f} {p:6.2f}")
    return "\n".join(out)


def _pmc(d: str, counter: str):
    con = _db(d)
    per = collections.defaultdict(list)
    for name, disp, cn, val in con.execute("select name, dispatch_id, counter_name, counter_value from pmc_events"):
        if cn == counter:
            per[name].append((disp, val))
    return per


def _calibration(per, calib_bytes: float) -> float:
    copies = [v for name, lst in per.items() if "copyBuffer" in name for _, v in lst]
    if not copies:
        raise SystemExit("no __amd_rocclr_copyBuffer dispatch to calibrate against")
    return calib_bytes / (max(copies) * 1024.0)


def traffic(fetch_dir: str, write_dir: str, calib_bytes: float, prec: str) -> tuple[str, float]:
    t = "double" if prec == "fp64" else "float"
    fetch, write = _pmc(fetch_dir, "FETCH_SIZE"), _pmc(write_dir, "WRITE_SIZE")
    ff, fw = _calibration(fetch, calib_bytes), _calibration(write, calib_bytes)
    lines = [f"calibration: plain copy of {calib_bytes / 2**20:.0f} MiB reads FETCH_SIZE x{1 / ff:.4f}, "
             f"WRITE_SIZE x{1 / fw:.4f} of its bytes -> scale factors fetch {ff:.4f}, write {fw:.4f}",
             f"{'pass':34s} {'dispatches':>10s} {'fetch_MB':>10s} {'write_MB':>10s} {'total_MB':>10s}  kernel"]
    total = 0.0
    missing = []
    found = []
    for label, pat in PASS_PATTERNS:
        rx = re.compile(pat.format(t=t))
        names = [n for n in fetch if rx.search(_short(n)) or rx.search(n)]
        if not names:
            missing.append(label)
        else:
            found.append((label, names))
    if missing:
        # a sum over part of the passes is not the bytes of a matvec
        raise SystemExit(f"passes not found in {fetch_dir}: {', '.join(missing)}")
    for label, names in found:
        name = max(names, key=lambda n: len(fetch[n]))   # the width in use, not a tuning run
        fv = [v for _, v in fetch[name]]
        wv = [v for _, v in write.get(name, [])]
        fb = sum(fv) / len(fv) * 1024.0 * ff
        wb = (sum(wv) / len(wv) * 1024.0 * fw) if wv else 0.0
        total += fb + wb
        lines.append(f"{label:34s} {len(fv):10d} {fb / 1e6:10.1f} {wb / 1e6:10.1f} {(fb + wb) / 1e6:10.1f}  {_short(name)}")
    lines.append(f"bytes_per_matvec={total:.0f}")
    return "\n".join(lines), total


def counters(dirs: list[str], patterns: list[str]) -> str:
    """Mean per dispatch of every PMC counter in the runs `dirs` (one run per
    counter set), for the kernel variant with the most dispatches among the
    names matching each pattern (regular expressions on the short name)."""
    per = collections.defaultdict(lambda: collections.defaultdict(list))   # name -> counter -> values
    for d in dirs:
        con = _db(d)
        for name, cn, val in con.execute("select name, counter_name, counter_value from pmc_events"):
            per[_short(name)][cn].append(val)
    cols, heads = [], []
    for pat in patterns:
        rx = re.compile(pat)
        names = [n for n in per if rx.search(n)]
        if not names:
            raise SystemExit(f"no kernel matches {pat!r}")
        name = max(names, key=lambda n: max(len(v) for v in per[n].values()))
        cols.append(per[name])
        heads.append(name if len(name) <= 60 else name[:57] + "...")
    allc = sorted({c for col in cols for c in col})
    out = [f"{'counter':24s} " + "  ".join(f"{h:>60s}" for h in heads)]
    for c in allc:
        vals = [(sum(col[c]) / len(col[c])) if col.get(c) else float("nan") for col in cols]
        out.append(f"{c:24s} " + "  ".join(f"{v:60.4g}" for v in vals))
    return "\n".join(out)


def window(d: str, every: str, skip: int = 0, top: int = 16) -> str:
    con = _db(d)
    rows = list(con.execute("select name, start, end from kernels order by start"))
    marks = [i for i, r in enumerate(rows) if re.search(every, r[0])]
    if len(marks) < skip + 2:
        raise SystemExit(f"fewer than {skip + 2} launches of {every!r}")
    win = rows[marks[skip]:marks[-1]]
    n = len(marks) - 1 - skip
    span = rows[marks[-1]][1] - win[0][1]
    busy = collections.defaultdict(float)
    for name, t0, t1 in win:
        busy[_short(name)] += t1 - t0
    total = sum(busy.values())
    out = [f"iterations {n} (launches of {every!r} after the first {skip}): span {span / 1e6 / n:.3f} ms, "
           f"kernels {total / 1e6 / n:.3f} ms, idle {(span - total) / 1e6 / n:.3f} ms per iteration"]
    for k, v in sorted(busy.items(), key=lambda x: -x[1])[:top]:
        out.append(f"  {k:110s} {v / 1e6 / n:8.3f} ms")
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python3 -m marie_vie.rocprof_summary")
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("dir")
    a.add_argument("--top", type=int, default=25)
    b = sub.add_parser("traffic")
    b.add_argument("fetch_dir")
    b.add_argument("write_dir")
    b.add_argument("--calib", type=float, required=True, help="bytes of the calibration copy")
    b.add_argument("--prec", default="fp64", choices=["fp64", "fp32"])
    cc = sub.add_parser("counters")
    cc.add_argument("dirs", nargs="+")
    cc.add_argument("--kernel", action="append", required=True, help="regex of a kernel's short name")
    w = sub.add_parser("window")
    w.add_argument("dir")
    w.add_argument("--every", required=True, help="regex of the kernel launched once per iteration")
    w.add_argument("--skip", type=int, default=0, help="launches of it before the timed loop")
    args = ap.parse_args(argv)
    if args.cmd == "window":
        print(window(args.dir, args.every, args.skip))
    elif args.cmd == "stats":
        print(stats(args.dir, args.top))
    elif args.cmd == "counters":
        print(counters(args.dirs, args.kernel))
    else:
        text, _ = traffic(args.fetch_dir, args.write_dir, args.calib, args.prec)
        print(text)


if __name__ == "__main__":
    sys.exit(main())
