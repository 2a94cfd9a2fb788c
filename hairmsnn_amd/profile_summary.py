"""Summaries of rocprofv3 result databases (rocpd SQLite) as markdown tables
and, for HBM traffic, the per-kernel JSON that bench.py reads.

    python -m hairmsnn_amd.profile_summary trace RUN.db
    python -m hairmsnn_amd.profile_summary calls RUN.db KERNEL [KERNEL ...]
    python -m hairmsnn_amd.profile_summary pmc RUN.db
    python -m hairmsnn_amd.profile_summary flops RUN.db [RUN.db ...] --json out.json
    python -m hairmsnn_amd.profile_summary traffic --fetch F.db --write W.db \
        --copy-fetch CF.db --copy-write CW.db --copy-mib 1024 --json out.json

trace: calls, total and mean duration per kernel. calls: a kernel's
launches split into the library calls that made them (a streaming call
ends with its drain: launches below 5 % of the kernel's median duration,
followed by a full-size one), with the mean duration per call. pmc: the mean per launch
of every counter collected (summed over a dispatch's instances). traffic:
FETCH_SIZE and WRITE_SIZE (KiB per dispatch) scaled by what the same
counters read for the device-to-device copy kernel of --copy-mib MiB (which
reads and writes that many bytes), giving HBM bytes per launch. flops:
FP32 VALU operations per launch (SQ_INSTS_VALU_FLOPS_FP32 and its
transcendental counterpart count per wave instruction, so x 64 lanes) and
matrix-core operations (SQ_INSTS_VALU_MFMA_MOPS_* x 512), with the rate over
each launch's mean duration.
"""
from __future__ import annotations

import argparse
import collections
import json
import re
import sqlite3
import subprocess
import sys
from pathlib import Path

COPY_KERNEL = "__amd_rocclr_copyBuffer"


_DEMANGLED: dict[str, str] = {}


def _nested_name(name: str) -> str:
    """The function name and a template argument of a mangled nested name
    (_ZN<len><id>...I<arg>E...), for names c++filt leaves alone (_Float16
    arguments)."""
    if not name.startswith("_ZN"):
        return name
    i, ids = 3, []
    while i < len(name) and name[i].isdigit():
        j = i
        while j < len(name) and name[j].isdigit():
            j += 1
        n = int(name[i:j])
        ids.append(name[j:j + n])
        i = j + n
    if not ids:
        return name
    args = {"DF16_": "_Float16", "DF16b": "__bf16", "f": "float", "d": "double"}
    if name.startswith("I", i):
        for code, arg in args.items():
            if name.startswith(code, i + 1):
                return f"{ids[-1]}<{arg}>"
    return ids[-1]


def _demangle(name: str) -> str:
    if not name.startswith("_Z"):
        return name
    if name not in _DEMANGLED:
        out = name
        try:
            r = subprocess.run(["c++filt", name], stdout=subprocess.PIPE, text=True, timeout=10)
            out = r.stdout.strip() or name
        except (OSError, subprocess.SubprocessError):
            pass
        _DEMANGLED[name] = out if not out.startswith("_Z") else _nested_name(out)
    return _DEMANGLED[name]


def short_name(name: str) -> str:
    """Kernel name without its argument list, return type and namespace."""
    n = _demangle(name.strip()).replace("(anonymous namespace)::", "")
    if n.startswith("void "):
        n = n[5:]
    depth, cut = 0, len(n)
    for i, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = i
            break
    n = n[:cut]
    n = re.sub(r"\bhm::", "", n)
    if n.startswith("rocprim::"):
        n = "rocprim " + ("radix sort" if "radix_sort" in n else n.split("::")[-1][:40])
    return n


def _connect(path: str) -> sqlite3.Connection:
    if not Path(path).exists():
        raise SystemExit(f"no such database: {path}")
    return sqlite3.connect(path)


def kernel_times(db: str) -> dict[str, tuple[int, float]]:
    """kernel -> (calls, total ns)."""
    con = _connect(db)
    out: dict[str, list] = collections.defaultdict(lambda: [0, 0.0])
    for name, dur in con.execute("select name, duration from kernels"):
        k = short_name(name)
        out[k][0] += 1
        out[k][1] += float(dur)
    return {k: (v[0], v[1]) for k, v in out.items()}


def counters(db: str) -> dict[str, dict[str, float]]:
    """kernel -> counter -> mean per launch; '_launches' and '_us' (mean duration)."""
    con = _connect(db)
    per_dispatch: dict[tuple[int, str], dict[str, float]] = collections.defaultdict(dict)
    dur: dict[int, float] = {}
    q = ("select dispatch_id, kernel_name, counter_name, sum(value), max(duration) from counters_collection "
         "group by dispatch_id, counter_name")
    for d, k, c, v, du in con.execute(q):
        per_dispatch[(d, short_name(k))][c] = float(v)
        dur[d] = float(du)
    sums: dict[str, dict[str, float]] = collections.defaultdict(lambda: collections.defaultdict(float))
    for (d, k), cs in per_dispatch.items():
        s = sums[k]
        s["_launches"] += 1
        s["_us"] += dur[d] / 1e3
        for c, v in cs.items():
            s[c] += v
    out = {}
    for k, s in sums.items():
        n = s["_launches"]
        out[k] = {c: (v / n if c != "_launches" else v) for c, v in s.items()}
    return out


def cmd_trace(a) -> None:
    kt = kernel_times(a.db)
    total = sum(v[1] for v in kt.values()) or 1.0
    print("| kernel | calls | total ms | avg us | % |")
    print("|---|---:|---:|---:|---:|")
    for k, (n, ns) in sorted(kt.items(), key=lambda kv: -kv[1][1])[: a.top]:
        print(f"| {k} | {n} | {ns / 1e6:.3f} | {ns / n / 1e3:.1f} | {100 * ns / total:.1f} |")


def cmd_calls(a) -> None:
    con = _connect(a.db)
    rows = [(short_name(n), float(d)) for n, d in con.execute("select name, duration from kernels order by start")]
    print("| kernel | call | launches | total ms | avg us |")
    print("|---|---:|---:|---:|---:|")
    for k in a.kernels:
        d = [x for n, x in rows if n == k]
        if not d:
            print(f"| {k} | - | 0 | - | - |")
            continue
        med = sorted(d)[len(d) // 2]
        calls = [[d[0]]]
        for prev, cur in zip(d, d[1:]):
            if prev < 0.05 * med and cur > 0.5 * med:
                calls.append([])
            calls[-1].append(cur)
        for i, c in enumerate(calls):
            print(f"| {k} | {i} | {len(c)} | {sum(c) / 1e6:.3f} | {sum(c) / len(c) / 1e3:.1f} |")


def cmd_pmc(a) -> None:
    cs = counters(a.db)
    names = sorted({c for v in cs.values() for c in v if not c.startswith("_")})
    print("| kernel | launches | avg us | " + " | ".join(names) + " |")
    print("|---|---:|---:|" + "---:|" * len(names))
    for k, v in sorted(cs.items(), key=lambda kv: -kv[1]["_us"] * kv[1]["_launches"])[: a.top]:
        vals = " | ".join(f"{v.get(c, float('nan')):.4g}" for c in names)
        print(f"| {k} | {int(v['_launches'])} | {v['_us']:.1f} | {vals} |")


def cmd_flops(a) -> None:
    per: dict[str, dict[str, float]] = {}
    for db in a.dbs:
        for k, v in counters(db).items():
            per.setdefault(k, {}).update(v)
    print("| kernel | launches | avg us | VALU FP32 GFLOP / launch | VALU TFLOP/s | MFMA GFLOP / launch | MFMA TFLOP/s |")
    print("|---|---:|---:|---:|---:|---:|---:|")
    out = {}
    for k, v in sorted(per.items(), key=lambda kv: -kv[1]["_us"] * kv[1]["_launches"]):
        valu = 64.0 * (v.get("SQ_INSTS_VALU_FLOPS_FP32", 0.0) + v.get("SQ_INSTS_VALU_FLOPS_FP32_TRANS", 0.0))
        mfma = 512.0 * sum(x for c, x in v.items() if c.startswith("SQ_INSTS_VALU_MFMA_MOPS_"))
        if valu == 0 and mfma == 0:
            continue
        us = v["_us"]
        out[k] = {"launches": int(v["_launches"]), "avg_us": us, "valu_flops_per_launch": valu,
                  "mfma_flops_per_launch": mfma}
        if len(out) <= a.top:
            print(f"| {k} | {int(v['_launches'])} | {us:.1f} | {valu / 1e9:.3f} | {valu / (us * 1e-6) / 1e12:.1f} | "
                  f"{mfma / 1e9:.3f} | {mfma / (us * 1e-6) / 1e12:.1f} |")
    if a.json:
        Path(a.json).write_text(json.dumps({"workload": a.workload, "kernels": out}, indent=1) + "\n")


def cmd_traffic(a) -> None:
    fetch, write = counters(a.fetch), counters(a.write)
    cf, cw = counters(a.copy_fetch), counters(a.copy_write)
    copy_bytes = a.copy_mib * 2.0 ** 20
    f_kib = cf[COPY_KERNEL]["FETCH_SIZE"]
    w_kib = cw[COPY_KERNEL]["WRITE_SIZE"]
    fetch_scale = copy_bytes / (f_kib * 1024.0)
    write_scale = copy_bytes / (w_kib * 1024.0)
    print(f"The copy kernel ({COPY_KERNEL}) reads FETCH_SIZE = {f_kib / 1024:.1f} MiB and WRITE_SIZE = "
          f"{w_kib / 1024:.1f} MiB per {a.copy_mib} MiB copy, so FETCH_SIZE is scaled by {fetch_scale:.3f} and "
          f"WRITE_SIZE by {write_scale:.3f} below.\n")
    print("| kernel | launches | avg us | FETCH_SIZE MiB (raw) | read MB (scaled) | WRITE_SIZE MiB (raw) | "
          "written MB (scaled) | HBM GB/s |")
    print("|---|---:|---:|---:|---:|---:|---:|---:|")
    out = {}
    for k in sorted(fetch, key=lambda k: -fetch[k]["_us"] * fetch[k]["_launches"]):
        if k == COPY_KERNEL or k not in write:
            continue
        fr = fetch[k]["FETCH_SIZE"]
        wr = write[k]["WRITE_SIZE"]
        rb = fr * 1024.0 * fetch_scale
        wb = wr * 1024.0 * write_scale
        us = fetch[k]["_us"]
        out[k] = {"launches": int(fetch[k]["_launches"]), "read_bytes_per_launch": rb,
                  "write_bytes_per_launch": wb, "traffic_bytes_per_launch": rb + wb, "avg_us": us}
        if len(out) <= a.top:
            print(f"| {k} | {int(fetch[k]['_launches'])} | {us:.1f} | {fr / 1024:.1f} | {rb / 1e6:.1f} | "
                  f"{wr / 1024:.1f} | {wb / 1e6:.1f} | {(rb + wb) / (us * 1e-6) / 1e9:.0f} |")
    if a.json:
        Path(a.json).write_text(json.dumps({"workload": a.workload, "fetch_scale": fetch_scale,
                                            "write_scale": write_scale, "kernels": out}, indent=1) + "\n")


def cmd_mfma(a) -> None:
    """Per kernel: launches and mean duration from a kernel trace, HBM bytes
    from FETCH_SIZE / WRITE_SIZE runs (FETCH_SIZE times --fetch-scale, the
    copy calibration), matrix-core FLOP per launch and the fraction of cycles
    the matrix cores were busy. rocprofv3 sums GRBM_GUI_ACTIVE over the 8 XCDs
    and SQ_VALU_MFMA_BUSY_CYCLES over all SIMDs, so busy = MFMA busy cycles /
    (GRBM_GUI_ACTIVE / 8 x SIMDs)."""
    kt = kernel_times(a.trace)
    fetch, write, mf = counters(a.fetch), counters(a.write), counters(a.mfma)
    print(f"| kernel | calls | avg us | read MB (FETCH_SIZE x {a.fetch_scale:.2f}) | written MB | "
          "MFMA FLOP (MOPS x 512) | MFMA busy % |")
    print("|---|---:|---:|---:|---:|---:|---:|")
    for k, (n, ns) in sorted(kt.items(), key=lambda kv: -kv[1][1])[: a.top]:
        rd = fetch.get(k, {}).get("FETCH_SIZE", float("nan")) * 1024.0 * a.fetch_scale / 1e6
        wr = write.get(k, {}).get("WRITE_SIZE", float("nan")) * 1024.0 / 1e6
        m = mf.get(k, {})
        flop = 512.0 * sum(x for c, x in m.items() if c.startswith("SQ_INSTS_VALU_MFMA_MOPS_"))
        gui = m.get("GRBM_GUI_ACTIVE", 0.0)
        busy = 100.0 * m.get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0) / (gui / 8.0 * a.simds) if gui else float("nan")
        print(f"| {k} | {n} | {ns / n / 1e3:.1f} | {rd:.1f} | {wr:.1f} | {flop / 1e9:.2f} G | {busy:.1f} |")


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("db")
    t.add_argument("--top", type=int, default=20)
    c = sub.add_parser("calls")
    c.add_argument("db")
    c.add_argument("kernels", nargs="+")
    p = sub.add_parser("pmc")
    p.add_argument("db")
    p.add_argument("--top", type=int, default=12)
    fl = sub.add_parser("flops")
    fl.add_argument("dbs", nargs="+")
    fl.add_argument("--json")
    fl.add_argument("--workload", default="")
    fl.add_argument("--top", type=int, default=12)
    mm = sub.add_parser("mfma")
    for n in ("--trace", "--fetch", "--write", "--mfma"):
        mm.add_argument(n, required=True)
    mm.add_argument("--fetch-scale", type=float, default=2.0)
    mm.add_argument("--simds", type=int, default=1024)
    mm.add_argument("--top", type=int, default=16)
    f = sub.add_parser("traffic")
    for n in ("--fetch", "--write", "--copy-fetch", "--copy-write"):
        f.add_argument(n, required=True)
    f.add_argument("--copy-mib", type=int, required=True)
    f.add_argument("--json")
    f.add_argument("--workload", default="")
    f.add_argument("--top", type=int, default=12)
    a = ap.parse_args(argv)
    {"trace": cmd_trace, "calls": cmd_calls, "pmc": cmd_pmc, "flops": cmd_flops, "traffic": cmd_traffic,
     "mfma": cmd_mfma}[a.cmd](a)


if __name__ == "__main__":
    main(sys.argv[1:])
