"""Summarises rocprofv3 result databases (kernel trace and PMC passes) into
the short text kept under profiles/.

    python -m psyclonebench_amd.profile_summary --trace T.db --fetch F.db \
        --write W.db --report "variant 13=k_fused_wave<16, 1>=3556769792" \
        --copy-bytes 402948144

Each --report names a label, a substring of the kernel name and the
algorithmic bytes of one launch; the summary prints lines
"[label] traffic_bytes_per_launch: B" that the benchmark reads back.

Calibration: k_copy3 (the unfused path's now<-after copy) reads and writes a
known number of bytes; FETCH_SIZE and WRITE_SIZE of every kernel are scaled
by (known bytes / counter value) of that copy kernel.
"""
from __future__ import annotations

import argparse
import sqlite3


def _short(name: str) -> str:
    base = name.replace("(anonymous namespace)::", "")
    if base.startswith("void "):
        base = base[5:]
    return base.split("(")[0].split("::")[-1]


def kernel_stats(db: str) -> list[tuple[str, int, float, float]]:
    """(name, calls, total_us, average_us) per kernel from top_kernels."""
    c = sqlite3.connect(db)
    rows = c.execute("select name, total_calls, total_duration, average from top_kernels")
    return [(r[0], int(r[1]), float(r[2]), float(r[3])) for r in rows]


def counter_means(db: str, counter: str) -> dict[str, tuple[int, float, float]]:
    c = sqlite3.connect(db)
    q = ("select kernel_name, count(*), avg(value), avg(duration) from counters_collection "
         "where counter_name=? group by kernel_name")
    return {r[0]: (int(r[1]), float(r[2]) * 1024.0, float(r[3])) for r in c.execute(q, (counter,))}


def main() -> None:
    p = argparse.ArgumentParser()
    p.add_argument("--trace", required=True)
    p.add_argument("--fetch", help="FETCH_SIZE pass (with --write and --copy-bytes)")
    p.add_argument("--write", help="WRITE_SIZE pass")
    p.add_argument("--report", action="append", default=[],
                   help="LABEL=KERNEL_SUBSTRING=ALG_BYTES_PER_LAUNCH (repeatable)")
    p.add_argument("--copy-bytes", type=float,
                   help="bytes k_copy3 reads (= writes) per launch")
    p.add_argument("--counters", action="append", default=[],
                   help="a PMC pass's result database (repeatable): mean value per "
                        "dispatch of each of its counters, per kernel")
    p.add_argument("--flops", action="append", default=[],
                   help="LABEL=KERNEL_SUBSTRING=FLOP_PER_LAUNCH (repeatable): achieved "
                        "TFLOP/s from the trace's average duration")
    a = p.parse_args()

    stats = kernel_stats(a.trace)
    print("## kernel trace (rocprofv3 --kernel-trace --stats)")
    print(f"{'kernel':<40} {'calls':>6} {'avg_us':>10} {'total_ms':>10}")
    for name, calls, total, avg in stats:
        print(f"{_short(name)[:40]:<40} {calls:>6} {avg:>10.2f} {total / 1e3:>10.3f}")

    for rep in a.flops:
        label, sub, flop = rep.split("=")
        rows = [(name, avg) for name, _, _, avg in stats if sub in name]
        if not rows:
            raise SystemExit(f"no kernel matching {sub!r} in the trace")
        name, dur_us = rows[0]
        print()
        print(f"[{label}] kernel: {_short(name)}")
        print(f"[{label}] flop_per_launch: {float(flop):.4g}")
        print(f"[{label}] avg_duration_us (kernel trace): {dur_us:.2f}")
        print(f"[{label}] achieved_TFLOPs: {float(flop) / (dur_us * 1e-6) / 1e12:.1f}")

    for db in a.counters:
        c = sqlite3.connect(db)
        rows = c.execute("select kernel_name, counter_name, count(*), avg(value) from "
                         "counters_collection group by kernel_name, counter_name").fetchall()
        names = sorted({r[1] for r in rows})
        kernels = sorted({r[0] for r in rows})
        val = {(r[0], r[1]): float(r[3]) for r in rows}
        print()
        print(f"## PMC counters, mean per dispatch ({', '.join(names)})")
        for k in kernels:
            cells = "  ".join(f"{n}={val.get((k, n), float('nan')):.4g}" for n in names)
            print(f"{_short(k)[:40]:<40} {cells}")

    if not a.report:
        return
    if not (a.fetch and a.write and a.copy_bytes):
        raise SystemExit("--report needs --fetch, --write and --copy-bytes")

    fetch = counter_means(a.fetch, "FETCH_SIZE")
    write = counter_means(a.write, "WRITE_SIZE")
    copy_f = [v for k, v in fetch.items() if "k_copy3" in k][0][1]
    copy_w = [v for k, v in write.items() if "k_copy3" in k][0][1]
    kf = a.copy_bytes / copy_f
    kw = a.copy_bytes / copy_w
    print()
    print("## HBM traffic (rocprofv3 --pmc FETCH_SIZE, --pmc WRITE_SIZE)")
    print(f"calibration on k_copy3 ({a.copy_bytes:.0f} B read, same written): "
          f"FETCH_SIZE x {kf:.3f}, WRITE_SIZE x {kw:.3f}")
    print(f"{'kernel':<40} {'read_MB':>10} {'write_MB':>10} {'dur_us':>10}")
    for k in sorted(fetch):
        f = fetch[k][1] * kf
        w = write.get(k, (0, 0.0, 0.0))[1] * kw
        print(f"{_short(k)[:40]:<40} {f / 1e6:>10.1f} {w / 1e6:>10.1f} {fetch[k][2] / 1e3:>10.1f}")

    for rep in a.report:
        label, sub, alg = rep.split("=")
        alg_bytes = float(alg)
        key = [k for k in fetch if sub in k][0]
        traffic = fetch[key][1] * kf + write[key][1] * kw
        dur_us = [avg for name, _, _, avg in stats if sub in name][0]
        print()
        print(f"[{label}] kernel: {_short(key)}")
        print(f"[{label}] algorithmic_bytes_per_launch: {alg_bytes:.0f}")
        print(f"[{label}] traffic_bytes_per_launch: {traffic:.0f}")
        print(f"[{label}] traffic_over_algorithmic: {traffic / alg_bytes:.3f}")
        print(f"[{label}] avg_duration_us (kernel trace): {dur_us:.2f}")
        print(f"[{label}] achieved_algorithmic_GBps: {alg_bytes / (dur_us * 1e3):.1f}")
        print(f"[{label}] achieved_traffic_GBps: {traffic / (dur_us * 1e3):.1f}")


if __name__ == "__main__":
    main()
