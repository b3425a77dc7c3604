"""Summarise rocprofv3 SQLite outputs (rocpd databases) of bench.py runs.

    python -m plens_amd.rocpd_summary --trace T.db [--fetch F.db --write W.db] [--copy-bytes B]

--trace: kernel-trace database -> per-kernel calls / average duration (us).
--fetch/--write: databases of `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs -> average KiB
per dispatch per kernel. The streaming copy kernel (k_stream_copy, --copy-bytes per copy, 1 GiB by
default) calibrates both counters: factor = expected KiB / measured KiB.
"""
import argparse
import sqlite3
from collections import defaultdict


def short(name):
    name = name.replace("(anonymous namespace)::", "")
    return name.split("(")[0].replace("void ", "")


def trace_stats(db):
    con = sqlite3.connect(db)
    rows = con.execute("select name, total_calls, average from top_kernels").fetchall()
    return {short(n): (c, a) for n, c, a in rows}  # rocpd reports the average in us


def counter_stats(db, counter):
    con = sqlite3.connect(db)
    acc = defaultdict(list)
    for name, val in con.execute("select kernel_name, value from counters_collection where counter_name=?",
                                 (counter,)):
        acc[short(name)].append(val)
    return {k: sum(v) / len(v) for k, v in acc.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-bytes", type=float, default=float(1 << 30))
    ap.add_argument("--n-dof", type=float, default=2048000.0)
    a = ap.parse_args()
    if a.trace:
        print("kernel                               calls   avg_us")
        for k, (c, us) in sorted(trace_stats(a.trace).items(), key=lambda kv: -kv[1][0] * kv[1][1]):
            print(f"{k:36s} {c:6d} {us:9.2f}")
    if a.fetch and a.write:
        fe, wr = counter_stats(a.fetch, "FETCH_SIZE"), counter_stats(a.write, "WRITE_SIZE")
        exp_kib = a.copy_bytes / 1024.0
        cf = exp_kib / fe["k_stream_copy"] if fe.get("k_stream_copy") else 1.0
        cw = exp_kib / wr["k_stream_copy"] if wr.get("k_stream_copy") else 1.0
        print(f"\ncalibration (k_stream_copy, {a.copy_bytes:.0f} B read + written per copy): "
              f"FETCH factor {cf:.4f}, WRITE factor {cw:.4f}")
        print("kernel                               FETCH_KiB   WRITE_KiB   bytes/DoF (calibrated)")
        for k in sorted(fe):
            f, w = fe[k] * cf, wr.get(k, 0.0) * cw
            print(f"{k:36s} {fe[k]:11.1f} {wr.get(k, 0.0):11.1f}   {(f + w) * 1024 / a.n_dof:8.1f}")


if __name__ == "__main__":
    main()
