"""Summarise rocprofv3 SQLite outputs (rocpd databases) of bench.py runs.

    python -m plens_amd.rocpd_summary --trace T.db [--fetch F.db --write W.db] [--copy-bytes B]

--trace: kernel-trace database -> per-kernel calls / average duration (us).
--fetch/--write: databases of `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs -> average KiB
per dispatch per kernel. The streaming copy kernel (k_stream_copy, --copy-bytes per copy, 1 GiB by
default) calibrates both counters: factor = expected KiB / measured KiB. The calibrated HBM bytes
per dispatch of every kernel are also printed as `traffic_bytes_per_launch <kernel> <bytes>` lines.
--pmc: database of any other `rocprofv3 --pmc` run -> average value per dispatch of every counter.
--flops: with --pmc holding SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 (and, for the FP32-bulk path,
the same _F32 counters) and --trace, the FP64 and FP32 VALU flop rates per kernel (64 lanes per wave
instruction, FMA = 2 flops) against the 78.6 TFLOP/s FP64 and 157.3 TFLOP/s FP32 peaks (the FP32
peak counts packed FP32 instructions, which this build does not emit: its FP32 issue rate is that
of FP64, 78.6 TFLOP/s).
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
    ap.add_argument("--pmc")
    ap.add_argument("--flops", action="store_true")
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
        print()
        for k in sorted(fe):
            f, w = fe[k] * cf, wr.get(k, 0.0) * cw
            print(f"traffic_bytes_per_launch {k.replace(' ', '')} {(f + w) * 1024:.0f}")
    if a.pmc:
        con = sqlite3.connect(a.pmc)
        names = [r[0] for r in con.execute("select distinct counter_name from counters_collection")]
        per = {n: counter_stats(a.pmc, n) for n in names}
        kernels = sorted({k for d in per.values() for k in d})
        print("\nkernel                               " + " ".join(f"{n:>22s}" for n in names))
        for k in kernels:
            print(f"{k:36s} " + " ".join(f"{per[n].get(k, 0.0):22.1f}" for n in names))
        if a.flops and a.trace:
            tr = trace_stats(a.trace)
            for prec, peak in (("F64", 78.6e12), ("F32", 157.3e12)):
                if f"SQ_INSTS_VALU_FMA_{prec}" not in per:
                    continue
                print(f"\nkernel                               {prec} GFLOP/launch   avg_us   {prec} TFLOP/s   frac of {peak / 1e12:.1f}")
                for k in kernels:
                    fl = 64.0 * (per.get(f"SQ_INSTS_VALU_ADD_{prec}", {}).get(k, 0.0)
                                 + per.get(f"SQ_INSTS_VALU_MUL_{prec}", {}).get(k, 0.0)
                                 + per.get(f"SQ_INSTS_VALU_TRANS_{prec}", {}).get(k, 0.0)
                                 + 2.0 * per.get(f"SQ_INSTS_VALU_FMA_{prec}", {}).get(k, 0.0))
                    if k in tr and fl > 0:
                        us = tr[k][1]
                        print(f"{k:36s} {fl / 1e9:16.3f} {us:8.2f} {fl / (us * 1e-6) / 1e12:13.2f} {fl / (us * 1e-6) / peak:13.3f}")


if __name__ == "__main__":
    main()
