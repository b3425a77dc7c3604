"""Benchmark of the MI355X DGSEM hot path: DOF-stage-updates/s at N=4 on the curvilinear viscous
wall-flow workload (128x64x2 hex cells, N=4, 2,048,000 DoF x 5 vars, Mach 6, Re 1e5, isothermal
wall), FP64. One step = dt reduction + 5 x (calc_rhs + low-storage RK update), replayed from a HIP
graph with the state resident in HBM.

Prints ONE JSON line. `roofline` is for the dominant kernel (volume + RK epilogue) from HIP-event
durations on the handle's stream; `cpu_baseline` times the CPU oracle on a bounded sample of the
same workload on this host and also reports the one-step parity of the GPU state against it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DOF-stage-updates/s at N=4: plain FP64 vs mixed+learned, at equal L2 error"
HBM_PEAK_GBS = 8000.0


def kernel_bytes_per_dof(N, viscous=True):
    """Algorithmic HBM bytes per DoF of one launch of each kernel class (compulsory traffic)."""
    n, nf = (N + 1) ** 3, (N + 1) ** 2
    face_per_dof = 6.0 * nf / n  # face nodes per cell node
    b = {}
    # k_grad: read Q (5), Ja (9), detJ (1); write avars (9, viscous)
    b["grad"] = 8 * (5 + 9 + 1 + (9 if viscous else 0))
    b["alpha"] = 0.0
    # k_lo: blended cells only (none in the default state); early-exit blocks read alpha
    b["lo"] = 0.0
    # k_face: per face node read both states (2x5) and avars (2x9), normal (3); write both sides' fluxes
    # (2 x 2 x 5); amortised per DoF (each face node shared by two cells)
    b["face"] = 8 * face_per_dof / 2 * (10 + (18 if viscous else 0) + 3 + 10 * (2 if viscous else 1))
    # k_vol: read Q (5), Ja (9), detJ (1), avars (9), stored face fluxes (2 x 5 per face node),
    # R' and R'' (stage average 4/5 and 3/5 of 5); write Q (5), R (5); the subcell residual of
    # blended cells is not counted (none at the measured state)
    b["vol"] = 8 * (5 + 9 + 1 + (9 if viscous else 0) + face_per_dof * 5 * (2 if viscous else 1)
                    + 5 * 4 / 5 + 5 * 3 / 5 + 10)
    return b


def cpu_baseline(case, gpu_state_after_1, seconds_target=20.0):
    """Time the CPU oracle on one RK4 step of the same workload (bounded sample)."""
    from oracle.pyoracle import OraclePlens
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    o = OraclePlens(case, threads=threads)
    t0 = time.perf_counter()
    o.update(case.courant)
    wall = time.perf_counter() - t0
    rate = case.mesh.n_dof * 5 / wall
    err = float(np.linalg.norm(gpu_state_after_1 - o.get_state()) / np.linalg.norm(o.get_state()))
    o.close()
    return {"value": rate, "unit": "DOF-stage-updates/s", "cores": threads, "kind": "port",
            "sample": f"1 RK4 step (5 stages) of the full workload, {case.mesh.n_dof} DoF, wall {wall:.2f} s, "
                      f"OpenMP over cells"}, err


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cells-per-block", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the multi-GPU sharding of this path is a design note, "
              "not a built path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from plens_amd import GpuPlens, wall_flow
    case = wall_flow()
    N, n_dof = case.mesh.degree, case.mesh.n_dof
    print(f"[bench] workload wall_flow {case.meta['cells']} N={N} n_dof={n_dof}", file=sys.stderr, flush=True)

    g = GpuPlens(case, cells_per_block=args.cells_per_block)
    # one step from the initial state for the full-size parity check against the oracle
    g.update(case.courant)
    state1 = g.get_state()
    g.set_state(case.Q0)

    g.run(args.warmup, case.courant)
    t0 = time.perf_counter()
    g.run(args.steps, case.courant)
    wall = time.perf_counter() - t0
    dev_ms = g.last_run_ms()
    ms_per_step = dev_ms / args.steps
    value = n_dof * 5 * args.steps / (dev_ms * 1e-3)
    print(f"[bench] {args.steps} steps: device {dev_ms:.2f} ms, wall {wall * 1e3:.2f} ms", file=sys.stderr, flush=True)

    _, kt = g.run_profiled(args.profile_steps, case.courant)
    names = ["grad", "alpha", "face", "lo", "vol"]
    kb = kernel_bytes_per_dof(N, viscous=True)
    kernels = {}
    for i, k in enumerate(names):
        ms = kt[i]
        bytes_launch = kb[k] * n_dof
        kernels[k] = {"avg_ms": ms, "alg_bytes": bytes_launch,
                      "achieved_GBs": bytes_launch / (ms * 1e-3) / 1e9 if ms > 0 else None}
    dom = max(names, key=lambda k: kernels[k]["avg_ms"])
    ach = kernels[dom]["achieved_GBs"]
    stage_ms = float(sum(kt))
    roofline = {"bound": "hbm", "kernel": f"k_{dom}", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS, "traffic": None,
                "stage": {"alg_bytes_per_dof": 400.0, "stage_ms": stage_ms,
                          "achieved_GBs": 400.0 * n_dof / (stage_ms * 1e-3) / 1e9,
                          "frac": 400.0 * n_dof / (stage_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
                "kernels": kernels}

    # achievable HBM bandwidth on this device: streaming copy of 1 GiB (2 GiB of traffic per copy)
    import ctypes as C
    from plens_amd import _lib as L
    cms = C.c_double()
    copy_bytes = 1 << 30
    L.check(L.load().plens_gpu_stream_copy(copy_bytes, 10, C.byref(cms)), "stream_copy")
    roofline["stream_copy_GBs"] = 2 * copy_bytes / (cms.value * 1e-3) / 1e9

    cpu, parity = (None, None)
    if not args.no_cpu:
        cpu, parity = cpu_baseline(case, state1)
    g.close()
    out = {
        "metric": METRIC, "value": value, "unit": "DOF-stage-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded mesh perturbation, uniform Mach-6 inflow IC)",
        "config": {"workload": "wall_flow 128x64x2 hex, N=4, 2,048,000 DoF x 5, viscous, HLLC, isothermal wall",
                   "n_dof": n_dof, "degree": N, "variant": "plain FP64"},
        "roofline": roofline, "cpu_baseline": cpu, "parity_1step_rel_l2_vs_oracle": parity,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
