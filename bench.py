"""Benchmark of the MI355X DGSEM hot path: DOF-stage-updates/s at N=4 on the curvilinear viscous
wall-flow workload (128x64x2 hex cells, N=4, 2,048,000 DoF x 5 vars, Mach 6, Re 1e5, isothermal
wall). One step = dt reduction + 5 x (calc_rhs + low-storage RK update), replayed from a HIP graph
with the state resident in HBM.

`value` is the plain FP64 path. `variants` runs the same workload and step count from the same
initial state in the FP32-bulk arithmetic and in FP32-bulk with the learned troubled-cell indicator
(the network of plens_amd/indicator_mlp.npy) and reports each one's rate, its ratio to the plain
rate and its solution relative L2 difference from the plain FP64 state after the same steps.
`batched` advances the 96 seeded uneven instances (degrees 2-5, 4^3-32^3 cells) 50 steps, once
through one graph launch per step for all of them and once instance by instance.

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
# compulsory bytes per DoF-stage of the two-pass algorithm with the RK epilogue fused (viscous,
# FP64): Q read and write, R write, R' and R'' reads, the metric terms, 1/J, avars write and read
STAGE_BYTES_COMPULSORY = 8 * (5 + 5 + 5 + 5 * 4 / 5 + 5 * 3 / 5 + 9 + 1 + 2 * 9)
FP64_PEAK_TFS = 78.6   # MI355X vector (and matrix) FP64
FP32_PEAK_TFS = 157.3  # MI355X vector FP32


def kernel_bytes_per_dof(N, viscous=True, fb=8, face_maps=True):
    """Algorithmic HBM bytes per DoF of one launch of each kernel class (compulsory traffic of this
    design); fb is the byte size of the flux-arithmetic arrays (avars and the face terms: 8 for
    FP64, 4 for FP32-bulk); states, the RK residual registers and metric data are FP64 in every
    variant.
    Metric terms are derived from each cell's trilinear map: 24 doubles per cell (map_b per DoF).
    face_maps: the face kernel derives its face-node metric terms from the maps of the face's two
    cells (the HLLC instantiation on trilinear cells) instead of reading the per-face-node table."""
    n, nf = (N + 1) ** 3, (N + 1) ** 2
    face_per_dof = 6.0 * nf / n  # face nodes per cell node
    map_b = 24 * 8 / n
    v = 1 if viscous else 0
    b = {}
    # k_grad: read Q (5) and the map; write avars (9, viscous)
    b["grad"] = 8 * 5 + map_b + fb * 9 * v
    b["alpha"] = 0.0
    # k_lo: blended cells only; not counted
    b["lo"] = 0.0
    # k_face, per face node (each shared by two cells): read both states (2x5), the normal (3), the
    # face-node metric terms of both sides (10; or the two cells' maps, 2x24 per face) and both avars
    # (2x9, viscous); write both sides' weighted surface terms (2x5)
    metric = 2 * 24 * 8 / nf if face_maps else 8 * 10
    b["face"] = face_per_dof / 2 * (8 * (10 + 3) + metric + fb * (18 * v + 10))
    # k_vol: read Q (5), the map, avars (9, viscous), the weighted surface terms of its face nodes
    # (5 per face node), R' and R'' (stage averages 4/5 and 3/5 of 5); write Q (5), R (5)
    b["vol"] = 8 * (5 + 5) + map_b + 8 * (5 * 4 / 5 + 5 * 3 / 5 + 5) + fb * (9 * v + face_per_dof * 5)
    return b


def stage_bytes_per_dof(N, fb=8):
    """compulsory bytes per DoF-stage of the two-pass algorithm (lift, then residual) with the fused
    RK epilogue: Q read + write, R write, R' and R'' reads (stage averages), avars write + read, and
    the cells' trilinear maps (the metric terms are not read); the registers in FP64, avars in the
    flux-arithmetic type"""
    return 8 * (5 + 5) + 8 * (5 + 4 + 3) + fb * 18 + 24 * 8 / (N + 1) ** 3


INFINITY_CACHE_BYTES = 256 * 2 ** 20
STREAM_FROM_BYTES = 176 * 2 ** 20


def streaming_handle(n_dof, fb):
    """whether the library runs the streaming face and volume kernels for a handle of n_dof DoF with
    fb-byte flux-type values: the gathered states and auxiliary variables ((40 + 9 fb) B per DoF) fit
    in the Infinity Cache, and with the two RK registers (80 B per DoF) exceed 176 MiB of it"""
    reuse = (40 + 9 * fb) * n_dof
    return reuse <= INFINITY_CACHE_BYTES and reuse + 80 * n_dof > STREAM_FROM_BYTES


def measured_traffic(precision, kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary of this workload
    (calibrated rocprofv3 FETCH_SIZE + WRITE_SIZE passes), or None"""
    path = os.path.join(ROOT, "profiles", f"wall_flow_N4_{precision}_kernels.md")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                return float(parts[2]), os.path.relpath(path, ROOT)
    return None, None


def measured_compute(precision, kernel):
    """FP64 and FP32 VALU flops per launch of `kernel` (SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS} counter
    passes, 64 lanes per instruction, FMA = 2) and the fraction of time its SIMDs' VALUs were busy
    (SQ_ACTIVE_INST_VALU), from the committed counter summaries of this workload; None where absent"""
    key = kernel.replace(" ", "")
    flops, src = {}, None
    path = os.path.join(ROOT, "profiles", f"wall_flow_N4_{precision}_kernels.md")
    if os.path.exists(path):
        prec = None
        with open(path) as f:
            for line in f:
                if "GFLOP/launch" in line:
                    prec = line.split()[1]  # F64 / F32
                    continue
                if prec and not line.strip():
                    prec = None
                    continue
                if prec and line.startswith("k_"):
                    t = line.split()
                    if "".join(t[:-4]) == key:
                        flops[prec] = float(t[-4]) * 1e9
                        src = os.path.relpath(path, ROOT)
    busy, bsrc = None, None
    path = os.path.join(ROOT, "profiles", f"wall_flow_N4_{precision}_sq_mix.md")
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                t = [x.strip() for x in line.split("|")]
                if len(t) > 6 and t[1].replace(" ", "") == key:
                    busy, bsrc = float(t[5]), os.path.relpath(path, ROOT)
    return flops, src, busy, bsrc


# section timers of the oracle, named as the reference's TimerOutput sections (dgsem/plens_test.cc:903-976)
SECTIONS = ["Assert positivity", "Calculate aux vars", "Calculate blender", "Calculate surface fluxes",
            "ho inv residual + ho dif residual", "lo inv residual", "Calculating time step", "Updating cvars"]


def cpu_baseline(case, gpu_state_after_1):
    """Time the CPU oracle on one RK4 step of the same workload (bounded sample): on all host threads
    (the value), then on one core; section timers of the all-threads step."""
    from oracle.pyoracle import OraclePlens
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    o = OraclePlens(case, threads=threads)
    t0 = time.perf_counter()
    o.update(case.courant)
    wall = time.perf_counter() - t0
    sections = {k: float(v) for k, v in zip(SECTIONS, o.section_times())}
    rate = case.mesh.n_dof * 5 / wall
    err = float(np.linalg.norm(gpu_state_after_1 - o.get_state()) / np.linalg.norm(o.get_state()))
    o.close()
    o1 = OraclePlens(case, threads=1)
    t0 = time.perf_counter()
    o1.update(case.courant)
    wall1 = time.perf_counter() - t0
    o1.close()
    OraclePlens.set_threads(threads)
    return {"value": rate, "unit": "DOF-stage-updates/s", "cores": threads, "kind": "port",
            "sample": f"1 RK4 step (5 stages) of the full workload, {case.mesh.n_dof} DoF, wall {wall:.2f} s, "
                      f"OpenMP over cells",
            "one_core": {"value": case.mesh.n_dof * 5 / wall1, "cores": 1, "wall_s": wall1},
            "sections_s": sections}, err


def algorithmic_flops(case, state):
    """FP64 operations per DoF-stage of one RK4 step from `state`, counted by the flop-counting
    build of the CPU oracle (oracle/flopcount.h: add, sub, mul, div, sqrt, log, exp, pow = 1 each;
    the oracle has no FMA), single-threaded.

    ref: the algorithm as the reference writes it (every ordered node pair of a line and the node
    itself, one viscous flux per node pair, the dense change of basis, the subcell residual of every
    cell). min: the minimal algorithm of the roofline, from the same counts: each unordered pair of
    a line once plus the end-node diagonal terms (3 (N/2 + 2/(N+1)) pair fluxes per node instead of
    3 (N+1)), priced at the measured flux cost plus the reference's per-pair overhead (direction
    average, normalisation, weight, accumulation); one viscous flux per node and direction and a
    (N+1)-point contraction of its 4 non-zero components; the sum-factorised change of basis
    (3 (N+1) multiply-adds per node); the subcell residual only in blended cells. The BR1 lifting,
    surface fluxes, blender, time step and update are counted as the reference writes them."""
    from oracle.pyoracle import FLOP_BUCKETS, FLOPS_LIB_PATH, OraclePlens
    N, n_dof, n_cells = case.mesh.degree, case.mesh.n_dof, case.mesh.n_cells
    o = OraclePlens(case, lib_path=FLOPS_LIB_PATH)
    o.set_state(state)
    o.flop_counts(reset=True)
    b0 = o.blended_cell_stages()
    t0 = time.perf_counter()
    o.update(case.courant)
    wall = time.perf_counter() - t0
    f, k = o.flop_counts(reset=True)
    blended = (o.blended_cell_stages() - b0) / (5.0 * n_cells)
    o.close()
    ds = 5.0 * n_dof  # DoF-stages of the step
    ref = float(f.sum()) / ds
    pairs_ref, pairs_min = k[10] / ds, 3 * (N / 2 + 2 / (N + 1))
    save_pairs = (pairs_ref - pairs_min) * (f[10] + f[4]) / k[10]
    save_visc = ((k[13] / ds - 3) * f[13] / k[13] + (f[5] / ds - 3 * (N + 1) * 4 * 2)) if k[13] else 0.0
    save_cob = f[11] / ds - 3 * (N + 1) * 2
    save_lo = f[6] / ds * (1.0 - blended)
    fmin = ref - save_pairs - save_visc - save_cob - save_lo
    return {"ref_flop_per_dof_stage": ref, "min_flop_per_dof_stage": fmin,
            "sections_ref": {name: float(f[i]) / ds for i, name in enumerate(FLOP_BUCKETS)},
            "minimal_savings": {"unordered_pairs": save_pairs, "viscous_volume_contraction": save_visc,
                                "sum_factorised_change_of_basis": save_cob, "subcell_residual_blended_only": save_lo},
            "blended_cell_fraction": blended, "pair_fluxes_per_node_ref": float(pairs_ref),
            "pair_fluxes_per_node_min": pairs_min, "count_wall_s": wall,
            "sample": "1 RK4 step (5 stages) of the full workload from the timed run's final state, single thread"}


def fp64_sensitivity(case, args, plain_state):
    """relative L2 change of the plain FP64 state after warmup + steps when the initial state is
    perturbed by a relative 2^-24 (FP32 round-off) and 1e-15 (FP64 round-off) uniform noise: the
    spread that the workload itself puts on any change of arithmetic"""
    from plens_amd import GpuPlens
    rng = np.random.default_rng(3)
    out = {}
    for label, eps in (("input_perturbation_2^-24", 2.0 ** -24), ("input_perturbation_1e-15", 1e-15)):
        g = GpuPlens(case)
        g.set_state(case.Q0 * (1.0 + eps * rng.uniform(-1.0, 1.0, case.Q0.shape)))
        g.run(args.warmup + args.steps, case.courant)
        out[label] = float(np.linalg.norm(g.get_state() - plain_state) / np.linalg.norm(plain_state))
        g.close()
    return out


def run_variants(case, args, plain_value, plain_state):
    """the same workload, steps and initial state in the FP32-bulk arithmetic, without and with the
    learned indicator; each variant's state difference from the plain FP64 state is set beside the
    FP64 path's own spread under an FP32-sized perturbation of its input"""
    from plens_amd import GpuPlens, with_precision
    from plens_amd.indicator_study import load_trained
    from plens_amd.learned import stats, use_network
    net = load_trained()
    sens = fp64_sensitivity(case, args, plain_state)
    print(f"[bench] FP64 spread under input perturbations: {sens}", file=sys.stderr, flush=True)
    out = {"fp64_sensitivity": sens}
    for label, prec, learned in (("fp32_bulk", "fp32_bulk", False), ("fp32_bulk+learned", "fp32_bulk", True),
                                 ("fp64+learned", "fp64", True)):
        g = GpuPlens(with_precision(case, prec))
        if learned:
            use_network(g, net)
        g.run(args.warmup, case.courant)
        if learned:
            stats(g, reset=True)
        g.run(args.steps, case.courant)
        rate = case.mesh.n_dof * 5 * args.steps / (g.last_run_ms() * 1e-3)
        err = float(np.linalg.norm(g.get_state() - plain_state) / np.linalg.norm(plain_state))
        rec = {"value": rate, "vs_plain_fp64": rate / plain_value, "ms_per_step": g.last_run_ms() / args.steps,
               "rel_l2_vs_plain_fp64": err,
               "rel_l2_over_fp64_spread": err / max(sens["input_perturbation_2^-24"], 1e-300)}
        if learned:
            ev, fb = stats(g)
            rec["indicator_fallback_fraction"] = fb / max(ev, 1)
        out[label] = rec
        print(f"[bench] {label}: {rate:.4e} DOF-stage/s ({rate / plain_value:.3f}x), rel L2 vs FP64 {err:.2e}",
              file=sys.stderr, flush=True)
        g.close()
    return out


def fp32_bulk_errors(case):
    """FP32-bulk error where the plain path is not chaotic under rounding: the measured wall flow over
    its first 5 steps (before the blender switches wall cells) against the plain FP64 GPU state, and
    the periodic cube at full size (8^3, N=3) after its 100 steps against the FP64 CPU oracle"""
    from oracle.pyoracle import OraclePlens
    from plens_amd import GpuPlens, periodic_cube, with_precision
    out = {}
    a, b = GpuPlens(case), GpuPlens(with_precision(case, "fp32_bulk"))
    a.run(5, case.courant)
    b.run(5, case.courant)
    qa = a.get_state()
    out["wall_flow_5_steps_rel_l2_vs_fp64"] = float(np.linalg.norm(b.get_state() - qa) / np.linalg.norm(qa))
    a.close()
    b.close()
    cube = periodic_cube()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    o = OraclePlens(cube, threads=threads)
    o.run(100, cube.courant)
    qo = o.get_state()
    o.close()
    for label, prec in (("fp64", "fp64"), ("fp32_bulk", "fp32_bulk")):
        g = GpuPlens(with_precision(cube, prec))
        g.run(100, cube.courant)
        out[f"cube_8^3_100_steps_{label}_rel_l2_vs_oracle"] = float(np.linalg.norm(g.get_state() - qo) / np.linalg.norm(qo))
        g.close()
    print(f"[bench] FP32-bulk errors where rounding is not amplified: {out}", file=sys.stderr, flush=True)
    return out


def run_batched(n_instances=96, steps=50):
    """96 seeded uneven instances, 50 steps: one graph launch per step for all vs one by one"""
    from plens_amd import GpuBatch, GpuPlens, batch_instance
    t0 = time.perf_counter()
    cases = [batch_instance(i) for i in range(n_instances)]
    ndof = sum(c.mesh.n_dof for c in cases)
    cfl = [c.courant for c in cases]
    print(f"[bench] batched: {n_instances} instances, {ndof} DoF, built in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    seq_ms = 0.0
    seq = [GpuPlens(c) for c in cases]
    for g, c in zip(seq, cases):
        g.run(steps, c.courant)
        seq_ms += g.last_run_ms()
    tuning = os.path.join(ROOT, "plens_amd", "tuning.txt")
    b = GpuBatch([GpuPlens(c) for c in cases], tuning_db=tuning)
    b.run(steps, cfl)
    bms = b.last_run_ms()
    same = all(np.array_equal(b.get_state(i), seq[i].get_state()) for i in range(n_instances))
    out = {"instances": n_instances, "steps": steps, "n_dof": ndof,
           "value": ndof * 5 * steps / (bms * 1e-3), "sequential_value": ndof * 5 * steps / (seq_ms * 1e-3),
           "speedup_vs_sequential": seq_ms / bms, "bitwise_equal_to_sequential": same, "groups": b.groups()}
    print(f"[bench] batched {out['value']:.4e} vs sequential {out['sequential_value']:.4e} DOF-stage/s",
          file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cells-per-block", type=int, default=0)
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32_bulk"],
                    help="arithmetic of the main leg (fp64 is the metric's plain path; fp32_bulk for profiling)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the multi-GPU sharding of this path is a design note, "
              "not a built path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from plens_amd import GpuPlens, wall_flow, with_precision
    case = with_precision(wall_flow(), args.precision)
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
    final_plain = g.get_state()  # state after warmup + steps, the reference of the variants

    # per-kernel durations: every class bracketed by HIP events in a stepwise run (includes the
    # event and launch gaps), and the gradient, face and volume kernels launched back to back with
    # the parameters of the 5 stages in turn (one launch's device time, as a kernel trace reports it)
    _, kt = g.run_profiled(args.profile_steps, case.courant)
    names = ["grad", "alpha", "face", "lo", "vol"]
    b2b = {k: g.time_kernel(k, reps=10) for k in ("grad", "face", "vol")}
    kb = kernel_bytes_per_dof(N, viscous=True, fb=8 if args.precision == "fp64" else 4)
    kernels = {}
    for i, k in enumerate(names):
        ms = b2b.get(k, kt[i])
        bytes_launch = kb[k] * n_dof
        kernels[k] = {"avg_ms": ms, "avg_ms_event_bracketed": kt[i], "alg_bytes": bytes_launch,
                      "achieved_GBs": bytes_launch / (ms * 1e-3) / 1e9 if ms > 0 else None}
    dom = max(names, key=lambda k: kernels[k]["avg_ms"])
    ach = kernels[dom]["achieved_GBs"]
    stage_ms = float(sum(kernels[k]["avg_ms"] for k in names))
    fb = 8 if args.precision == "fp64" else 4
    stage_bytes = stage_bytes_per_dof(N, fb)
    # kernel name as rocprofv3 prints it: k_grad<N, type, one cell per workgroup>, k_vol<N, type,
    # extended scheme, one cell per workgroup[, streaming]>, k_face<N, type, extended scheme, HLLC in
    # the normal frame[, streaming]>, k_lo<N, type, extended scheme, HLLC> (here: HLLC surface flux,
    # Chandrashekhar volume flux, one cell per workgroup); the streaming instantiations run when the
    # gathered states and auxiliary variables fit in the Infinity Cache only without the RK registers
    one = "true" if (N >= 3 and args.cells_per_block in (0, 1)) else "false"
    st = ",true" if streaming_handle(n_dof, fb) else ""
    tail = {"grad": f",{one}>", "vol": f",false,{one}{st}>", "face": f",false,true{st}>", "lo": ",false,true>"}
    tname = ("k_alpha" if dom == "alpha" else
             f"k_{dom}<{N},{'double' if args.precision == 'fp64' else 'float'}" + tail[dom])
    traffic, tsrc = measured_traffic(args.precision, tname)
    roofline = {"bound": "hbm", "kernel": tname, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": tsrc,
                "alg_bytes_per_launch": kernels[dom]["alg_bytes"],
                "stage": {"alg_bytes_per_dof": stage_bytes, "stage_ms": stage_ms,
                          "achieved_GBs": stage_bytes * n_dof / (stage_ms * 1e-3) / 1e9,
                          "frac": stage_bytes * n_dof / (stage_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
                "kernels": kernels}

    # where the dominant kernel stands against the vector FP64 (and FP32) peaks: flops per launch from
    # the committed VALU counter passes over its live duration
    flops, fsrc, busy, bsrc = measured_compute(args.precision, tname)
    dom_ms = kernels[dom]["avg_ms"]
    roofline["compute"] = {
        "unit": "TFLOP/s", "flop_source": fsrc, "valu_busy": busy, "valu_busy_source": bsrc,
        "f64": ({"flop_per_launch": flops["F64"], "achieved": flops["F64"] / (dom_ms * 1e-3) / 1e12,
                 "peak": FP64_PEAK_TFS, "frac": flops["F64"] / (dom_ms * 1e-3) / 1e12 / FP64_PEAK_TFS}
                if "F64" in flops else None),
        "f32": ({"flop_per_launch": flops["F32"], "achieved": flops["F32"] / (dom_ms * 1e-3) / 1e12,
                 "peak": FP32_PEAK_TFS, "frac": flops["F32"] / (dom_ms * 1e-3) / 1e12 / FP32_PEAK_TFS}
                if "F32" in flops else None)}

    # achievable HBM bandwidth on this device: streaming copy of 1 GiB (2 GiB of traffic per copy)
    import ctypes as C
    from plens_amd import _lib as L
    cms = C.c_double()
    copy_bytes = 1 << 30
    L.check(L.load().plens_gpu_stream_copy(copy_bytes, 10, C.byref(cms)), "stream_copy")
    roofline["stream_copy_GBs"] = 2 * copy_bytes / (cms.value * 1e-3) / 1e9

    g.close()
    variants = None if args.no_variants else run_variants(case, args, value, final_plain)
    if variants is not None:
        variants["fp32_bulk_errors"] = fp32_bulk_errors(case)
    batched = None if args.no_batch else run_batched()
    cpu, parity = (None, None)
    if not args.no_cpu:
        cpu, parity = cpu_baseline(case, state1)
        # the roofline of the whole path (bound per DoF-stage): HBM time of the compulsory
        # bytes against FP64 time of the minimal algorithm's operations, at the timed rate
        af = algorithmic_flops(case, final_plain)
        t_hbm = STAGE_BYTES_COMPULSORY / (HBM_PEAK_GBS * 1e9)
        t_f64 = af["min_flop_per_dof_stage"] / (FP64_PEAK_TFS * 1e12)
        t_meas = 1.0 / value
        roofline["path"] = {"alg_flops": af, "bytes_per_dof_stage": STAGE_BYTES_COMPULSORY,
                            "hbm_s_per_dof_stage": t_hbm, "f64_s_per_dof_stage": t_f64,
                            "bound": "hbm" if t_hbm >= t_f64 else "f64",
                            "bound_rate": 1.0 / max(t_hbm, t_f64), "frac": max(t_hbm, t_f64) / t_meas}
        print(f"[bench] algorithmic flops per DoF-stage: reference {af['ref_flop_per_dof_stage']:.0f}, minimal "
              f"{af['min_flop_per_dof_stage']:.0f}; path at {roofline['path']['frac']:.3f} of its "
              f"{roofline['path']['bound']} bound", file=sys.stderr, flush=True)
    out = {
        "metric": METRIC, "value": value, "unit": "DOF-stage-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32+f64", "data": "synthetic (seeded mesh perturbation, uniform Mach-6 inflow IC)",
        "config": {"workload": "wall_flow 128x64x2 hex, N=4, 2,048,000 DoF x 5, viscous, HLLC, isothermal wall",
                   "n_dof": n_dof, "degree": N,
                   "variant": "plain FP64" if args.precision == "fp64" else "FP32-bulk"},
        "roofline": roofline, "cpu_baseline": cpu, "parity_1step_rel_l2_vs_oracle": parity,
        "variants": variants, "batched": batched,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
