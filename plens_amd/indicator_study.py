"""Learned troubled-cell indicator study: training data from the plain GPU path, training on the GPU
with the hand-written kernels of plens_amd/csrc/mlp.hip, and evaluation on a held-out shock tube.

Training data: plain runs of the periodic density-wave cube (8^3 cells, 100 steps) and of the
Mach-5 shock tube (128x8x8 cells, 200 steps), at degrees 3 and 4, sampled every 10 steps
(features: Legendre modes of rho*p over the cell mean, zero-padded to 125; label: the plain blender
value of BlenderCalculator::get_blender, dgsem/blender_calculator.cc:123-187). Seeds: cube 0-9,
tube 1-6; the tube seed 1001 is held out for evaluation.

Evaluation (held-out tube, seed 1001, 200 steps): the indicator's agreement with the plain one on
the held-out samples (trained and untrained network), the fallback rate of the fused inference
(plain value kept where |alpha_mlp - alpha_plain| > tolerance), and the density L2 and
shock-position differences of the learned run against the plain run.

    python -m plens_amd.indicator_study --weights plens_amd/indicator_mlp.npy --out study.json

The committed plens_amd/indicator_mlp.npy holds the FP32 parameters this produced (packed W1, b1,
W2, b2, w3, b3 as in plens_amd.learned.unpack); load_trained() returns them as a network.
"""
import argparse
import json
import os
import time

import numpy as np

from .cases import periodic_cube, shock_tube
from .learned import MLP, collect, init_params, stats, use_network
from .solver import GpuPlens

HELD_OUT_SEED = 1001
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "indicator_mlp.npy")


def load_trained(path=WEIGHTS):
    """the network trained by this study (FP32 parameters from the .npy file)"""
    return MLP(np.load(path))


def training_set(cube_seeds=range(10), tube_seeds=range(1, 7), degrees=(3, 4), every=10, log=print):
    Xs, ys = [], []
    for deg in degrees:
        for s in cube_seeds:
            X, y = collect(periodic_cube(degree=deg, n=8, seed=s, flux="HLLC"), steps=100, every=every)
            Xs.append(X)
            ys.append(y)
        for s in tube_seeds:
            X, y = collect(shock_tube(degree=deg, seed=s), steps=200, every=every)
            Xs.append(X)
            ys.append(y)
            log(f"  tube N={deg} seed {s}: {len(y)} samples, {np.mean(y > 0):.4f} blended")
    return np.concatenate(Xs), np.concatenate(ys).astype(np.float32)


def agreement(mlp, X, y, floor=0.02, alpha_max=0.5, tol=0.05):
    """agreement of the network's blender value (zero below the network floor, capped at alpha_max,
    as in the solver's inference kernel) with the plain one"""
    p = mlp.predict(X).astype(np.float64)
    p = np.where(p < floor, 0.0, np.minimum(p, alpha_max))
    err = np.abs(p - y)
    troubled = y > 0
    return {"mae": float(err.mean()), "fallback_fraction": float(np.mean(err > tol)),
            "troubled_mae": float(err[troubled].mean()) if troubled.any() else 0.0,
            "troubled_fallback_fraction": float(np.mean(err[troubled] > tol)) if troubled.any() else 0.0,
            "n": int(len(y)), "n_troubled": int(troubled.sum())}


def column_density(case, Q):
    """cell-mean density (GLL quadrature with detJ) averaged over the y-z cells of every x column"""
    mesh = case.mesh
    from . import _lib as L
    import ctypes as C
    N = mesh.degree
    NP = N + 1
    x = np.empty(NP)
    w = np.empty(NP)
    D = np.empty(NP * NP)
    C1 = np.empty(NP * NP)
    dp = C.POINTER(C.c_double)
    L.check(L.load().plens_basis_1d(N, x.ctypes.data_as(dp), w.ctypes.data_as(dp), D.ctypes.data_as(dp),
                                    C1.ctypes.data_as(dp)), "basis")
    w3 = np.einsum("i,j,k->kji", w, w, w).reshape(-1)  # node id i + NP j + NP^2 k
    rho = Q[0].reshape(mesh.n_cells, -1)
    vol = (mesh.detJ * w3).sum(1)
    cm = (rho * mesh.detJ * w3).sum(1) / vol
    nx, ny, nz = mesh.shape
    return cm.reshape(nz, ny, nx).mean(axis=(0, 1))


def shock_position(case, Q, rho1=1.0):
    """x where the column density first rises to the mean of the pre- and post-shock densities
    (linear interpolation between column centres)"""
    prof = column_density(case, Q)
    nx = case.mesh.shape[0]
    lo, hi = case.mesh.params.lo[0], case.mesh.params.hi[0]
    xc = lo + (np.arange(nx) + 0.5) * (hi - lo) / nx
    mid = 0.5 * (rho1 + prof[-1])
    k = int(np.argmax(prof >= mid))
    if k == 0:
        return float(xc[0])
    return float(xc[k - 1] + (mid - prof[k - 1]) / (prof[k] - prof[k - 1]) * (xc[k] - xc[k - 1]))


def evaluate(mlp, degree=3, steps=200, tol=0.05, seed=HELD_OUT_SEED):
    case = shock_tube(degree=degree, seed=seed)
    plain = GpuPlens(case)
    tp = plain.run(steps, case.courant)
    learned = GpuPlens(case)
    use_network(learned, mlp, tol)
    stats(learned, reset=True)
    tl = learned.run(steps, case.courant)
    ev, fb = stats(learned)
    Qp, Ql = plain.get_state(), learned.get_state()
    drho = float(np.linalg.norm(Ql[0] - Qp[0]) / np.linalg.norm(Qp[0]))
    xp, xl = shock_position(case, Qp), shock_position(case, Ql)
    dx = (case.mesh.params.hi[0] - case.mesh.params.lo[0]) / case.mesh.shape[0]
    return {"degree": degree, "steps": steps, "tolerance": tol, "t_plain": tp, "t_learned": tl,
            "density_rel_l2_vs_plain": drho, "shock_x_plain": xp, "shock_x_learned": xl,
            "shock_position_error": abs(xl - xp), "shock_position_error_cells": abs(xl - xp) / dx,
            "cells_evaluated": ev, "fallbacks": fb, "fallback_fraction": fb / max(ev, 1)}


def stage_cost(mlp, precision="fp64", steps=5):
    """per-stage kernel times of the wall-flow workload (N=4) with the plain and the learned indicator"""
    from .cases import wall_flow, with_precision
    case = with_precision(wall_flow(), precision)
    out = {}
    for label, net in (("plain", None), ("learned", mlp)):
        g = GpuPlens(case)
        if net is not None:
            use_network(g, net, 0.05)
        g.run(2, case.courant)
        g.run(steps, case.courant)
        ms = g.last_run_ms() / steps
        _, kt = g.run_profiled(steps, case.courant)
        out[label] = {"ms_per_step": ms, "dof_stage_per_s": 5 * case.mesh.n_dof / ms * 1e3,
                      "kernel_ms_per_stage": [float(v) for v in kt]}
        g.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--weights", default=None, help="write the trained FP32 parameters here (.npy)")
    ap.add_argument("--out", default=None)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--tube-seeds", type=int, default=6)
    a = ap.parse_args()
    log = lambda *x: print(*x, flush=True)  # noqa: E731
    t0 = time.time()
    X, y = training_set(tube_seeds=range(1, 1 + a.tube_seeds), log=log)
    log(f"training set: {len(y)} samples, {np.mean(y > 0):.4f} blended, {time.time() - t0:.1f} s")
    Xh, yh = collect(shock_tube(degree=3, seed=HELD_OUT_SEED), steps=200, every=10)
    untrained = MLP(init_params(0))
    res = {"n_train": int(len(y)), "held_out_seed": HELD_OUT_SEED,
           "untrained": agreement(untrained, Xh, yh)}
    mlp = MLP(init_params(0))
    t1 = time.time()
    hist = mlp.train(X, y, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=1)
    res["train_seconds"] = time.time() - t1
    res["loss_per_epoch"] = [float(v) for v in hist]
    res["trained"] = agreement(mlp, Xh, yh)
    log(f"trained {a.epochs} epochs in {res['train_seconds']:.1f} s; loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    log(f"held-out agreement untrained {res['untrained']}")
    log(f"held-out agreement trained   {res['trained']}")
    res["heldout_run"] = evaluate(mlp)
    log(f"held-out run: {res['heldout_run']}")
    res["heldout_run_untrained"] = evaluate(untrained)
    log(f"held-out run (untrained network): {res['heldout_run_untrained']}")
    res["cost_fp64"] = stage_cost(mlp, "fp64")
    res["cost_fp32_bulk"] = stage_cost(mlp, "fp32_bulk")
    log(f"cost fp64 {res['cost_fp64']}")
    log(f"cost fp32 {res['cost_fp32_bulk']}")
    if a.weights:
        np.save(a.weights, mlp.params())
    if a.out:
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
