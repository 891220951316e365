"""L-BFGS recovery of ADE material parameters from receiver IRs through the adjoint.

    python -m fases_amd.recover [--iterations 50] [--steps 3000] [--room 192,160,128]

Held-out targets: the seeded materials of the 192x160x128 differentiation room (3 wall
types with 4 real poles and 1 complex pair each, seed 1004) generate the target IRs;
the start point multiplies every residue, pole real part and R_inf by a log-uniform
factor in [1/1.2, 1.2] (seeded); the drive amplitude starts 20 % off. L-BFGS fits all of
them to the target IRs (sum of squared IR differences, FP32 storage, checkpoints every
64 steps) and prints one JSON line with the misfit history and the parameter errors.
"""
from __future__ import annotations

import argparse
import copy
import json
import sys
import time

import numpy as np

from . import workloads as W
from .autograd import RoomProblem, free_mask, lbfgs_recover, pack


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--iterations", type=int, default=50)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--room", default="192,160,128")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    args = ap.parse_args()
    room = tuple(int(v) for v in args.room.split(","))
    wl = W.config(4, room=room, nsteps=args.steps)
    sv = W.make_solver(wl, args.precision)
    target_ir, _, _, _ = sv.run(wl.nsteps)
    truth = pack(wl.materials)
    rng = np.random.default_rng(1004)
    mask = free_mask(wl.materials.K, wl.materials.L, truth.shape[0])
    start = truth * np.exp(rng.uniform(-np.log(1.2), np.log(1.2), truth.shape) * mask)
    sval0 = np.asarray(wl.src_sval, float) * 1.2
    prob = RoomProblem(sv, wl, checkpoint_every=64)
    t0 = time.time()
    log = lambda m: print("[recover] " + m, file=sys.stderr, flush=True)
    theta, sval, hist = lbfgs_recover(prob, target_ir, start, sval0, iterations=args.iterations, log=log)
    wall = time.time() - t0
    sv.close()
    used = mask > 0
    err0 = np.abs(start / truth - 1)[used]
    err1 = np.abs(theta / truth - 1)[used]
    out = {"room": room, "steps": wl.nsteps, "precision": args.precision, "iterations": args.iterations,
           "misfit_history": hist, "misfit_reduction": hist[-1] / hist[0] if hist[0] > 0 else None,
           "param_rel_err_start": {"median": float(np.median(err0)), "max": float(err0.max())},
           "param_rel_err_end": {"median": float(np.median(err1)), "max": float(err1.max())},
           "drive_rel_err_start": float(abs(sval0[0] / wl.src_sval[0] - 1)),
           "drive_rel_err_end": float(abs(sval[0] / wl.src_sval[0] - 1)), "wall_s": round(wall, 1)}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
