"""Benchmark of the FVM time-stepping core on one MI355X.

Workload (default): the 512x384x256-cell shoebox (dx = 0.01 m, 16-cell cPML on +x,
five ADE walls drawn from three seeded material types, K = 4 real poles + L = 2
complex pairs, Ricker 2 kHz source, 8 receivers), FP32 field storage. A step is one
full SSP-RK3 step (3 flux stages + sources + immersed boundary) over the room.

Prints ONE JSON line: Mcell-updates/s (room cells x steps / HIP-event time of the
step loop), the flux-stage kernel's achieved algorithmic HBM bandwidth vs the 8 TB/s
peak (and vs a streaming copy measured in the same run), its VALU issue rate, the IR
relative L2 vs the FP64 path over a fixed window that holds the direct sound at every
receiver (independent of --steps), and the FP64 CPU oracle timed on bounded samples on
this host. Beside the headline:
the same room with FP16 storage (per-field power-of-two scales, FP32 math) and its IR
error vs the FP64 and FP32 runs, the batched benchmark (32 seeded rooms of 0.9-7.9 M
cells, 1500-4000 steps each, FP32) run as one batch vs the same jobs run one after the
other, the adjoint gradient vs the plain forward, the scene classification, and the
geometrical acoustics (image sources + ray tracer) of a tessellated room next to the
reference's own ray tracer on the host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# non-packed FP32 VALU issue: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz (lane-ops/s)
VALU_PEAK_LANE_OPS = 256 * 4 * 16 * 2.4e9
# an IR whose FP64 reference stays below this [Pa] has no signal to compare yet
IR_FLOOR_PA = 1e-6
METRIC = "Mcell-updates/s + % HBM roofline; batched-source & adjoint vs plain, same GPU"
ESZ = {"fp64": 8, "fp32": 4, "fp16": 2}
DTYPE = {"fp64": "f64", "fp32": "f32", "fp16": "f16-storage/f32-math"}


def log(msg: str) -> None:
    print("[bench] " + msg, file=sys.stderr, flush=True)


def geometry_timing(room, h) -> dict:
    """Scene classification (signed distance + material id) on the host (OpenMP) and on
    the GPU, with the results compared bit for bit: the bench room, and a room of the
    same cell count per side ratio meshed with 12 n^2 triangles."""
    import tempfile

    from fases_amd.solver import Scene
    from fases_amd.workloads import write_tessellated_room

    def both(make):
        t = time.time()
        a = make(None)
        th = time.time() - t
        t = time.time()
        b = make(0)
        tg = time.time() - t
        same = bool(np.array_equal(a.S.view(np.uint32), b.S.view(np.uint32)) and np.array_equal(a.oid, b.oid))
        return {"host_s": round(th, 3), "gpu_s": round(tg, 3), "speedup": round(th / tg, 1), "bitwise_equal": same,
                "cells": int(np.prod(a.dims))}

    Scene.shoebox(8, 8, 8, h, device=0)         # device initialised outside the timing
    out = {"room": "shoebox %s cells dx=%g (12 triangles)" % ("x".join(map(str, room)), h),
           **both(lambda d: Scene.shoebox(*room, h, device=d))}
    tess = (256, 192, 128)
    with tempfile.TemporaryDirectory() as tmp:
        p = os.path.join(tmp, "tess.obj")
        nt = write_tessellated_room(p, tess[0] * 0.02, tess[1] * 0.02, tess[2] * 0.02, 11)
        out["tessellated"] = {"room": "%s cells dx=0.02, walls split into 11x11 quads (%d triangles)"
                                      % ("x".join(map(str, tess)), nt),
                              **both(lambda d: Scene.from_obj(p, 0.02, device=d))}
    out["host_threads"] = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    log("geometry: bench room host %.2f s / GPU %.3f s (equal %s); tessellated host %.2f s / GPU %.3f s (equal %s)"
        % (out["host_s"], out["gpu_s"], out["bitwise_equal"], out["tessellated"]["host_s"],
           out["tessellated"]["gpu_s"], out["tessellated"]["bitwise_equal"]))
    return out


def accuracy_window(wl) -> tuple:
    """Steps until the direct sound has passed every receiver: the longest source-receiver
    distance at 0.5 cells per step (CFL 0.5) plus three times the wavelet's peak delay,
    rounded up to 100 steps, at least 600. Returns (steps, last arrival step)."""
    s = np.asarray(wl.src_ijk[0], float)
    r = np.asarray(wl.rec_ijk, float)
    arrive = float(np.sqrt(((r - s) ** 2).sum(axis=1)).max()) / (wl.params.c * wl.params.dt / wl.h)
    peak = float(np.argmax(np.abs(wl.u_s))) / 2.0       # u_s is sampled at dt/2
    return max(600, int(np.ceil((arrive + 3.0 * peak) / 100.0)) * 100 + 100), int(round(arrive))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def profile_json(name: str):
    p = ROOT / "profiles" / name
    try:
        return json.loads(p.read_text()) if p.exists() else None
    except Exception:
        return None


class stdout_to_stderr:
    """Sends what C code prints to stdout (file descriptor 1) to stderr meanwhile, so the
    bench's stdout stays one JSON line."""

    def __enter__(self):
        import ctypes
        self._libc = ctypes.CDLL(None)
        sys.stdout.flush()
        self._libc.fflush(None)
        self._saved = os.dup(1)
        os.dup2(2, 1)

    def __exit__(self, *exc):
        self._libc.fflush(None)
        os.dup2(self._saved, 1)
        os.close(self._saved)
        return False


def ga_timing(n_rays: int) -> dict:
    """Geometrical acoustics of a 12 x 9 x 4 m room whose walls are split into 16 x 16
    quads (3072 triangles, six wall groups): image sources to order 3 (GPU path validity)
    and the ray tracer (n_rays per source, 2 sources, 4 receivers, 1 s at 44.1 kHz),
    beside the reference's own rt_trace_late built from its sources (one core) on the
    same rays, with the echogram energies compared per band and receiver."""
    import tempfile

    from fases_amd import ga as GA
    from fases_amd.solver import Scene
    from fases_amd.workloads import write_tessellated_room

    rng = np.random.default_rng(2024)
    src = np.array([(2.1, 3.0, 1.5), (9.4, 6.2, 1.2)], np.float32)
    rec = np.array([(6.0, 4.5, 1.6), (10.5, 1.5, 1.1), (3.2, 7.7, 2.4), (7.9, 2.4, 3.1)], np.float32)
    alpha = rng.uniform(0.05, 0.4, (6, 8)).astype(np.float32)
    sctr = rng.uniform(0.05, 0.3, 6).astype(np.float32)
    m_air = np.array([1e-5, 4e-5, 1.2e-4, 3e-4, 6e-4, 1.1e-3, 2.8e-3, 8.5e-3], np.float32)
    src_p = np.ones(2, np.float32)
    c, fs, Nt = 343.0, 44100.0, 44100
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "tess.obj")
        ntri = write_tessellated_room(path, 12.0, 9.0, 4.0, 16)
        sc = Scene.from_obj(path, np.float32(0.05))
        prm = GA.GAParams(device=0, ism_order=3, n_rays=n_rays, c=c, fs=fs, Nt=Nt)
        GA.run(sc, src[:1], rec[:1], GA.GAParams(device=0, ism_order=1, n_rays=64, c=c, fs=fs, Nt=64),
               src_pressure=src_p[:1], alpha=alpha, sctr=sctr, m_air=m_air).close()      # warm-up
        t = time.time()
        g = GA.run(sc, src, rec, prm, src_pressure=src_p, alpha=alpha, sctr=sctr, m_air=m_air)
        wall = time.time() - t
        info = g.info()
        es, ed = g.energy()
        g.close()
        out = {"workload": "12x9x4 m room, walls in 16x16 quads (%d triangles), ISM order 3, %d rays x 2 sources, "
                           "4 receivers, 1 s at 44.1 kHz" % (ntri, n_rays),
               "ism_ms": round(info["ism_ms"], 2), "image_paths": info["paths"], "rt_ms": round(info["rt_ms"], 2),
               "rt_kernel_ms": round(info["rt_kernel_ms"], 2), "wall_s": round(wall, 3),
               "rays_per_s": round(2 * n_rays / (info["rt_kernel_ms"] * 1e-3), 1)}
        from oracle import oracle as O
        if O.ref_available():
            t = time.time()
            with stdout_to_stderr():
                res, red = O.ref_rt(path, np.float32(0.05), src, src_p, rec, alpha, m_air, sctr, c, fs, Nt, n_rays,
                                    100000, 0.12, info["order_used"])
            el = time.time() - t
            mine = (es + ed).sum(axis=2)
            ref = (res + red).sum(axis=2)
            out["cpu_reference"] = {"kind": "reference", "cores": 1, "rt_s": round(el, 3),
                                    "sample": "the reference's rt_build + rt_trace_late (raytrace3D.h) built from "
                                              "its sources without OpenMP, same scene and rays"}
            out["rt_speedup_vs_cpu_reference"] = round(el / (info["rt_kernel_ms"] * 1e-3), 1)
            out["energy_max_rel_diff_vs_reference"] = float(np.abs(mine - ref).max() / ref.max())
    log("geometrical acoustics: ISM %.1f ms, ray tracer %.1f ms (%s rays/s)%s"
        % (out["ism_ms"], out["rt_kernel_ms"], out["rays_per_s"],
           ", reference CPU %.2f s" % out["cpu_reference"]["rt_s"] if "cpu_reference" in out else ""))
    return out


def chunked_run(sv, steps, what, chunk=1000):
    """IR blocks of `steps` steps run in chunks, with a progress line per chunk (the
    FP64 path over thousands of steps of the largest room takes minutes)."""
    out, done = [], 0
    while done < steps:
        n = min(chunk, steps - done)
        out.append(sv.run(n)[0])
        done += n
        log("%s: %d / %d steps" % (what, done, steps))
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64", "fp16"])
    ap.add_argument("--room", default="512,384,256")
    ap.add_argument("--cpu-steps", type=int, default=2, help="oracle steps for the CPU baseline sample (0: skip)")
    ap.add_argument("--accuracy-steps", type=int, default=0,
                    help="window of the FP32/FP16-vs-FP64 IR comparison (0: until the direct sound has passed "
                         "every receiver)")
    ap.add_argument("--no-fp64-check", action="store_true")
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--batch-jobs", type=int, default=32, help="jobs of the batched benchmark (0: skip)")
    ap.add_argument("--adjoint-steps", type=int, default=3000, help="steps of the adjoint benchmark (0: skip)")
    ap.add_argument("--tile-z", type=int, default=0, help="flux kernel z depth per block (0: autotune)")
    ap.add_argument("--tile-rows", type=int, default=0, help="flux kernel waves per block, 8 or 16 (0: autotune)")
    ap.add_argument("--tile-cols", type=int, default=0, help="flux kernel columns per block, 64 or 62 (0: autotune)")
    ap.add_argument("--no-geometry", action="store_true", help="skip the host vs GPU scene classification timing")
    ap.add_argument("--ga-rays", type=int, default=10000,
                    help="rays per source of the geometrical-acoustics benchmark (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the FVM core has no multi-GPU path in this build "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        return 2

    from fases_amd import workloads as W

    room = tuple(int(v) for v in args.room.split(","))
    total = args.warmup + args.steps
    t0 = time.time()
    wl = W.config(2, room=room, nsteps=total, precision=args.precision)
    acc_steps, last_arrival = accuracy_window(wl)
    if args.accuracy_steps > 0:
        acc_steps = args.accuracy_steps
    if acc_steps > total and not args.no_fp64_check:
        # the waveform must cover the accuracy window the run continues into
        wl = W.config(2, room=room, nsteps=acc_steps, precision=args.precision)
    if args.tile_z or args.tile_rows or args.tile_cols:
        import dataclasses
        wl = dataclasses.replace(wl, params=dataclasses.replace(wl.params, tile_z=args.tile_z, tile_rows=args.tile_rows,
                                                                tile_cols=args.tile_cols))
    sc = wl.scene
    log("scene %s padded %s built in %.1f s" % (room, sc.dims, time.time() - t0))

    sv = W.make_solver(wl, args.precision)
    info = sv.info()
    log("device bytes %.2f GB, boundary cells %d, updated cells %d" % (info["dev_bytes"] / 1e9, info["n_bnd"],
                                                                       info["n_upd"]))
    ir_w, _, ms_w, _ = sv.run(args.warmup)
    log("warmup %d steps %.1f ms" % (args.warmup, ms_w))
    t1 = time.time()
    ir_t, _, ms, kms = sv.run(args.steps, kernel_timing=True)
    wall = time.time() - t1
    tile = sv.tile_config()
    # the accuracy window continues the same run past the timed steps
    rest = max(0, acc_steps - total)
    ir_rest = chunked_run(sv, rest, "%s accuracy window" % args.precision) if (rest and not args.no_fp64_check) else []
    sv.close()
    cells = wl.room_cells
    mcells = cells * args.steps / (ms * 1e-3) / 1e6
    log("timed %d steps: %.1f ms (%.3f ms/step), flux kernels %.1f ms, host wall %.2f s -> %.0f Mcell/s"
        % (args.steps, ms, ms / args.steps, kms, wall, mcells))

    esz = ESZ[args.precision]
    launches = 3 * args.steps
    alg_bytes_launch = info["n_upd"] * 32 * esz / 3.0
    avg_launch_s = kms * 1e-3 / launches
    achieved = alg_bytes_launch / avg_launch_s / 1e9
    # HBM traffic and VALU instructions per launch come from rocprofv3 counter passes of the
    # same command (profiles/), which a profiler-free run cannot read: copied, with their source
    tag = "%s_%s" % ("x".join(map(str, room)), args.precision)
    pmc = profile_json("flux_stage_%s_traffic.json" % tag) or {}
    traffic = pmc.get("hbm_bytes_per_launch")
    valu_instr = pmc.get("valu_wave_instr_per_launch")
    from fases_amd.solver import hbm_copy_gbs
    copy_gbs = hbm_copy_gbs(0, 2 << 30, 5)
    log("streaming copy (2 GiB): %.0f GB/s" % copy_gbs)

    from fases_amd.metrics import decay_error_db, rel_l2
    win = max(acc_steps, total)
    ir = np.concatenate([ir_w, ir_t] + ir_rest, axis=1)
    ir_err, ir64, acc_note = None, None, None
    if not args.no_fp64_check and args.precision != "fp64":
        sv64 = W.make_solver(wl, "fp64")
        ir64 = np.concatenate(chunked_run(sv64, win, "fp64 reference"), axis=1)
        sv64.close()
        ir = ir[:, :win]
        if np.abs(ir64).max() < IR_FLOOR_PA:
            acc_note = "FP64 IR below %.0e Pa over the window: no signal to compare" % IR_FLOOR_PA
        else:
            ir_err = rel_l2(ir, ir64)
        log("IR rel L2 vs FP64 path over %d steps: %s (|IR|max %.3e)" % (win, ir_err, np.abs(ir64).max()))
    accuracy = {"window_steps": win, "last_direct_arrival_step": last_arrival,
                "ir64_max_pa": float(np.abs(ir64).max()) if ir64 is not None else None, "note": acc_note}

    fp16 = None
    if not args.no_fp16 and args.precision == "fp32":
        sv16 = W.make_solver(wl, "fp16")
        ir16w, _, _, _ = sv16.run(args.warmup)
        ir16t, _, ms16, kms16 = sv16.run(args.steps, kernel_timing=True)
        ir16r = chunked_run(sv16, max(0, win - total), "fp16 accuracy window") if ir64 is not None else []
        log("fp16 storage run done")
        sv16.close()
        ir16 = np.concatenate([ir16w, ir16t] + ir16r, axis=1)[:, :ir.shape[1]]
        ok64 = ir64 is not None and ir_err is not None
        fp16 = {"value": round(cells * args.steps / (ms16 * 1e-3) / 1e6, 2), "unit": "Mcell-updates/s",
                "ms_per_step": round(ms16 / args.steps, 5),
                "flux_roofline_frac": round(info["n_upd"] * 32 * 2 / 3.0 / (kms16 * 1e-3 / launches) / 1e9
                                            / HBM_PEAK_GBS, 4),
                "ir_rel_l2_vs_fp32": rel_l2(ir16, ir) if ok64 else None,
                "ir_rel_l2_vs_fp64": rel_l2(ir16, ir64) if ok64 else None,
                "decay_err_db_vs_fp64": decay_error_db(ir16, ir64) if ok64 else None,
                "decay_err_db_vs_fp32": decay_error_db(ir16, ir) if ok64 else None,
                "rescale_every": wl.params.fp16_rescale_every,
                "note": "same room and accuracy window (%d steps); decay error = max |EDC difference| down "
                        "to -40 dB" % ir16.shape[1]}
        log("fp16 storage: %.0f Mcell/s, IR rel L2 vs FP64 %s, vs FP32 %s"
            % (fp16["value"], fp16["ir_rel_l2_vs_fp64"], fp16["ir_rel_l2_vs_fp32"]))

    batch = None
    if args.batch_jobs > 0:
        from fases_amd.solver import batch_run
        jobs = W.batch_jobs(args.batch_jobs)
        work = sum(int(np.prod(j.room)) * j.nsteps for j in jobs)
        # device time from events on each side, and host wall time around the same calls
        # (run set-up, autotuning, graph capture and the IR copies included)
        seq_ms, seq_ir, seq_wall = 0.0, [], 0.0
        for j in jobs:
            s1 = W.make_solver(j)
            tw = time.time()
            r, _, el, _ = s1.run(j.nsteps)
            seq_wall += time.time() - tw
            s1.close()
            seq_ms += el
            seq_ir.append(r)
        log("batched benchmark: %d jobs sequential %.0f ms" % (len(jobs), seq_ms))
        svs = [W.make_solver(j) for j in jobs]
        tw = time.time()
        b_ir, b_ms = batch_run(svs, [j.nsteps for j in jobs])
        b_wall = time.time() - tw
        for s1 in svs:
            s1.close()
        same = all(np.array_equal(a, b) for a, b in zip(b_ir, seq_ir))
        batch = {"jobs": len(jobs), "cell_steps": work, "batched_ms": round(b_ms, 1),
                 "sequential_ms": round(seq_ms, 1), "speedup": round(seq_ms / b_ms, 3),
                 "batched_wall_s": round(b_wall, 2), "sequential_wall_s": round(seq_wall, 2),
                 "speedup_wall": round(seq_wall / b_wall, 3),
                 "batched_value": round(work / (b_ms * 1e-3) / 1e6, 1),
                 "sequential_value": round(work / (seq_ms * 1e-3) / 1e6, 1), "unit": "Mcell-updates/s",
                 "irs_bitwise_equal": bool(same),
                 "workload": "seed 1003: rooms U[96..256]x[96..192]x[96..160] cells, dx 0.02, 1-3 wall types, "
                             "Ricker f_p U[500,4000] Hz, 1500-4000 steps, 2 receivers, FP32; sequential = "
                             "fases_run per job (autotuned tile), batched = fases_batch_run (hipGraph replay)"}
        log("batched %.0f ms vs sequential %.0f ms: x%.2f, IRs identical %s" % (b_ms, seq_ms, seq_ms / b_ms, same))

    adjoint = None
    if args.adjoint_steps > 0:
        # 192x160x128 room, 3 materials (4 real poles + 1 pair), Ricker source, 8 receivers,
        # FP32: plain forward vs gradient (checkpointed forward + reverse sweep)
        wa = W.config(4, nsteps=args.adjoint_steps)
        sa = W.make_solver(wa, "fp32")
        sa.run(20)
        sa.set_fields(wa.initial_pressure())
        ir_a, _, fwd_ms, _ = sa.run(wa.nsteps)
        sa.set_fields(wa.initial_pressure())
        t_a = time.time()
        ir_c = sa.adjoint_forward(wa.nsteps, 64)
        t_b = time.time()
        gm_a, gs_a, _ = sa.adjoint(2.0 * ir_c)
        t_c = time.time()
        info_a = sa.info()
        sa.close()
        lin = 2.0 * float((ir_c ** 2).sum()) / wa.src_sval[0]
        adjoint = {"workload": "192x160x128 room dx 0.02, 3 materials (K=4, L=1), Ricker source, 8 receivers, FP32 "
                               "(FP64 within radius %d of the source), %d steps, checkpoints every 64 steps"
                               % (wa.params.source_patch, wa.nsteps),
                   "forward_ms": round(fwd_ms, 1), "checkpointed_forward_ms": round((t_b - t_a) * 1e3, 1),
                   "reverse_ms": round((t_c - t_b) * 1e3, 1),
                   "gradient_over_forward": round((t_c - t_a) * 1e3 / fwd_ms, 3),
                   "checkpoint_and_slot_field_bytes": int(16 * info_a["n_cells"] * (3 * 64 + 1 + -(-wa.nsteps // 64))),
                   "drive_gradient": float(gs_a[0]), "drive_gradient_linear_estimate": lin,
                   "note": "J = sum ir^2; the IR is nearly linear in the drive, so dJ/dsval ~ 2 J / sval; "
                           "gradient and FD / FP64 agreement are tested in tests/test_gpu_adjoint.py and, on "
                           "this room at full length, in tests/test_gpu_adjoint_192x160x128.py (FP64 vs FD, FP32 "
                           "vs FP64); FP64 radius 32 around the source (source_patch)"}
        log("adjoint: forward %.0f ms, gradient %.0f ms (x%.2f)" % (fwd_ms, (t_c - t_a) * 1e3,
                                                                  adjoint["gradient_over_forward"]))

    geometry = None
    if not args.no_geometry:
        geometry = geometry_timing(room, wl.h)

    ga = ga_timing(args.ga_rays) if args.ga_rays > 0 else None

    cpu = None
    if args.cpu_steps > 0:
        from oracle import oracle as O
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        t2 = time.time()
        O.run_workload(wl, nsteps=args.cpu_steps)
        el = time.time() - t2
        # the 64x48x40 rigid box in full (600 steps) on the CPU and on the GPU's FP64 path
        w0 = W.config(0)
        t3 = time.time()
        r0 = O.run_workload(w0)["ir"]
        el0 = time.time() - t3
        s0 = W.make_solver(w0, "fp64")
        g0, _, ms0, _ = s0.run(w0.nsteps)
        s0.close()
        c0 = {"workload": "rigid box 64x48x40 dx 0.05, Gaussian pulse, 600 steps, FP64 (full run)",
              "cpu_value": round(w0.room_cells * w0.nsteps / el0 / 1e6, 3), "cpu_s": round(el0, 2),
              "gpu_fp64_value": round(w0.room_cells * w0.nsteps / (ms0 * 1e-3) / 1e6, 1),
              "gpu_vs_cpu_ir_rel_l2": rel_l2(g0, r0)}
        cpu = {"value": round(cells * args.cpu_steps / el / 1e6, 3), "unit": "Mcell-updates/s", "cores": threads,
               "kind": "port", "cpu_model": cpu_model(),
               "sample": "FP64 oracle (OpenMP) on the same %s room, first %d steps, %.1f s"
                         % ("x".join(map(str, room)), args.cpu_steps, el),
               "rigid_box_full": c0}
        log("cpu baseline %.2f Mcell/s on %d threads (%s); rigid box in full %.1f s, GPU FP64 vs CPU IR %.1e"
            % (cpu["value"], threads, cpu["cpu_model"], el0, c0["gpu_vs_cpu_ir_rel_l2"]))

    out = {
        "metric": METRIC,
        "value": round(mcells, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic",
        "config": {"workload": "shoebox %s cells dx=0.01 m, %s storage, 16-cell cPML +x, 5 ADE walls "
                               "(3 types, K=4 L=2), Ricker 2 kHz, 8 receivers" % ("x".join(map(str, room)),
                                                                                   args.precision.upper()),
                   "room_cells": cells, "padded_cells": info["n_cells"], "boundary_cells": info["n_bnd"],
                   "updated_cells": info["n_upd"], "parallelism": "single GPU",
                   "tile": "%dx%dx%d" % tile},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this command, scaled by the "
                                        "copy calibration, copied from profiles/flux_stage_%s_traffic.json" % tag)
                                       if traffic else None,
                     "kernel": "flux stage (WENO5-Z + Godunov + SSP-RK3 + cPML)",
                     "alg_bytes_per_launch": int(alg_bytes_launch),
                     "avg_launch_us": round(avg_launch_s * 1e6, 1),
                     "avg_launch_source": "HIP events around the flux stages of every 16th timed step "
                                          "(the stream they run on)",
                     "copy_kernel_gbs": round(copy_gbs, 1),
                     "frac_of_copy": round(achieved / copy_gbs, 4) if copy_gbs > 0 else None,
                     "valu": ({"wave_instr_per_launch": valu_instr,
                               "lane_ops_per_updated_cell": round(valu_instr * 64 / info["n_upd"], 1),
                               "achieved_lane_ops_per_s": valu_instr * 64 / avg_launch_s,
                               "peak_lane_ops_per_s": VALU_PEAK_LANE_OPS,
                               "frac": round(valu_instr * 64 / avg_launch_s / VALU_PEAK_LANE_OPS, 4),
                               "source": "SQ_INSTS_VALU per launch from profiles/flux_stage_%s_traffic.json, "
                                         "launch time from this run" % tag} if valu_instr else None)},
        "ir_rel_l2_vs_fp64": ir_err,
        "accuracy": accuracy,
        "fp16_storage": fp16,
        "batched": batch,
        "adjoint": adjoint,
        "geometry": geometry,
        "geometrical_acoustics": ga,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
