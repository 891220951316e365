"""Batched multi-job loop (fases_batch_run) and the tile autotuner.

Each job of a batch runs the same per-cell arithmetic as a single-context fases_run, so
the bar is bit-for-bit equality with the single runs (IRs and the final fields, which
also checks the host-side rotation bookkeeping after odd and even step counts)."""
import dataclasses
import sys

import numpy as np
import pytest

from fases_amd import _lib
from fases_amd import workloads as W
from fases_amd._lib import FasesError
from fases_amd.solver import batch_run

pytestmark = pytest.mark.gpu


def announce(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def jobs():
    out = []
    for room, n, seed, pml in (((40, 30, 24), 61, 1, True), ((52, 26, 30), 40, 2, False), ((30, 34, 20), 75, 3, True)):
        out.append(W.ade_room(room=room, h=0.02, nsteps=n, seed=seed, n_rec=2, K=2, L=1, pml_cells=6, pml_x_hi=pml,
                              precision="fp32"))
    # a sine-driven job and one without sources (initial pulse only)
    s = out[1]
    out[1] = dataclasses.replace(s, src_opt=_lib.FASES_SRC_SINE, u_s=None, src_sval=[2.0], src_freq=[700.0])
    r = W.rigid_box(room=(26, 22, 18), h=0.05, nsteps=33, ic_center=(8, 8, 8), rec=(18, 14, 10))
    out.append(dataclasses.replace(r, params=dataclasses.replace(r.params, precision="fp32")))
    return out


EXTRA = 5   # steps run after the batch (the workloads' waveforms cover them)


def test_batch_equals_single_runs_bitwise():
    js = jobs()
    steps = [wl.nsteps - EXTRA for wl in js]
    announce("batch of %d jobs, steps %s" % (len(js), steps))
    single_ir, single_f = [], []
    for wl, n in zip(js, steps):
        sv = W.make_solver(wl, "fp32")
        ir, _, _, _ = sv.run(n)
        single_ir.append(ir)
        single_f.append(sv.get_fields())
        sv.close()
    svs = [W.make_solver(wl, "fp32") for wl in js]
    irs, ms = batch_run(svs, steps)
    assert ms > 0
    for q, sv in enumerate(svs):
        assert np.array_equal(irs[q], single_ir[q]), q
        f = sv.get_fields()
        for a, b in zip(f, single_f[q]):
            assert np.array_equal(a, b), q
    # the jobs continue from the batch's end state like a single context would
    more = [sv.run(EXTRA)[0] for sv in svs]
    for sv in svs:
        sv.close()
    for q, wl in enumerate(js):
        sv = W.make_solver(wl, "fp32")
        ir, _, _, _ = sv.run(wl.nsteps)
        sv.close()
        assert np.array_equal(ir[:, steps[q]:], more[q]), q


def test_tiled_result_independent_of_tile_shape():
    """Every face is computed by the same operations whichever block holds it, so the
    width (64 columns with edge work, 62 without: the x seams fall on different cells),
    height and z depth of a block change nothing in the result."""
    wl = W.ade_room(room=(150, 30, 70), h=0.02, nsteps=30, seed=4, n_rec=3, K=2, L=1, pml_cells=6, precision="fp32")
    announce("tile shape independence")
    res = []
    for cols, rows, tz in ((64, 16, 16), (64, 16, 32), (64, 16, 64), (64, 16, 128), (64, 8, 32), (62, 8, 64),
                           (62, 16, 16), (62, 8, 16)):
        w2 = dataclasses.replace(wl, params=dataclasses.replace(wl.params, tile_z=tz, tile_rows=rows, tile_cols=cols))
        sv = W.make_solver(w2, "fp32")
        ir, _, _, _ = sv.run(wl.nsteps)
        assert sv.tile_config() == (cols, rows, tz)
        res.append((ir, sv.get_fields()[0]))
        sv.close()
    for ir, p in res[1:]:
        assert np.array_equal(ir, res[0][0]) and np.array_equal(p, res[0][1])


def test_autotuner_picks_a_candidate():
    wl = W.ade_room(room=(40, 30, 24), h=0.02, nsteps=4, seed=1, n_rec=1, K=2, L=1, pml_cells=6, precision="fp32")
    announce("autotuner")
    sv = W.make_solver(wl, "fp32")
    assert sv.tile_config()[2] == 0
    sv.run(2)
    tx, ty, tz = sv.tile_config()
    sv.close()
    assert tx in (62, 64) and ty in (8, 16) and (tz in (16, 32, 64, 128) or 4 <= tz < 16)


def test_batch_rejects_bad_input():
    wl = jobs()[0]
    announce("batch errors")
    a = W.make_solver(wl, "fp32")
    b = W.make_solver(wl, "fp64")
    with pytest.raises(FasesError, match="FP32"):
        batch_run([a, b], [3, 3])
    with pytest.raises(FasesError, match="twice"):
        batch_run([a, a], [3, 3])
    a.close()
    b.close()


def test_seeded_32_job_batch_equals_sequential_runs():
    """The 32-job batched benchmark at full length (seed 1003: rooms of 96^3 .. 256x192x160
    cells, 1-3 wall types, Ricker 500-4000 Hz, 1500-4000 steps, FP32; the jobs drop out of
    the batch at different steps): the batched IRs equal the jobs run one after the other
    through fases_run, bit for bit. A difference that starts at one wall cell late in a run
    (a product fused into an add in one kernel and not in the other) reaches the receivers
    only hundreds of steps later, so shortened runs do not show it."""
    js = W.batch_jobs(32)
    steps = [wl.nsteps for wl in js]
    announce("seeded 32-job batch, %d-%d steps per job" % (min(steps), max(steps)))
    seq = []
    for q, (wl, n) in enumerate(zip(js, steps)):
        sv = W.make_solver(wl, "fp32")
        ir, _, _, _ = sv.run(n)
        sv.close()
        seq.append(ir)
        if q % 8 == 7:
            announce("sequential: %d / 32 jobs" % (q + 1))
    svs = [W.make_solver(wl, "fp32") for wl in js]
    irs, ms = batch_run(svs, steps)
    for sv in svs:
        sv.close()
    announce("batched: %.0f ms" % ms)
    for q in range(len(js)):
        assert irs[q].shape == seq[q].shape, q
        assert np.array_equal(irs[q], seq[q]), q
    # the jobs hear their sources within the window (the comparison is not of zeros)
    assert sum(np.abs(ir).max() > 1e-6 for ir in irs) >= 24


@pytest.fixture(scope="module")
def small_batch():
    specs = [((40, 30, 24), 160, 31, True), ((36, 44, 28), 170, 32, False), ((48, 26, 30), 150, 33, True),
             ((30, 30, 30), 180, 34, False)]
    js = [W.ade_room(room=room, h=0.02, nsteps=n, seed=seed, n_rec=2, K=2, L=1, pml_cells=6, pml_x_hi=pml,
                     precision="fp32") for room, n, seed, pml in specs]
    announce("batch of 4 uneven jobs")
    svs = [W.make_solver(wl, "fp32") for wl in js]
    irs, _ = batch_run(svs, [wl.nsteps for wl in js])
    for sv in svs:
        sv.close()
    return js, irs


@pytest.mark.parametrize("q", [0, 1, 2, 3])
def test_batched_jobs_match_fp64_oracle(small_batch, q):
    """Batched FP32 jobs of uneven size against the FP64 oracle: IR relative L2 <= 1e-5."""
    from oracle import oracle as O
    js, irs = small_batch
    announce("batched job %d vs FP64 oracle" % q)
    ref = O.run_workload(js[q])["ir"]
    assert np.abs(ref).max() > 1e-3
    err = np.linalg.norm(irs[q] - ref) / np.linalg.norm(ref)
    print("job %d: IR rel L2 vs FP64 oracle %.2e" % (q, err))
    assert err <= 1e-5
