"""The 256 x 192 x 128 room at full length (dx 0.02 m, 16-cell cPML on +x, five ADE walls
each with 4 real poles and 2 complex pairs drawn from seed 1001, Ricker 2 kHz, 4 receivers,
4000 steps at CFL 0.5): FP32-storage IRs against the FP64 path over all 4000 steps. The
FP64 path is the oracle's arithmetic bit for bit on the rooms of tests/test_gpu_parity.py;
here the oracle itself would take about 40 minutes on the host. Bars: relative L2 <= 1e-5
and the energy-decay curves within 0.01 dB down to -40 dB. A second test runs the oracle
itself at this shape for a 60-step window."""
import dataclasses
import sys

import numpy as np
import pytest

from fases_amd import workloads as W
from fases_amd.metrics import decay_error_db, rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu

STEPS = 4000


def run_chunked(wl, prec, chunk=1000):
    sv = W.make_solver(wl, prec)
    out = []
    for s0 in range(0, STEPS, chunk):
        out.append(sv.run(min(chunk, STEPS - s0))[0])
        print("[256x192x128] %s %d / %d steps" % (prec, s0 + len(out[-1][0]), STEPS), flush=True)
        sys.stdout.flush()
    sv.close()
    return np.concatenate(out, axis=1)


def test_fp32_matches_fp64_over_4000_steps():
    wl = W.config(1, precision="fp32")
    assert wl.room == (256, 192, 128) and wl.nsteps == STEPS
    ir64 = run_chunked(wl, "fp64")
    ir32 = run_chunked(wl, "fp32")
    assert np.all(np.isfinite(ir32)) and np.abs(ir64).max() > 0
    err = rel_l2(ir32, ir64)
    dec = decay_error_db(ir32, ir64)
    print("[256x192x128] fp32 vs fp64: IR rel L2 %.3e, decay error %.2e dB" % (err, dec), flush=True)
    assert err <= 1e-5
    assert dec <= 0.01


def test_fp64_equals_oracle_at_full_shape():
    """The FP64 GPU path against the FP64 oracle at this room's full shape (6.95 M padded
    cells), 60 steps: the Ricker source plus a Gaussian pulse (sigma 3 cells, 1 Pa) placed
    6 cells from the +x cPML band and 7 cells from the y- and z- ADE walls, so that the
    cPML, the ADE walls and the source all act within the window. Bar: every field's
    relative L2 <= 1e-10 (the same arithmetic, bit for bit in practice)."""
    wl = W.config(1, nsteps=60)
    Nx, Ny, Nz = wl.dims
    centre = (Nx - 3 - 16 - 6, 3 + 7, 3 + 7)
    wl = dataclasses.replace(wl, ic={"center": centre, "sigma": 3.0, "amp": 1.0})
    sv = W.make_solver(wl, "fp64")
    ir, _, _, _ = sv.run(wl.nsteps)
    fields = sv.get_fields()
    sv.close()
    print("[256x192x128] fp64 GPU done; oracle on the host ...", flush=True)
    ref = O.run_workload(wl, want_fields=True)
    errs = [rel_l2(fields[f], ref["fields"][f]) for f in range(4)]
    print("[256x192x128] fp64 GPU vs oracle, fields rel L2 %s, IR max |diff| %.2e" %
          (" ".join("%.2e" % e for e in errs), np.abs(ir - ref["ir"]).max()), flush=True)
    assert all(np.abs(fields[f]).max() > 1e-3 for f in range(4))
    assert max(errs) <= 1e-10
    assert np.abs(ir - ref["ir"]).max() <= 1e-10 * np.abs(fields[0]).max()
