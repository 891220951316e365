"""The 256 x 192 x 128 room at full length (dx 0.02 m, 16-cell cPML on +x, five ADE walls
each with 4 real poles and 2 complex pairs drawn from seed 1001, Ricker 2 kHz, 4 receivers,
4000 steps at CFL 0.5): FP32-storage IRs against the FP64 path over all 4000 steps. The
FP64 path is the oracle's arithmetic bit for bit on the rooms of tests/test_gpu_parity.py;
here the oracle itself would take about 40 minutes on the host. Bars: relative L2 <= 1e-5
and the energy-decay curves within 0.01 dB down to -40 dB."""
import sys

import numpy as np
import pytest

from fases_amd import workloads as W
from fases_amd.metrics import decay_error_db, rel_l2

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
