"""The 512 x 384 x 256 room at full size (dx 0.01 m, 16-cell cPML on +x, five ADE walls
of three types, Ricker 2 kHz, 8 receivers): the FP32- and FP16-storage IRs against the
FP64 path over the first 400 steps (the FP64 path is the oracle's arithmetic bit for bit
on the smaller rooms of tests/test_gpu_parity.py; at this size the oracle itself would
take hours on the host). Tolerances: FP32 relative L2 <= 1e-5; FP16 storage <= 2e-3."""
import numpy as np
import pytest

from fases_amd import workloads as W
from fases_amd.metrics import rel_l2

pytestmark = pytest.mark.gpu

STEPS = 400


@pytest.fixture(scope="module")
def room():
    wl = W.config(2, nsteps=STEPS, precision="fp32")
    print("[full] scene", wl.room, flush=True)
    sv = W.make_solver(wl, "fp64")
    ir64 = sv.run(STEPS)[0]
    sv.close()
    print("[full] fp64 done, |IR|max %.3e" % np.abs(ir64).max(), flush=True)
    return wl, ir64


def test_fp32_full_room_matches_fp64(room):
    wl, ir64 = room
    sv = W.make_solver(wl, "fp32")
    ir = sv.run(STEPS)[0]
    sv.close()
    assert np.all(np.isfinite(ir)) and np.abs(ir64).max() > 0
    err = rel_l2(ir, ir64)
    print("[full] fp32 IR rel L2 %.3e" % err, flush=True)
    assert err <= 1e-5


def test_fp16_storage_full_room_matches_fp64(room):
    wl, ir64 = room
    sv = W.make_solver(wl, "fp16")
    ir = sv.run(STEPS)[0]
    sv.close()
    assert np.all(np.isfinite(ir))
    err = rel_l2(ir, ir64)
    print("[full] fp16 IR rel L2 %.3e" % err, flush=True)
    assert err <= 2e-3
