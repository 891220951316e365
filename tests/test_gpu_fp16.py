"""FP16 field storage with per-field power-of-two scales (FP32 arithmetic, FP64 ADE
states, receivers and source patch), against the FP64 oracle and the FP32 path.

FP16 keeps 11 significant bits, so the bar here is a reported accuracy, not parity:
IR relative L2 <= 1e-3 and the energy-decay curve within 0.5 dB down to -40 dB."""
import dataclasses
import sys

import numpy as np
import pytest

from fases_amd import workloads as W
from fases_amd.metrics import decay_error_db, rel_l2
from fases_amd._lib import FasesError
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def announce(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def small(nsteps=400, **kw):
    return W.ade_room(room=(40, 32, 24), h=0.02, nsteps=nsteps, seed=7, n_rec=3, K=4, L=2, pml_cells=8, **kw)


def run(wl, prec):
    sv = W.make_solver(wl, prec)
    ir, _, _, _ = sv.run(wl.nsteps)
    sv.close()
    return ir


def test_fp16_set_get_round_trip_rounds_once():
    """Host fields enter as round-to-nearest-even binary16 of value * 2^-e with e chosen so
    the field's largest |value| is stored in [32, 64); get_fields returns stored * 2^e."""
    wl = small()
    announce("fp16 set/get round trip")
    sv = W.make_solver(wl, "fp16")
    rng = np.random.default_rng(0)
    N = int(np.prod(wl.dims))
    x = [rng.standard_normal(N) * a for a in (3.0, 1e-3, 2e-2, 5e-5)]
    x[1][:5] = [0.0, 1e-30, -1e-30, 2.0 ** -40, 7e-9]          # zeros and subnormal-range values
    sv.set_fields(*x)
    y = sv.get_fields()
    sv.close()
    for q in range(4):
        e = np.frexp(np.abs(x[q]).max())[1] - 6
        ref = (x[q] * 2.0 ** -e).astype(np.float16).astype(np.float64) * 2.0 ** e
        assert np.array_equal(y[q], ref), q
        assert 32.0 <= np.abs(x[q]).max() * 2.0 ** -e < 64.0


def test_fp16_ade_room_vs_fp64_and_fp32():
    wl = small()
    announce("fp16 ADE room %s" % (wl.dims,))
    ref = O.run_workload(wl)["ir"]
    h16, h32 = run(wl, "fp16"), run(wl, "fp32")
    e16, e32 = rel_l2(h16, ref), rel_l2(h32, ref)
    d16 = decay_error_db(h16, ref)
    print("fp16 IR rel L2 %.3e (fp32 %.3e), decay error %.3f dB" % (e16, e32, d16))
    assert e16 <= 1e-3 and d16 <= 0.5
    assert e32 <= 1e-5


@pytest.mark.parametrize("period", [7, 1000])
def test_fp16_rescale_period(period):
    """Rescaling is exact (powers of two), so the period changes only when values round."""
    wl = small(nsteps=300)
    w2 = dataclasses.replace(wl, params=dataclasses.replace(wl.params, fp16_rescale_every=period))
    announce("fp16 rescale period %d" % period)
    ref = O.run_workload(wl)["ir"]
    assert rel_l2(run(w2, "fp16"), ref) <= 1e-3


def test_fp16_decaying_tail():
    """A rigid-walled room with one absorbing wall: the field decays by orders of magnitude
    and the rescale keeps the FP16 values normal, so the decay curve follows FP64."""
    wl = W.ade_room(room=(36, 30, 24), h=0.02, nsteps=1500, seed=11, n_rec=2, K=2, L=1, pml_cells=8)
    announce("fp16 decaying tail %s %d steps" % (wl.dims, wl.nsteps))
    ref = O.run_workload(wl)["ir"]
    h16 = run(wl, "fp16")
    d = decay_error_db(h16, ref, -40.0)
    print("fp16 decay error to -40 dB: %.3f dB; IR rel L2 %.3e" % (d, rel_l2(h16, ref)))
    assert d <= 0.5


def test_fp16_overflow_is_an_error():
    """A drive far beyond the FP16 range overflows the stored source cell; the first
    rescale pass finds the non-finite value and fases_run fails (no silent Inf)."""
    wl = small(nsteps=40)
    wl = dataclasses.replace(wl, src_sval=[1e12])
    announce("fp16 overflow")
    sv = W.make_solver(wl, "fp16")
    with pytest.raises(FasesError, match="overflow"):
        sv.run(wl.nsteps)
    sv.close()
