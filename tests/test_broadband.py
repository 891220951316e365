"""Broadband IR post-processing (fases_amd/broadband.py) against vectors of the
reference's own common/allIRs.py (tests/golden/broadband_ref.npz, written by
tests/golden/make_broadband_golden.py), plus the identities the method rests on: the
deconvolution of a convolution with the pulse is the Tikhonov filter of the room
response (and recovers it inside the pulse's band), the
cosine shoulders are amplitude- / power-complementary, the filter with no edge is the
identity. No GPU."""
import os
from pathlib import Path

import numpy as np
import pytest

from fases_amd import broadband as B

GOLD = Path(__file__).resolve().parent / "golden" / "broadband_ref.npz"


def _close(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    scale = max(float(np.abs(b).max()), 1e-300)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12 * scale)


def _project(tmp: Path, g, cfg_key="cfg", header=True, pref=True, only=None):
    (tmp / "config").mkdir(parents=True)
    out = tmp / "output"
    out.mkdir()
    (tmp / "config" / "config.txt").write_text("".join("%r\n" % float(v) for v in g[cfg_key]))
    if header:
        (out / "header.txt").write_text("46 36 31 600 0.92 0.72 0.62 0 5 %f scene.obj 0.02 3\n" % float(g["dt_header"]))
    if pref:
        g["pref"].astype(np.float64).tofile(tmp / "config" / "p_ref.bin")
    for k in g.files:
        if k.startswith("in/") and (only is None or k[3:].startswith(only)):
            g[k].tofile(out / k[3:])
    return str(out)


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def test_pulse_and_deconvolution_match_reference(tmp_path, gold):
    out = _project(tmp_path, gold)
    run = B.load_run(out)
    _close(run.pulse_fs, gold["full/pulse_fs"])
    assert run.lag == int(gold["full/lag"])
    _close(B.low_deconv(gold["in/IR0_0.bin"], run), gold["full/low_deconv_IR0"])


def test_broadband_pairs_match_reference(tmp_path, gold):
    out = _project(tmp_path, gold)
    for rec in (0, 1, 2):                     # pair 2 has geometrical bands only
        bb, lo, hi = B.broadband_ir(out, rec, 0, log=lambda *a: None)
        _close(bb, gold["full/bb%d" % rec])
        _close(lo, gold["full/lo%d" % rec])
        _close(hi, gold["full/hi%d" % rec])
    assert not lo.any()


def test_delay_fallback_without_pulse_matches_reference(tmp_path, gold):
    out = _project(tmp_path, gold, pref=False)
    assert B.load_run(out).lag == int(gold["nopref/lag"])
    bb, lo, hi = B.broadband_ir(out, 0, 0, log=lambda *a: None)
    _close(bb, gold["nopref/bb0"])
    _close(lo, gold["nopref/lo0"])


def test_bands_only_run_matches_reference(tmp_path, gold):
    out = _project(tmp_path, gold, cfg_key="cfg_ga", header=False, pref=False, only="HIR0_0_")
    bb, lo, hi = B.broadband_ir(out, 0, 0, log=lambda *a: None)
    _close(bb, gold["gaonly/bb0"])
    assert not lo.any() and np.array_equal(bb, hi)


def test_filter_matches_reference(gold):
    x = gold["filt/x"]
    edges = [(None, 707.1067811865476, False, True), (707.1067811865476, 1414.2, True, False),
             (1414.2, 2828.4, False, False), (5656.85, None, False, False), (None, None, False, False)]
    for i, (hp, lp, hpp, lpp) in enumerate(edges):
        _close(B.band_filter(x, hp, lp, hp_pow=hpp, lp_pow=lpp), gold["filt/y%d" % i])


def test_deconvolution_is_the_tikhonov_filter_of_the_room_response():
    """low_deconv(h * q) = h filtered by |Q|^2 / (|Q|^2 + reg max |Q|^2): with the full
    linear convolution inside the zero-padded transform the deconvolution is exact up to
    the regularisation, at an FVM rate equal to the output rate (no resampling). With a
    small reg, h comes back wherever the pulse has energy."""
    rng = np.random.default_rng(0)
    x = (np.arange(240) - 120) / 14.0
    q = (1 - x ** 2) * np.exp(-x ** 2 / 2)
    q = q / np.abs(q).max()
    run = B.Run(f_cutoff_sim=1000.0, f_seam=B.seam(1000.0), src_opt=1, dt=1.0 / B.FS, pulse_fs=q)
    h = rng.standard_normal(4000)
    y = np.convolve(h, q)
    n = 2 * y.size
    P = np.abs(np.fft.rfft(q, n=n)) ** 2
    for reg in (1e-2, 1e-6):
        want = np.fft.irfft(np.fft.rfft(h, n=n) * P / (P + reg * P.max()), n=n)[:y.size]
        np.testing.assert_allclose(B.low_deconv(y, run, reg=reg), want, atol=1e-9 * np.abs(want).max())
    # a room response inside the pulse's band is recovered (all but the lowest bins: a
    # Ricker pulse has no energy at DC)
    hb = np.fft.irfft(np.fft.rfft(h) * (np.fft.rfftfreq(h.size) < 0.02), n=h.size)
    hr = B.low_deconv(np.convolve(hb, q), run, reg=1e-6)[:h.size]
    assert np.corrcoef(hb, hr)[0, 1] > 0.99


def test_shoulders_are_complementary():
    f = np.linspace(0, 22050, 40001)
    for e in (88.4, 707.1, 5656.9):
        hi, lo = B.shoulder(f, e, True), B.shoulder(f, e, False)
        np.testing.assert_allclose(hi + lo, 1.0, atol=1e-15)
        np.testing.assert_allclose(np.sqrt(hi) ** 2 + np.sqrt(lo) ** 2, 1.0, atol=1e-15)
        assert hi[f <= e / 2 ** (1 / 6)].max() == 0 and lo[f >= e * 2 ** (1 / 6)].max() == 0
    x = np.random.default_rng(1).standard_normal(999)
    np.testing.assert_allclose(B.band_filter(x, None, None), x, atol=1e-12)


def test_seam_rule():
    assert B.seam(500.0) == 500 * 2 ** 0.5 and B.seam(353.0) == 250 * 2 ** 0.5 and B.seam(20.0) == 63 * 2 ** 0.5
    assert B.seam(20000.0) == 20000.0


def test_write_all_and_fp64_input(tmp_path, gold):
    out = _project(tmp_path, gold)
    Path(out, "IR_broad_9_9.npy").write_bytes(b"stale")
    done = B.write_all(out, log=lambda *a: None)
    assert done == [(0, 0), (1, 0), (2, 0)]
    assert not os.path.exists(os.path.join(out, "IR_broad_9_9.npy"))
    _close(np.load(os.path.join(out, "IR_broad_0_0.npy")), gold["full/bb0"])
    # an FP64 receiver IR replaces the float32 file where asked
    ir64 = gold["in/IR0_0.bin"].astype(np.float64) * (1 + 1e-7)
    ir64.tofile(os.path.join(out, "IR0_0.f64.bin"))
    assert B.discover_pairs(out) == [(0, 0), (1, 0), (2, 0)]
    lo32 = B.broadband_ir(out, 0, 0, log=lambda *a: None)[1]
    lo64 = B.broadband_ir(out, 0, 0, fp64=True, log=lambda *a: None)[1]
    run = B.load_run(out)
    want = B.band_filter(B.low_deconv(ir64, run), None, run.f_seam, lp_pow=True)
    _close(lo64[:want.size], want)
    assert not lo64[want.size:].any()          # padded to the geometrical bands' length
    assert not np.array_equal(lo32, lo64)
