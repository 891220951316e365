"""Project formats and source waveforms (host C++ via the C ABI), pinned by the
reference's own fixtures and by closed forms / library routines."""
import numpy as np
import pytest
import scipy.signal

from fases_amd import _lib
from fases_amd.solver import (read_config, read_fit_files, read_poles, read_receivers, read_sources,
                              source_pref, source_waveform, write_f32)
from fases_amd._lib import FasesError


def test_config_default_project(golden):
    v = read_config(str(golden / "default_config.txt"))
    assert v.dtype == np.float32
    assert list(v[:5]) == [np.float32(0.0275), np.float32(0.7), np.float32(1000.0), np.float32(22.0), np.float32(0.5)]
    assert v.size == 12 and v[11] == 1.0 and v[6] == 0.0


def test_config_comments_and_blank_lines(tmp_path):
    p = tmp_path / "config.txt"
    p.write_text("0.05 # h\n\n# a comment line\n  0.7\t# tmax\nabc\n1e3\n")
    v = read_config(str(p))
    # lines without a leading float are skipped, as in main3D.cu:114-128
    assert list(v) == [np.float32(0.05), np.float32(0.7), np.float32(1000.0)]


def test_config_missing_file():
    with pytest.raises(FasesError) as e:
        read_config("/nonexistent/config.txt")
    assert e.value.code == _lib.FASES_E_IO


def test_poles_floor(tmp_path, golden):
    assert read_poles(str(golden / "bras_rs1" / "config" / "poles.txt")) == (2, 5)
    p = tmp_path / "poles.txt"
    p.write_text("0 0\n")
    assert read_poles(str(p)) == (1, 1)


def test_fit_sample_known_values(golden):
    m = read_fit_files([str(golden / "ade_benchmark" / "sample.fit")], 2, 3)
    assert m.Rinf[0] == 0.5021750604967876
    assert m.A[0].tolist() == [328.5183890078403, -0.46803927460631534]
    assert m.zeta[0].tolist() == [1071.6694737064222, 145.55896191182228]
    assert m.beta[0, 2] == 529.8926564688727 and m.C[0, 0] == 123.06056772771578


def test_fit_zero_poles_and_padding(golden):
    m = read_fit_files([str(golden / "bras_fit" / "default_Material0.fit"),
                        str(golden / "bras_fit" / "default_Material1.fit"), None], 1, 4)
    assert m.Rinf.tolist() == [0.99, 0.2981064553812348, 0.0]
    assert np.all(m.A == 0) and np.all(m.B[0] == 0) and np.all(m.B[2] == 0)
    assert m.alpha[1, 3] == 283.3825205916775 and m.beta[1, 3] == 387.53098010301716


def test_fit_matches_reference_reader(golden):
    from oracle import oracle as O
    if not O.ref_available():
        pytest.skip("reference checker not built here")
    fits = sorted((golden / "bras_rs1" / "input" / "fit").glob("*.fit"))
    K, L = read_poles(str(golden / "bras_rs1" / "config" / "poles.txt"))
    m = read_fit_files([str(f) for f in fits], K, L)
    for i, f in enumerate(fits):
        R, A, Z, B, Cc, al, be = O.ref_read_fit(str(f), i, K, L, len(fits))
        # the reference parses into float32: our FP64 values must round to the same floats
        assert np.float32(m.Rinf[i]) == R[i]
        for ours, ref, n in ((m.A, A, K), (m.zeta, Z, K), (m.B, B, L), (m.C, Cc, L), (m.alpha, al, L), (m.beta, be, L)):
            assert np.array_equal(ours[i].astype(np.float32), ref[i * n:(i + 1) * n])


def test_fit_malformed(tmp_path):
    p = tmp_path / "bad.fit"
    p.write_text("# R_inf\n0.5\n# K L\n2 1\n# A_k zeta_k\n1.0 2.0\n")
    with pytest.raises(FasesError) as e:
        read_fit_files([str(p)], 2, 1)
    assert e.value.code == _lib.FASES_E_PARSE
    p.write_text("# R_inf\n0.5\n# K L\n3 0\n")
    with pytest.raises(FasesError) as e:
        read_fit_files([str(p)], 2, 1)   # K_local > K_m
    assert e.value.code == _lib.FASES_E_PARSE
    with pytest.raises(FasesError) as e:
        read_fit_files([str(tmp_path / "missing.fit")], 1, 1)
    assert e.value.code == _lib.FASES_E_IO


def test_sources_and_receivers(tmp_path, golden):
    h = np.float32(0.045)
    dims = (250, 140, 120)
    cell, lp, f0, xyz = read_sources(str(golden / "bras_rs1" / "input" / "src" / "src.txt"), h, dims, 350.0)
    i, j, k = [int(np.float32(x) / h) + 3 for x in (3.379, 2.985, 2.141)]
    assert cell.tolist() == [i + dims[0] * (j + dims[1] * k)]
    assert lp[0] == np.float32(86.02) and f0[0] == 50.0
    p = tmp_path / "src.txt"
    p.write_text("# x y z Lp\n1.0 1.0 1.0 80\n100 1 1 80\n")   # 4 columns: default f0; 2nd outside -> dropped
    cell, lp, f0, _ = read_sources(str(p), h, dims, 350.0)
    assert cell.size == 1 and f0[0] == 350.0
    r = tmp_path / "rec.txt"
    r.write_text("1 1 1\n2 2 2\n")
    rc, _ = read_receivers(str(r), h, dims)
    assert rc.size == 2
    r.write_text("1 1 1\n100 1 1\n")
    with pytest.raises(FasesError) as e:
        read_receivers(str(r), h, dims)
    assert e.value.code == _lib.FASES_E_RANGE


def test_ricker_waveform_closed_form():
    dt = 2.9154e-05
    Nt = 400
    f = 2000.0
    u = source_waveform(_lib.FASES_SRC_RICKER, dt, f, 4000.0, Nt)
    dth = 0.5 * float(np.float32(dt))
    sig = 1.0 / (np.sqrt(2.0) * np.pi * f)
    lag = np.ceil(5.0 * sig / dth)
    x = (np.arange(2 * Nt + 1) - lag) * dth / sig
    np.testing.assert_allclose(u, x * np.exp(-0.5 * x * x), rtol=0, atol=1e-15)
    # radiated pressure ~ du/dt is a Ricker wavelet: its spectrum peaks at f
    pr, peak = source_pref(u, Nt, dt, 0.02, 1.2)
    spec = np.abs(np.fft.rfft(pr, 1 << 16))
    fr = np.fft.rfftfreq(1 << 16, dth)
    assert abs(fr[np.argmax(spec)] - f) < 0.02 * f
    assert peak == pytest.approx(np.abs(pr).max())


def test_pcs_waveform_is_bilinear_msd_filter():
    """PCS surface velocity = Gaussian force through the bilinear-transformed (prewarped at
    f0) mass-spring-damper V/F = s/(M s^2 + R s + K): checked with scipy's bilinear+lfilter."""
    dt, Nt, f0, fc = 5e-05, 600, 25.0, 1414.2
    u = source_waveform(_lib.FASES_SRC_PCS, dt, f0, fc, Nt)
    dth = 0.5 * float(np.float32(dt))
    w0 = 2 * np.pi * f0
    beta = w0 / np.tan(0.5 * w0 * dth)
    M = 0.025
    R, K = w0 * M, M * w0 * w0
    b, a = scipy.signal.bilinear([1.0, 0.0], [M, R, K], fs=beta / 2.0)
    sig = 0.425 / float(np.float32(fc))
    lag = int(np.ceil(8.5 * sig / dth))
    t = np.arange(2 * Nt + 1)
    F = np.exp(-((t - lag) * dth) ** 2 / (2 * sig * sig))
    ref = scipy.signal.lfilter(b, a, F)
    np.testing.assert_allclose(u, ref, rtol=0, atol=1e-9 * np.abs(ref).max())


def test_one_pascal_scaling():
    """Lp,1m = 20 log10(1/2e-5) dB maps to a 1 Pa peak of the 1 m reference pressure."""
    dt, Nt, h, rho = 2.9154e-05, 300, 0.02, 1.2
    u = source_waveform(_lib.FASES_SRC_RICKER, dt, 2000.0, 4000.0, Nt)
    pr, peak = source_pref(u, Nt, dt, h, rho)
    sval = 2e-5 * 10 ** ((20 * np.log10(1 / 2e-5)) / 20) / peak
    assert np.abs(pr * sval).max() == pytest.approx(1.0, rel=1e-12)


def test_write_f32_roundtrip(tmp_path):
    x = np.random.default_rng(0).standard_normal(1001)
    p = tmp_path / "IR0_0.bin"
    write_f32(str(p), x)
    assert p.stat().st_size == 4 * 1001
    y = np.fromfile(p, "<f4")
    assert np.array_equal(y, x.astype(np.float32))
