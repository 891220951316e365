"""Writes tests/golden/broadband_ref.npz: inputs and outputs of the reference's broadband
IR post-processing (common/allIRs.py: low_deconv, _filt, broadband_ir) on synthetic
projects, for tests/test_broadband.py. Run where the reference checkout exists:

    python tests/golden/make_broadband_golden.py

The Tikhonov deconvolution after polyphase resampling and the band merge have no closed
form to pin them, so the vectors come from the reference module itself (imported, its
functions called on files this script writes). Three projects: a full run (receiver IRs,
radiated pulse p_ref.bin, geometrical octave bands, one pair with bands only), the same
run without p_ref.bin (the delay fallback) and a bands-only run (wave band below 63 Hz).
"""
import os
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF = Path(os.environ.get("FASES_REFERENCE", "/root/reference")) / "common"

H, F_IN, NT = 0.02, 500.0, 600
DT = float(np.float32(np.float32(0.5) * np.float32(H)) / np.float32(343.0))
BANDS = (250, 500, 1000, 2000, 4000, 8000)


def ricker_pref(nt, dt, f0=400.0):
    """Pulse at 1 m sampled at dt/2 (2 nt + 1 samples), unit peak."""
    t = np.arange(2 * nt + 1) * (dt / 2)
    s = 1.0 / (np.sqrt(2) * np.pi * f0)
    ts = np.ceil(5 * s / (dt / 2)) * (dt / 2)
    x = (t - ts) / s
    q = -x * np.exp(-x * x / 2)
    return q / np.abs(q).max()


def write_project(root: Path, cfg, header: bool, pref, files: dict):
    (root / "config").mkdir(parents=True)
    (root / "output").mkdir()
    (root / "config" / "config.txt").write_text("".join("%r\n" % float(v) for v in cfg))
    if header:
        (root / "output" / "header.txt").write_text(
            "%d %d %d %d %f %f %f %d %d %f %s %f %d\n" % (46, 36, 31, NT, 0.92, 0.72, 0.62, 0, 5, DT, "scene.obj", H, 3))
    if pref is not None:
        pref.astype(np.float64).tofile(root / "config" / "p_ref.bin")
    for name, arr in files.items():
        arr.tofile(root / "output" / name)


def main():
    sys.path.insert(0, str(REF))
    import allIRs as A                                   # noqa: E402  (the reference module)

    rng = np.random.default_rng(2024)
    pref = ricker_pref(NT, DT)
    pulse = pref[::2]                                    # at the FVM rate
    files = {}
    for rec in (0, 1):
        room = np.zeros(NT)
        taps = rng.integers(40, NT - 100, 12)
        room[taps] = rng.standard_normal(12) * np.exp(-taps / 300.0)
        room += 1e-3 * rng.standard_normal(NT) * np.exp(-np.arange(NT) / 250.0)
        files["IR%d_0.bin" % rec] = np.convolve(room, pulse)[:NT].astype(np.float32)
        for b in BANDS:
            files["HIR%d_0_%d.bin" % (rec, b)] = (rng.standard_normal(1200) * np.exp(-np.arange(1200) / 300.0)
                                                  ).astype(np.float32)
    for b in (1000, 2000):
        files["HIR2_0_%d.bin" % b] = (rng.standard_normal(1000) * np.exp(-np.arange(1000) / 250.0)).astype(np.float32)
    cfg = [H, NT * DT, F_IN, 20.0, 0.5, 12, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 400.0, 0]
    out = {"cfg": np.array(cfg), "dt_header": np.array(float("%f" % DT)), "pref": pref}
    for k, v in files.items():
        out["in/" + k] = v
    cwd = os.getcwd()
    try:
        for tag, with_pref in (("full", True), ("nopref", False)):
            d = Path(tempfile.mkdtemp())
            write_project(d, cfg, True, pref if with_pref else None, files)
            os.chdir(d / "output")
            A._ready = False
            A._init()
            out[tag + "/lag"] = np.array(A.lag)
            if with_pref:
                out[tag + "/pulse_fs"] = A.pulse_fs
                out[tag + "/low_deconv_IR0"] = A.low_deconv(files["IR0_0.bin"])
            for rec in (0, 1, 2) if with_pref else (0,):
                bb, lo, hi = A.broadband_ir(rec, 0)
                out["%s/bb%d" % (tag, rec)], out["%s/lo%d" % (tag, rec)], out["%s/hi%d" % (tag, rec)] = bb, lo, hi
            os.chdir(cwd)
        # bands only: wave band below 63 Hz, no header.txt (the FVM did not run)
        d = Path(tempfile.mkdtemp())
        cfg_ga = list(cfg)
        cfg_ga[2] = 50.0
        ga = {k: v for k, v in files.items() if k.startswith("HIR0_0_")}
        write_project(d, cfg_ga, False, None, ga)
        os.chdir(d / "output")
        A._ready = False
        bb, lo, hi = A.broadband_ir(0, 0)
        out["gaonly/bb0"], out["gaonly/lo0"], out["gaonly/hi0"] = bb, lo, hi
        out["cfg_ga"] = np.array(cfg_ga)
        os.chdir(cwd)
        # the filter alone, on every edge combination broadband_ir uses
        x = rng.standard_normal(777)
        out["filt/x"] = x
        for i, (hp, lp, hpp, lpp) in enumerate([(None, 707.1067811865476, False, True), (707.1067811865476, 1414.2, True, False),
                                                (1414.2, 2828.4, False, False), (5656.85, None, False, False),
                                                (None, None, False, False)]):
            out["filt/y%d" % i] = A._filt(x, hp, lp, hp_pow=hpp, lp_pow=lpp)
    finally:
        os.chdir(cwd)
    np.savez_compressed(HERE / "broadband_ref.npz", **out)
    print("wrote", HERE / "broadband_ref.npz", len(out), "arrays")


if __name__ == "__main__":
    main()
