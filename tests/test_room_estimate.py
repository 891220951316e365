"""Reverberation estimate of the preflight (T60.bin, f_Schroeder.bin, A_tot.bin): mesh
group areas, .mat absorption loading and the Eyring estimate against the reference's own
host code (oracle/_ref: parsing3D.h mesh_group_areas, ism3D.h ism_load_impedances_layered
and estimate_T60s), bit for bit; the ISO 9613-1 air attenuation against the standard's
tabulated value and its limits; the CLI's --estimate files against the library calls."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from fases_amd._lib import load
from fases_amd.solver import Scene

from tests.golden.make_geometry_golden import write_box

OCT = np.array([63, 125, 250, 500, 1000, 2000, 4000, 8000], np.float32)


def _f(a):
    return np.ascontiguousarray(a, np.float32).ctypes.data_as(C.POINTER(C.c_float))


def areas(sc):
    A = np.zeros(max(len(sc.group_names), 1), np.float32)
    tot = C.c_float()
    assert load().fases_scene_group_areas(sc._h, _f(A), C.byref(tot)) == 0
    return A, np.float32(tot.value)


def air(T, alt=0.0, hm=1.1):
    m = np.zeros(8, np.float32)
    assert load().fases_air_attenuation(T, alt, hm, _f(OCT), 8, _f(m)) == 0
    return m


def t60(V, c, A_tot, A, alpha, m_air):
    T = np.zeros(8, np.float32)
    fs = C.c_float()
    A = np.ascontiguousarray(A, np.float32)
    alpha = np.ascontiguousarray(alpha, np.float32)
    m_air = np.ascontiguousarray(m_air, np.float32)
    assert load().fases_estimate_t60(V, c, A_tot, A.size, _f(A), _f(alpha), _f(m_air), _f(T), C.byref(fs)) == 0
    return T, np.float32(fs.value)


def test_box_areas_closed_form(tmp_path):
    p = tmp_path / "box.obj"
    write_box(p, 30, 20, 25, 0.034)
    A, tot = areas(Scene.from_obj(str(p), np.float32(0.034)))
    X, Y, Z = [np.float32(n) * np.float32(0.034) for n in (30, 20, 25)]
    want = np.array([Y * Z, Y * Z, X * Z, X * Z, X * Y, X * Y], np.float64)
    np.testing.assert_allclose(A, want, rtol=1e-5)
    np.testing.assert_allclose(tot, want.sum(), rtol=1e-5)


def test_air_attenuation_iso9613():
    """ISO 9613-1 at 20 C and 101.325 kPa: about 4.7 dB/km at 1 kHz near 50 % relative
    humidity (1.1 % molar vapour is ~47 % RH at 20 C); attenuation rises with frequency,
    and altitude (lower pressure) changes it."""
    T = np.float32(293.15)
    m = air(T)
    db_km = m.astype(np.float64) / 0.115 * 1000.0      # Np/m (the dB value x 0.115) -> dB/km
    assert 4.0 < db_km[4] < 5.5, db_km
    assert np.all(np.diff(m) > 0)
    assert not np.array_equal(air(T, alt=1500.0), m)


def test_group_areas_match_reference(golden, tmp_path):
    from oracle import oracle as O
    if not O.ref_available():
        pytest.skip("reference checker not built here")
    p = tmp_path / "box.obj"
    write_box(p, 41, 23, 19, 0.0275)
    for path, h in [(str(golden / "scenes/room_modes_shoebox.obj"), 0.2),
                    (str(golden / "scenes/default_project_room.obj"), 0.15), (str(p), 0.0275)]:
        A, tot = areas(Scene.from_obj(path, np.float32(h)))
        Ar, totr = O.ref_group_areas(path, np.float32(h))
        assert np.array_equal(A.view(np.uint32), Ar.view(np.uint32)) and tot == totr


def test_t60_matches_reference(golden, tmp_path):
    """Areas, .mat loading (comments, values spread over lines, a short file, a missing
    file) and the Eyring estimate, with the same air attenuation, bit for bit."""
    from oracle import oracle as O
    if not O.ref_available():
        pytest.skip("reference checker not built here")
    path = str(golden / "scenes/default_project_room.obj")
    h = np.float32(0.15)
    sc = Scene.from_obj(path, h)
    mat = tmp_path / "mat"
    mat.mkdir()
    rng = np.random.default_rng(7)
    names = sc.group_names
    for q, g in enumerate(names):
        if q == 1:
            continue                                   # missing: all zero
        a = rng.uniform(0.02, 0.6, 8)
        if q == 2:
            (mat / (g + ".mat")).write_text("# absorption\n%s\n%s # tail\n" % (
                " ".join("%.4f" % v for v in a[:3]), " ".join("%.4f" % v for v in a[3:])))
        elif q == 3:
            (mat / (g + ".mat")).write_text(" ".join("%.4f" % v for v in a[:5]) + "\n")   # 5 of 8
        else:
            (mat / (g + ".mat")).write_text(" ".join("%.5f" % v for v in a) + "\n")
    alpha = np.zeros((len(names), 8), np.float32)
    for q, g in enumerate(names):
        load().fases_read_mat(str(mat / (g + ".mat")).encode(), _f(alpha[q]), None)
    A, tot = areas(sc)
    V, c = np.float32(53.7), np.float32(343.2)
    m = air(np.float32(295.15))
    T, fs = t60(V, c, tot, A, alpha, m)
    Tr, alr = O.ref_t60(path, h, str(mat), V, c, m)
    assert np.array_equal(alr, alpha)
    assert np.array_equal(T.view(np.uint32), Tr.view(np.uint32)), (T, Tr)
    assert fs == np.float32(2000.0) * np.sqrt(T[3] / V)
    # uniform absorption, no air: Eyring's closed form 55.25 V / (-c S ln(1 - a))
    Tu, _ = t60(V, c, tot, A, np.full((len(names), 8), 0.01, np.float32), np.zeros(8, np.float32))
    np.testing.assert_allclose(Tu, 55.25 / 343.2 * V / (tot * -np.log(1 - 0.01)), rtol=1e-5)


def test_rigid_room_band_does_not_decay():
    T, fs = t60(np.float32(10.0), np.float32(343.0), np.float32(30.0), np.array([30.0], np.float32),
                np.zeros((1, 8), np.float32), air(np.float32(293.15)))
    assert np.all(T == np.float32(1e9)) and fs == 0.0


def test_missing_mat_is_io_error_with_zero_alpha(tmp_path):
    a = np.ones(8, np.float32)
    n = C.c_int(-1)
    assert load().fases_read_mat(str(tmp_path / "none.mat").encode(), _f(a), C.byref(n)) == -2
    assert np.all(a == 0) and n.value == 0


def test_cli_estimate_writes_reverberation_files(tmp_path):
    from tests.test_cli import make_project, run_cli
    paths = make_project(tmp_path)
    root = Path(paths["config"]).parent.parent
    (root / "input" / "mat").mkdir(parents=True, exist_ok=True)
    for q, g in enumerate(("west", "east", "south", "north", "floor", "ceiling")):
        (root / "input" / "mat" / (g + ".mat")).write_text(" ".join(["%.2f" % (0.05 + 0.05 * q)] * 8) + "\n")
    r = run_cli(paths, "--estimate")
    assert r.returncode == 0, r.stderr
    cdir = Path(paths["config"]).parent
    T = np.fromfile(cdir / "T60.bin", np.float32)
    fs = np.fromfile(cdir / "f_Schroeder.bin", np.float32)
    A_tot = np.fromfile(cdir / "A_tot.bin", np.float32)
    V = np.fromfile(cdir / "Vol.bin", np.float32)[0]
    sc = Scene.from_obj(paths["scene"], np.float32(0.02))
    A, tot = areas(sc)
    assert A_tot.size == 1 and A_tot[0] == tot
    alpha = np.array([[0.05 + 0.05 * q] * 8 for q in range(6)], np.float32)
    Temp = np.float32(np.float32(20.0) + np.float32(273.15))
    c = np.sqrt(np.float32(1.4) * np.float32(287.05) * Temp, dtype=np.float32)
    Tw, fw = t60(V, c, tot, A, alpha, air(Temp))
    assert np.array_equal(T, Tw) and fs[0] == fw
    assert "[OK] f_Schroeder" in r.stdout
