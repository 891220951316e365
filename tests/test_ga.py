"""Geometrical acoustics (fases_ga_*): image sources and ray tracer against the reference's
own ism_solve and rt_trace_late (3D-solver/include/ism3D.h, raytrace3D.h), compiled from
the reference sources into oracle/_ref, and the band-IR / arrival synthesis against its
Python restatement (oracle/ga_oracle.py, main3D.cu:1600-1801).

Image sources are bit-exact (host and GPU validity alike): same reflectors, image tree,
path validity, lengths, directions, incidence angles and curvature gains, in the same
order. Scenes: axis-aligned shoebox, the reference's default project room, a room with a
pillar (occlusion), a room with a tessellated sphere (proxy patches with curvature) and
the BRAS RS1 room. The ray tracer adds its energies with float atomics in another order
than the reference's loop, and its device exp / sin / cos differ from the host's by an
ulp, so its echograms are compared by energy per band and receiver and over short windows.
A closed form also pins the first-order images of a shoebox on its own."""
import math
import sys
from pathlib import Path

import numpy as np
import pytest

from fases_amd import ga as GA
from fases_amd.solver import Scene
from oracle import ga_oracle as GO
from oracle import oracle as O

ROOT = Path(__file__).resolve().parent.parent
SCENES = ROOT / "tests" / "golden" / "scenes"

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference checker not built (oracle/_ref)")


def announce(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def write_obj(path, V, faces):
    with open(path, "w") as f:
        f.write("o scene\n")
        for v in V:
            f.write("v %.9g %.9g %.9g\n" % tuple(float(a) for a in v))
        for name, polys in faces:
            f.write("g %s\n" % name)
            for p in polys:
                f.write("f " + " ".join(str(i + 1) for i in p) + "\n")


def box(L):
    X, Y, Z = L
    V = [(0, 0, 0), (X, 0, 0), (X, Y, 0), (0, Y, 0), (0, 0, Z), (X, 0, Z), (X, Y, Z), (0, Y, Z)]
    F = [("west", [(3, 7, 4, 0)]), ("east", [(1, 5, 6, 2)]), ("south", [(0, 4, 5, 1)]), ("north", [(2, 6, 7, 3)]),
         ("floor", [(0, 1, 2, 3)]), ("ceiling", [(4, 7, 6, 5)])]
    return V, F


def icosphere(level):
    t = (1 + 5 ** 0.5) / 2
    V = [np.array(v, float) for v in ((-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
                                      (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1))]
    V = [v / np.linalg.norm(v) for v in V]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7),
         (9, 8, 1)]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = V[a] + V[b]
                V.append(m / np.linalg.norm(m))
                cache[k] = len(V) - 1
            return cache[k]
        F2 = []
        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            F2 += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = F2
    return V, F


def scene_file(tmp, name):
    """(path, h, sources, receivers) of a test scene."""
    if name == "shoebox":
        return str(SCENES / "room_modes_shoebox.obj"), 0.05, [(1.0, 1.2, 1.1)], [(4.0, 2.5, 1.6), (2.2, 3.1, 0.9)]
    if name == "default_room":
        return str(SCENES / "default_project_room.obj"), 0.05, [(1.0, 1.0, 1.2)], [(2.5, 2.0, 1.3), (1.7, 2.9, 1.0)]
    if name == "pillar":
        V, F = box((6.0, 4.5, 3.2))
        x0, x1, y0, y1 = 2.6, 3.4, 1.8, 2.6
        n = len(V)
        V += [(x0, y0, 0), (x1, y0, 0), (x1, y1, 0), (x0, y1, 0), (x0, y0, 3.2), (x1, y0, 3.2), (x1, y1, 3.2),
              (x0, y1, 3.2)]
        F.append(("pillar", [(n + 0, n + 4, n + 7, n + 3), (n + 2, n + 6, n + 5, n + 1), (n + 1, n + 5, n + 4, n + 0),
                             (n + 3, n + 7, n + 6, n + 2)]))
        p = tmp / "pillar.obj"
        write_obj(p, V, F)
        return str(p), 0.05, [(1.0, 2.2, 1.5)], [(5.2, 2.1, 1.4), (3.0, 0.7, 2.0)]
    if name == "sphere_room":
        V, F = box((6.0, 5.0, 4.0))
        Vs, Fs = icosphere(2)
        n = len(V)
        V += [tuple(np.array((3.2, 2.4, 1.9)) + 0.8 * v) for v in Vs]
        F.append(("ball", [(a + n, b + n, c + n) for a, b, c in Fs]))
        p = tmp / "sphere_room.obj"
        write_obj(p, V, F)
        return str(p), 0.05, [(1.0, 1.0, 1.2), (5.1, 4.0, 3.0)], [(2.0, 4.0, 1.5), (5.0, 1.0, 2.5)]
    if name == "bras_rs1":
        from tests.test_gpu_bras_rs1 import rs1_obj, sketch_box
        p = tmp / "rs1.obj"
        p.write_text(rs1_obj(*sketch_box()))
        return str(p), 0.045, [(3.379, 2.985, 2.141)], [(7.621, 2.985, 2.141)]
    raise ValueError(name)


ISM_CASES = [("shoebox", 3), ("default_room", 4), ("pillar", 3), ("sphere_room", 3), ("bras_rs1", 3)]


def ism_equal(g, ref, ns, nr):
    used, direct, imgs = ref
    assert g.info()["order_used"] == used
    n = 0
    for s in range(ns):
        for r in range(nr):
            assert g.direct(s, r) == direct[s, r], (s, r)
            mine, theirs = g.images(s, r), imgs[s][r]
            assert len(mine) == len(theirs), (s, r, len(mine), len(theirs))
            for a, b in zip(mine, theirs):
                assert a.distance == b[0]
                assert np.array_equal(a.doa, b[1])
                assert a.amp == b[2]
                assert np.array_equal(a.group, b[3])
                assert np.array_equal(a.cos_inc, b[4])
                assert np.array_equal(a.angle, b[5])
            n += len(mine)
    return n


@needs_ref
@pytest.mark.parametrize("name,order", ISM_CASES)
def test_image_sources_on_the_host_equal_the_reference(tmp_path, name, order):
    path, h, src, rec = scene_file(tmp_path, name)
    sc = Scene.from_obj(path, np.float32(h))
    g = GA.run(sc, src, rec, GA.GAParams(device=-1, ism_order=order, n_rays=0, Nt=10))
    n = ism_equal(g, O.ref_ism(path, np.float32(h), src, rec, order), len(src), len(rec))
    assert n > 10 * len(src) * len(rec)
    if name == "sphere_room":        # proxy patches of the ball give curved-mirror gains
        amps = [im.amp for s in range(len(src)) for r in range(len(rec)) for im in g.images(s, r)]
        assert any(a != 1.0 for a in amps)


def test_first_order_images_of_a_shoebox_closed_form(tmp_path):
    """Source and receiver in a box: the direct distance and the six first-order image
    distances are |r - mirror_w(s)| for each wall w; the incidence cosine of the wall
    reflection is the normal distance sum over the path length."""
    L = (5.0, 4.0, 3.0)
    V, F = box(L)
    p = tmp_path / "box.obj"
    write_obj(p, V, F)
    s, r = np.array((1.3, 1.1, 0.9)), np.array((3.6, 2.7, 2.2))
    sc = Scene.from_obj(str(p), np.float32(0.05))
    g = GA.run(sc, [s], [r], GA.GAParams(device=-1, ism_order=1, n_rays=0, Nt=10))
    assert g.direct(0, 0) == pytest.approx(np.linalg.norm(r - s), rel=1e-6)
    imgs = g.images(0, 0)
    assert len(imgs) == 6 and all(len(im.group) == 1 for im in imgs)
    names = sc.group_names
    for im in imgs:
        wall = names[int(im.group[0])]
        axis, at = {"west": (0, 0.0), "east": (0, L[0]), "south": (1, 0.0), "north": (1, L[1]), "floor": (2, 0.0),
                    "ceiling": (2, L[2])}[wall]
        m = s.copy()
        m[axis] = 2 * at - s[axis]
        d = np.linalg.norm(r - m)
        assert im.distance == pytest.approx(d, rel=2e-6), wall
        cos_expected = (abs(s[axis] - at) + abs(r[axis] - at)) / d
        assert im.cos_inc[0] == pytest.approx(cos_expected, rel=1e-5), wall
        assert im.amp == 1.0
    assert sorted(names[int(im.group[0])] for im in imgs) == sorted(names)


def ga_tables(ng, seed=3):
    rng = np.random.default_rng(seed)
    alpha = rng.uniform(0.05, 0.6, (ng, 8)).astype(np.float32)
    sctr = rng.uniform(0.0, 0.4, ng).astype(np.float32)
    m_air = np.array([1e-5, 4e-5, 1.2e-4, 3e-4, 6e-4, 1.1e-3, 2.8e-3, 8.5e-3], np.float32)
    return alpha, sctr, m_air


def images_for_oracle(g, ns, nr):
    return [[[{"distance": im.distance, "amp": im.amp, "group": im.group, "angle": im.angle, "doa": im.doa}
              for im in g.images(s, r)] for r in range(nr)] for s in range(ns)]


def check_band_irs(g, sc, src, rec, src_p, alpha, sctr, m_air, c, fs, Nt, bands, rtol):
    ns, nr = len(src), len(rec)
    direct = np.array([[g.direct(s, r) for r in range(nr)] for s in range(ns)], np.float32)
    imgs = images_for_oracle(g, ns, nr)
    es, ed = g.energy()
    ng = len(sc.group_names)
    for b in bands:
        hf, rt = g.band_irs(b)
        ohf, ort = GO.band_irs(direct, imgs, src_p, es, ed, alpha.ravel(), sctr, ng, m_air, c, fs, Nt, b)
        np.testing.assert_allclose(hf, ohf, rtol=rtol, atol=rtol * np.abs(ohf).max())
        np.testing.assert_allclose(rt, ort, rtol=rtol, atol=rtol * max(np.abs(ort).max(), 1e-30))
        assert np.abs(hf).max() > 0
    for r in range(nr):
        a = g.arrivals(r)
        oa = GO.arrivals(direct, imgs, src_p, src, rec, alpha.ravel(), sctr, ng, m_air, c, fs, Nt, r)
        assert a.shape == oa.shape
        np.testing.assert_allclose(a, oa, rtol=rtol, atol=1e-12)


def test_band_irs_and_arrivals_match_the_synthesis_restatement(tmp_path):
    """Image sources only (no rays): the direct sound and the images as windowed-sinc
    fractional delays with per-band oblique reflection factors, scattering loss and air
    attenuation."""
    path, h, src, rec = scene_file(tmp_path, "default_room")
    sc = Scene.from_obj(path, np.float32(h))
    alpha, sctr, m_air = ga_tables(len(sc.group_names))
    c, fs, Nt = 343.0, 44100.0, 2205
    src_p = np.array([0.7], np.float32)
    g = GA.run(sc, src, rec, GA.GAParams(device=-1, ism_order=3, n_rays=0, c=c, fs=fs, Nt=Nt), src_pressure=src_p,
               alpha=alpha, sctr=sctr, m_air=m_air)
    check_band_irs(g, sc, src, rec, src_p, alpha, sctr, m_air, c, fs, Nt, bands=[0, 4, 7], rtol=2e-6)


# ------------------------------------------------------------------------------------
# GPU
# ------------------------------------------------------------------------------------

@pytest.mark.gpu
@needs_ref
@pytest.mark.parametrize("name,order", ISM_CASES + [("shoebox", 7)])
def test_image_sources_on_the_gpu_equal_the_reference(tmp_path, name, order):
    path, h, src, rec = scene_file(tmp_path, name)
    announce("image sources on the GPU: %s order %d" % (name, order))
    sc = Scene.from_obj(path, np.float32(h))
    g = GA.run(sc, src, rec, GA.GAParams(device=0, ism_order=order, n_rays=0, Nt=10))
    ism_equal(g, O.ref_ism(path, np.float32(h), src, rec, order), len(src), len(rec))


def rt_case(tmp_path, name, scatter):
    path, h, src, rec = scene_file(tmp_path, name)
    sc = Scene.from_obj(path, np.float32(h))
    alpha, sctr, m_air = ga_tables(len(sc.group_names))
    if not scatter:
        sctr = np.zeros_like(sctr)
    c, fs, Nt, n_rays = 343.0, 44100.0, 8820, 20000
    src_p = np.ones(len(src), np.float32)
    return path, h, src, rec, sc, alpha, sctr, m_air, c, fs, Nt, n_rays, src_p


@pytest.mark.gpu
@needs_ref
@pytest.mark.parametrize("name,scatter", [("shoebox", False), ("shoebox", True), ("pillar", True),
                                          ("sphere_room", True)])
def test_ray_tracer_matches_the_reference(tmp_path, name, scatter):
    path, h, src, rec, sc, alpha, sctr, m_air, c, fs, Nt, n_rays, src_p = rt_case(tmp_path, name, scatter)
    announce("ray tracer %s scatter=%s" % (name, scatter))
    g = GA.run(sc, src, rec, GA.GAParams(device=0, ism_order=2, n_rays=n_rays, c=c, fs=fs, Nt=Nt),
               src_pressure=src_p, alpha=alpha, sctr=sctr, m_air=m_air)
    es, ed = g.energy()
    res, red = O.ref_rt(path, np.float32(h), src, src_p, rec, alpha, m_air, sctr, c, fs, Nt, n_rays, 100000, 0.12,
                        g.info()["order_used"])
    info = g.info()
    print("kernel %.1f ms" % info["rt_kernel_ms"])
    for mine, ref, what in ((es, res, "spec"), (ed, red, "diff")):
        tot, rtot = mine.sum(axis=2), ref.sum(axis=2)
        if rtot.max() == 0:
            assert mine.max() == 0, what
            continue
        # energy per band and receiver
        tol = 1e-5 if not scatter else 0.03
        np.testing.assert_allclose(tot, rtot, rtol=tol, err_msg=what)
        # and its time distribution over 10 ms windows carrying at least 1 % of the energy
        w = 441
        nw = Nt // w
        mw = mine[:, :, :nw * w].reshape(8, -1, nw, w).sum(axis=3)
        rw = ref[:, :, :nw * w].reshape(8, -1, nw, w).sum(axis=3)
        big = rw > 0.01 * rtot[:, :, None]
        assert np.allclose(mw[big], rw[big], rtol=(1e-4 if not scatter else 0.15)), what
    if not scatter:
        assert ed.max() == 0 and red.max() == 0
        # specular: every detection is the reference's, up to the float summation order
        np.testing.assert_allclose(es, res, rtol=1e-4, atol=1e-6 * es.max())


@pytest.mark.gpu
def test_band_irs_with_rays_match_the_synthesis_restatement(tmp_path):
    path, h, src, rec, sc, alpha, sctr, m_air, c, fs, Nt, n_rays, src_p = rt_case(tmp_path, "shoebox", True)
    announce("band IRs with rays")
    g = GA.run(sc, src, rec, GA.GAParams(device=0, ism_order=2, n_rays=n_rays, c=c, fs=fs, Nt=Nt),
               src_pressure=src_p, alpha=alpha, sctr=sctr, m_air=m_air)
    check_band_irs(g, sc, src, rec, src_p, alpha, sctr, m_air, c, fs, Nt, bands=[2, 6], rtol=2e-6)


# ------------------------------------------------------------------------------------
# Pins of the IR-synthesis restatement (oracle/ga_oracle.py) on its own
# ------------------------------------------------------------------------------------

@needs_ref
def test_synthesis_helpers_equal_the_reference_build():
    """R_oblique (parsing3D.h:1178) and rt_sign (raytrace3D.h:69) of the restatement
    against the reference's own functions built from its sources, bit for bit."""
    L = O.ref_lib()
    rng = np.random.default_rng(9)
    al = np.concatenate([rng.uniform(0, 1, 400), [0.0, 0.99, 0.9900001, 1.0, 1e-6, 2e-5]]).astype(np.float32)
    cs = np.concatenate([rng.uniform(-1, 1, 400), [1.0, 0.0, 0.5, -1.0, 1e-3, 0.999]]).astype(np.float32)
    for a, c in zip(al, cs):
        assert np.float32(GO.r_oblique(a, c)).view(np.uint32) == np.float32(L.ref_r_oblique(a, c)).view(np.uint32), (a, c)
    for s in list(rng.integers(0, 2 ** 32, 2000, dtype=np.uint64)) + [0, 1, 2 ** 32 - 1, 2654435761]:
        assert GO.rt_sign(int(s)) == L.ref_rt_sign(int(s)), s


def test_oblique_reflection_closed_forms():
    """A locally reacting surface matched to the random-incidence alpha reflects
    sqrt(1 - alpha) at normal incidence and -1 at grazing incidence; alpha > 0.99 is
    anechoic and alpha -> 0 rigid."""
    for a in (0.02, 0.1, 0.3, 0.55, 0.8, 0.95):
        assert GO.r_oblique(a, 1.0) == pytest.approx(np.sqrt(1 - a), rel=2e-6)
        assert GO.r_oblique(a, 0.0) == -1.0
        # R rises monotonically from -1 (grazing) to sqrt(1 - a) (normal)
        r = [float(GO.r_oblique(a, c)) for c in np.linspace(0, 1, 21)]
        assert all(x < y for x, y in zip(r, r[1:]))
    assert GO.r_oblique(0.995, 0.7) == 0.0
    assert GO.r_oblique(0.0, 0.3) == 1.0


def test_fractional_delay_deposit_closed_forms():
    """The windowed-sinc deposit of one arrival: at an integer sample delay only that
    sample receives the gain (sinc zeros at the other integers, up to the rounding of
    sinf at float multiples of pi); at a half-sample delay
    the taps are symmetric about it and sum to the gain within the window's DC error;
    direct sound of a source at distance d arrives with gain p / d exp(-m d)."""
    Nt = 200
    row = np.zeros(Nt, np.float32)
    GO._deposit(row, np.float32(50.0), np.float32(0.7), Nt)
    assert row[50] == np.float32(0.7) and np.abs(np.delete(row, 50)).max() < 1e-6 * 0.7
    row = np.zeros(Nt, np.float32)
    GO._deposit(row, np.float32(80.5), np.float32(1.0), Nt)
    # taps 64 .. 96 around int(80.5) = 80: n and 161 - n mirror each other about 80.5
    np.testing.assert_allclose(row[65:81], row[96:80:-1], rtol=1e-5, atol=1e-9)
    assert abs(float(row.sum()) - 1.0) < 0.02 and np.argmax(row) in (80, 81)
    # direct sound only: c = fs so the delay in samples is the distance in metres
    c = fs = 1000.0
    d = 37.0
    direct = np.array([[d]], np.float32)
    m_air = np.full(8, 1e-3, np.float32)
    zero = np.zeros((8, 1, Nt), np.float32)
    hf, rt = GO.band_irs(direct, [[[]]], np.array([2.0], np.float32), zero, zero, np.zeros(8, np.float32),
                         np.zeros(1, np.float32), 1, m_air, c, fs, Nt, 3)
    g = 2.0 / d * np.exp(-1e-3 * d)
    assert hf[0, 37] == pytest.approx(g, rel=1e-6) and np.abs(np.delete(hf[0], 37)).max() < 1e-6 * g
    assert not rt.any()
    # one normally incident first-order image: gain p amp sqrt(1 - alpha) sqrt(1 - s) / d exp(-m d)
    im = {"distance": 61.0, "amp": 1.0, "group": np.array([0]), "angle": np.array([0.0], np.float32),
          "doa": np.zeros(3, np.float32)}
    alpha = np.full(8, 0.36, np.float32)
    hf, _ = GO.band_irs(np.array([[0.0]], np.float32), [[[im]]], np.array([1.0], np.float32), zero, zero, alpha,
                        np.array([0.19], np.float32), 1, m_air, c, fs, Nt, 5)
    g = 0.8 * 0.9 / 61.0 * np.exp(-1e-3 * 61.0)
    assert hf[0, 61] == pytest.approx(g, rel=2e-6) and np.abs(np.delete(hf[0], 61)).max() < 1e-6 * g
    # the ray-traced energy enters as 4 pi sign(t) sqrt(e) at its sample
    es = zero.copy()
    es[5, 0, 120] = 0.25
    hf, rt = GO.band_irs(np.array([[0.0]], np.float32), [[[]]], np.array([1.0], np.float32), es, zero, alpha,
                         np.zeros(1, np.float32), 1, m_air, c, fs, Nt, 5)
    assert abs(rt[0, 120]) == pytest.approx(4 * np.pi * 0.5, rel=1e-6) and np.count_nonzero(rt) == 1
    assert np.array_equal(hf, rt)
