"""GPU scene classification (fases_scene_*_dev, fvm_geometry.hip) against the host
classification, bit for bit: signed distance S and material id oid.

The host classification is itself pinned to the reference's geometry code
(tests/test_geometry.py: committed fixtures and the live reference build), so equality
here carries that parity over. Scenes cover what the kernels branch on: axis-aligned
walls on and off cell faces, non-axis-aligned triangles (every Voronoi region of the
point-triangle distance, generic ray crossings), a non-convex room (several crossings
per +X ray), more triangles than one LDS chunk (1024) and coplanar tessellations whose
equal nearest distances must resolve to the first triangle, as on the host."""
import numpy as np
import pytest

from fases_amd.solver import Scene

from tests.golden.make_geometry_golden import write_box

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _same(host, gpu):
    assert gpu.dims == host.dims
    bad = np.flatnonzero(_bits(gpu.S) != _bits(host.S))
    assert bad.size == 0, "S differs at %d cells, first %s: gpu %r host %r" % (
        bad.size, bad[:3], gpu.S[bad[:3]], host.S[bad[:3]])
    bad = np.flatnonzero(gpu.oid != host.oid)
    assert bad.size == 0, "oid differs at %d cells, first %s: gpu %r host %r" % (
        bad.size, bad[:3], gpu.oid[bad[:3]], host.oid[bad[:3]])


def _write_obj(path, V, faces):
    """faces: list of (group name, [(i, j, k), ...]) with 0-based vertex indices."""
    with open(path, "w") as f:
        f.write("o scene\n")
        for v in V:
            f.write("v %.9g %.9g %.9g\n" % tuple(float(a) for a in v))
        for name, tris in faces:
            f.write("g %s\n" % name)
            for t in tris:
                f.write("f %d %d %d\n" % tuple(i + 1 for i in t))


def _icosphere(level):
    t = (1 + 5 ** 0.5) / 2
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    V = [np.array(v, float) / np.linalg.norm(v) for v in V]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        mid = {}

        def m(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                p = V[a] + V[b]
                V.append(p / np.linalg.norm(p))
                mid[key] = len(V) - 1
            return mid[key]
        F = [f for (a, b, c) in F for f in
             ((a, m(a, b), m(c, a)), (b, m(b, c), m(a, b)), (c, m(c, a), m(b, c)), (m(a, b), m(b, c), m(c, a)))]
    return np.array(V), F


def test_shoeboxes_match_host():
    for (nx, ny, nz, h) in [(30, 20, 25, 0.034), (17, 13, 11, 0.05), (64, 48, 40, 0.05), (41, 23, 19, 0.0275),
                            (5, 4, 3, 0.1)]:
        print("[geo] shoebox", nx, ny, nz, h, flush=True)
        _same(Scene.shoebox(nx, ny, nz, np.float32(h)), Scene.shoebox(nx, ny, nz, np.float32(h), device=0))


@pytest.mark.parametrize("src,h", [("scenes/room_modes_shoebox.obj", 0.2), ("scenes/default_project_room.obj", 0.15),
                                   ("scenes/room_modes_shoebox.obj", 0.07)])
def test_reference_scenes_match_host_and_fixture(golden, src, h):
    print("[geo] scene", src, h, flush=True)
    p = str(golden / src)
    host = Scene.from_obj(p, np.float32(h))
    gpu = Scene.from_obj(p, np.float32(h), device=0)
    _same(host, gpu)
    if h in (0.2, 0.15):
        g = np.load(golden / "geometry_ref.npz")
        name = "room_modes_h0.2" if h == 0.2 else "default_room_h0.15"
        assert np.array_equal(_bits(gpu.S), _bits(g[name + "/S"])) and np.array_equal(gpu.oid, g[name + "/oid"])


def test_box_obj_matches_shoebox_on_gpu(tmp_path):
    p = tmp_path / "box.obj"
    write_box(p, 30, 20, 25, 0.034)
    _same(Scene.shoebox(30, 20, 25, np.float32(0.034), device=0), Scene.from_obj(str(p), np.float32(0.034), device=0))


def test_tilted_icosphere(tmp_path):
    """320 generic triangles, 8 groups (one per octant of the face centroid), the
    sphere rotated off the axes and shifted so no vertex sits on a cell centre."""
    V, F = _icosphere(2)
    a, b = 0.37, 0.61
    Rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    Rx = np.array([[1, 0, 0], [0, np.cos(b), -np.sin(b)], [0, np.sin(b), np.cos(b)]])
    V = (V @ (Rx @ Rz).T) * 0.45 + np.array([0.013, -0.007, 0.021])
    groups = {}
    for f in F:
        c = V[list(f)].mean(0)
        groups.setdefault("oct%d" % (4 * (c[0] > 0) + 2 * (c[1] > 0) + (c[2] > 0)), []).append(f)
    p = tmp_path / "sphere.obj"
    _write_obj(p, V, sorted(groups.items()))
    for h in (0.03, 0.0437):
        print("[geo] icosphere h", h, flush=True)
        host = Scene.from_obj(str(p), np.float32(h))
        gpu = Scene.from_obj(str(p), np.float32(h), device=0)
        _same(host, gpu)
        # the classification found inside cells and several materials
        assert (host.S > 0).sum() > 1000 and len(set(np.unique(host.oid)) - {0, 255}) == 8


def test_room_with_pillar(tmp_path):
    """Non-convex: an inward-facing pillar inside the room, so +X rows through it cross
    four walls; the pillar is its own material."""
    X, Y, Z = 1.3, 0.9, 0.7
    V = [(0, 0, 0), (X, 0, 0), (X, Y, 0), (0, Y, 0), (0, 0, Z), (X, 0, Z), (X, Y, Z), (0, Y, Z)]
    x0, x1, y0, y1 = 0.52, 0.81, 0.33, 0.57
    V += [(x0, y0, 0), (x1, y0, 0), (x1, y1, 0), (x0, y1, 0), (x0, y0, Z), (x1, y0, Z), (x1, y1, Z), (x0, y1, Z)]
    quads = {"walls": [(3, 7, 4, 0), (1, 5, 6, 2), (0, 4, 5, 1), (2, 6, 7, 3)], "floor": [(0, 1, 2, 3)],
             "ceiling": [(4, 7, 6, 5)],
             "pillar": [(8, 12, 15, 11), (10, 14, 13, 9), (9, 13, 12, 8), (11, 15, 14, 10)]}
    faces = [(g, [t for (a, b, c, d) in q for t in ((a, b, c), (a, c, d))]) for g, q in quads.items()]
    p = tmp_path / "pillar.obj"
    _write_obj(p, V, faces)
    for h in (0.02, 0.0311):
        print("[geo] pillar h", h, flush=True)
        host = Scene.from_obj(str(p), np.float32(h))
        _same(host, Scene.from_obj(str(p), np.float32(h), device=0))
        assert 4 in set(np.unique(host.oid))


def test_tessellated_room_chunks_and_ties(tmp_path):
    """Every wall split into 11 x 11 quads (1452 triangles, > 1024): coplanar triangles
    give equal nearest distances along shared edges, which must go to the first
    triangle, and the band and nearest-triangle passes run over several chunks."""
    from fases_amd.workloads import write_tessellated_room
    p = tmp_path / "tess.obj"
    assert write_tessellated_room(p, 0.8, 0.6, 0.5, 11) == 1452
    print("[geo] tessellated room", flush=True)
    host = Scene.from_obj(str(p), np.float32(0.02))
    _same(host, Scene.from_obj(str(p), np.float32(0.02), device=0))


def test_bad_device_index_is_an_error():
    from fases_amd._lib import FasesError
    with pytest.raises(FasesError):
        Scene.shoebox(8, 8, 8, np.float32(0.1), device=99)


def test_ray_crossing_more_than_the_list_holds_is_an_error(tmp_path):
    """4100 stacked planes across x (two triangles each, a ray meets one of them): every
    +X ray through them crosses more than the 4096 abscissae the sign pass holds per
    row, and the call returns an error (the kernel drops the extra crossings and
    reports the count; nothing is written past the list)."""
    from fases_amd._lib import FasesError
    V, tris = [], []
    for q in range(4100):
        x = 0.01 + q * 0.0002
        b = len(V)
        V += [(x, 0, 0), (x, 0.3, 0), (x, 0.3, 0.3), (x, 0, 0.3)]
        tris += [(b, b + 1, b + 2), (b, b + 2, b + 3)]
    p = tmp_path / "planes.obj"
    _write_obj(p, V, [("planes", tris)])
    with pytest.raises(FasesError, match="crosses"):
        Scene.from_obj(str(p), np.float32(0.05), device=0)
