"""Scene classification: product (host C++) vs the reference's own geometry code, plus
invariants the geometry fixes (signed distance of an axis-aligned box, band/oid
consistency, brute-force nearest-triangle distance)."""
import numpy as np
import pytest

from fases_amd.solver import Scene, boundary_map

from tests.golden.make_geometry_golden import CASES, write_box  # noqa: E402


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("name,src,box,h", CASES)
def test_matches_reference_fixture(golden, tmp_path, name, src, box, h):
    g = np.load(golden / "geometry_ref.npz")
    if src == "box":
        p = tmp_path / "box.obj"
        write_box(p, *box, h)
        sc = Scene.from_obj(str(p), np.float32(h))
        sb = Scene.shoebox(*box, np.float32(h))
        assert np.array_equal(_bits(sb.S), _bits(sc.S)) and np.array_equal(sb.oid, sc.oid)
    else:
        sc = Scene.from_obj(str(golden / src), np.float32(h))
    assert tuple(g[name + "/dims"]) == sc.dims
    assert np.array_equal(_bits(sc.S), _bits(g[name + "/S"]))
    assert np.array_equal(sc.oid, g[name + "/oid"])
    bmap, nB = boundary_map(sc.oid)
    assert np.array_equal(bmap, g[name + "/bmap"])
    assert nB == int((g[name + "/bmap"] >= 0).sum())


def test_matches_live_reference_build(tmp_path):
    from oracle import oracle as O
    if not O.ref_available():
        pytest.skip("reference checker not built here")
    for (nx, ny, nz, h) in [(64, 48, 40, 0.05), (41, 23, 19, 0.0275), (20, 20, 20, 0.01)]:
        p = tmp_path / "b.obj"
        write_box(p, nx, ny, nz, h)
        dims, S, oid = O.ref_scene(str(p), np.float32(h))
        sc = Scene.shoebox(nx, ny, nz, np.float32(h))
        assert sc.dims == dims
        assert np.array_equal(_bits(sc.S), _bits(S)) and np.array_equal(sc.oid, oid)


def test_box_signed_distance_and_band():
    """Axis-aligned box with walls on cell faces (h a power of two: exact arithmetic):
    S = distance to the nearest wall plane inside the 8h band (negative outside the
    room), clamped at +-8h; oid set exactly on interior cells with 0 <= S <= 3h, with
    the id of the nearest wall's group away from edges."""
    nx, ny, nz, h = 24, 20, 18, 0.0625
    sc = Scene.shoebox(nx, ny, nz, np.float32(h))
    assert sc.dims == (nx + 6, ny + 6, nz + 6)
    Nx, Ny, Nz = sc.dims
    S = sc.S.reshape(Nz, Ny, Nx)
    oid = sc.oid.reshape(Nz, Ny, Nx)
    k, j, i = np.meshgrid(np.arange(Nz), np.arange(Ny), np.arange(Nx), indexing="ij")
    c = [(a + 0.5) * h for a in (i, j, k)]
    lo, hi = 3 * h, [(n + 3) * h for n in (nx, ny, nz)]
    d_in = np.minimum.reduce([c[0] - lo, hi[0] - c[0], c[1] - lo, hi[1] - c[1], c[2] - lo, hi[2] - c[2]])
    inside = d_in > 0
    assert np.array_equal(S > 0, inside)
    band = 8 * h
    sel = inside & (d_in < band - 1e-6)
    np.testing.assert_allclose(S[sel], d_in[sel], rtol=0, atol=1e-6)
    assert np.all(S[inside & (d_in > band)] == np.float32(band))
    tol = np.float32(3 * h)
    core = np.zeros_like(inside)
    core[1:-1, 1:-1, 1:-1] = True
    expect_band = core & (S >= 0) & (S <= tol)
    assert np.array_equal(oid != 0xFF, expect_band)
    # nearest wall group (west,east,south,north,floor,ceiling = 1..6) on the face centres
    dist = np.stack([c[0] - lo, hi[0] - c[0], c[1] - lo, hi[1] - c[1], c[2] - lo, hi[2] - c[2]])
    srt = np.sort(dist, axis=0)
    unique = expect_band & (srt[1] - srt[0] > 0.5 * h)
    assert np.array_equal(oid[unique], (np.argmin(dist, axis=0) + 1)[unique].astype(np.uint8))


def _tri_dist(p, a, b, c):
    """Exact point-triangle distance by projection + edge clamping (float64)."""
    ab, ac = b - a, c - a
    n = np.cross(ab, ac)
    n /= np.linalg.norm(n)
    q = p - np.dot(p - a, n) * n
    # barycentric test
    v0, v1, v2 = ac, ab, q - a
    d00, d01, d11, d20, d21 = v0 @ v0, v0 @ v1, v1 @ v1, v2 @ v0, v2 @ v1
    den = d00 * d11 - d01 * d01
    u = (d11 * d20 - d01 * d21) / den
    v = (d00 * d21 - d01 * d20) / den
    if u >= 0 and v >= 0 and u + v <= 1:
        return np.linalg.norm(p - q)

    def seg(p, x, y):
        t = np.clip(np.dot(p - x, y - x) / np.dot(y - x, y - x), 0, 1)
        return np.linalg.norm(p - (x + t * (y - x)))
    return min(seg(p, a, b), seg(p, b, c), seg(p, c, a))


def test_brute_force_distance_tiny_tetra(tmp_path):
    """Non-axis-aligned closed mesh: |S| equals the brute-force min distance over all
    triangles (to float rounding) inside the band, and the sign marks the inside."""
    p = tmp_path / "tet.obj"
    p.write_text("v 0 0 0\nv 1.0 0.1 0.05\nv 0.2 0.9 0.1\nv 0.15 0.2 0.8\n"
                 "g a\nf 1 3 2\ng b\nf 1 2 4\ng c\nf 2 3 4\ng d\nf 1 4 3\n")
    h = np.float32(0.1)
    sc = Scene.from_obj(str(p), h)
    Nx, Ny, Nz = sc.dims
    V = np.array([[0, 0, 0], [1.0, 0.1, 0.05], [0.2, 0.9, 0.1], [0.15, 0.2, 0.8]], np.float32).astype(float)
    V += 3 * float(h)
    T = [(0, 2, 1), (0, 1, 3), (1, 2, 3), (0, 3, 2)]
    S = sc.S.reshape(Nz, Ny, Nx)
    checked = 0
    for k in range(Nz):
        for j in range(Ny):
            for i in range(Nx):
                q = (np.array([i, j, k]) + 0.5) * float(h)
                d = min(_tri_dist(q, V[a], V[b], V[c]) for a, b, c in T)
                if d < 7.5 * float(h):
                    assert abs(abs(S[k, j, i]) - d) < 1e-5
                    checked += 1
    assert checked > 100
    assert (sc.S > 0).sum() > 0


def test_device_count_and_gpu_classification_without_device():
    """fases_device_count always succeeds; asking for the GPU classification where no
    device is visible is an error (FasesError), never a silent host run."""
    import ctypes as C
    from fases_amd._lib import FasesError, load
    n = C.c_int(-1)
    assert load().fases_device_count(C.byref(n)) == 0 and n.value >= 0
    if n.value > 0:
        pytest.skip("a HIP device is visible: covered by tests/test_gpu_geometry.py")
    with pytest.raises(FasesError):
        Scene.shoebox(8, 8, 8, np.float32(0.1), device=0)
