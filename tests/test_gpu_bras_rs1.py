"""The reference's own full project, 3D-solver/tests/BRAS_RS1 (a byte copy sits in
tests/golden/bras_rs1): h = 0.045 m, t_max = 0.15 s, PCS source, poles.txt "2 5", cPML on
five faces (U R L F B; the floor, D, reflects), the real multi-pole .fit files.

Its shipped sketch.obj names its groups floor / wall_1..4 / ceiling, but the project holds
no floor.fit or wall_N.fit: the reference looks each group up as input/fit/<group>.fit
(main3D.cu:622-634) and read_fit_file exits on a missing one (filterABC.h:43-47), so the
project as shipped stops there, and so does fases3D. The runs below use the same room
(sketch.obj's vertices) with its groups named after the project's fits, as the original
RS1 scene names them: the floor is the absorbing plate, split into 2 x 2 tiles of four
C-RockFon_Sonar_G fits (K = 2, L = 5), and the walls and ceiling are the project's
wall_x0/x1/y0/y1 and ceiling layers."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from fases_amd import project as P
from fases_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
CLI = ROOT / "fases_amd" / "fases3D"
SRC = ROOT / "tests" / "golden" / "bras_rs1"
ORACLE_STEPS = 200


def sketch_box():
    xs, ys, zs = [], [], []
    for line in (SRC / "input" / "scene" / "sketch.obj").read_text().splitlines():
        t = line.split()
        if t and t[0] == "v":
            xs.append(float(t[1])); ys.append(float(t[2])); zs.append(float(t[3]))
    return max(xs), max(ys), max(zs)


def rs1_obj(lx, ly, lz):
    """The sketch room with the project's layer names; the floor as 2 x 2 plate tiles."""
    v = []

    def vert(x, y, z):
        v.append((x, y, z))
        return len(v)

    out = []
    hx, hy = lx / 2, ly / 2
    tiles = ["C-RockFon_Sonar_G", "C-RockFon_Sonar_G.001", "C-RockFon_Sonar_G.002", "C-RockFon_Sonar_G.003"]
    for q, (x0, y0) in enumerate(((0, 0), (hx, 0), (0, hy), (hx, hy))):
        a, b = vert(x0, y0, 0), vert(x0 + hx, y0, 0)
        c, d = vert(x0 + hx, y0 + hy, 0), vert(x0, y0 + hy, 0)
        out += ["g " + tiles[q], "f %d %d %d %d" % (a, b, c, d)]
    c8 = [vert(*p) for p in ((0, 0, 0), (lx, 0, 0), (lx, ly, 0), (0, ly, 0), (0, 0, lz), (lx, 0, lz), (lx, ly, lz),
                             (0, ly, lz))]
    faces = {"wall_x0": (4, 8, 5, 1), "wall_x1": (2, 6, 7, 3), "wall_y0": (1, 5, 6, 2), "wall_y1": (3, 7, 8, 4),
             "ceiling": (5, 8, 7, 6)}
    for g, f in faces.items():
        out += ["g " + g, "f %d %d %d %d" % tuple(c8[i - 1] for i in f)]
    return "\n".join(["# BRAS RS1 room (sketch.obj extents) with the project's layer names"]
                     + ["v %r %r %r" % p for p in v] + out) + "\n"


def project(tmp: Path, precision: int, scene_text=None) -> dict:
    root = tmp / "BRAS_RS1"
    shutil.copytree(SRC, root)
    (root / "output").mkdir()
    scene = root / "input" / "scene" / "sketch.obj"
    if scene_text is not None:
        scene = root / "input" / "scene" / "rs1.obj"
        scene.write_text(scene_text)
    cfg = root / "config" / "config.txt"
    cfg.write_text(cfg.read_text() + "@precision %d\n" % precision)
    return {"scene": str(scene), "config": str(cfg), "src": str(root / "input" / "src" / "src.txt"),
            "rec": str(root / "input" / "rec" / "rec.txt")}


def run_cli(paths, *args, timeout=900):
    cmd = [str(CLI), paths["scene"], paths["config"], paths["src"], paths["rec"], *args]
    print("[case] " + " ".join(Path(c).name for c in cmd), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)


def test_shipped_project_stops_at_its_missing_fit_file(tmp_path):
    paths = project(tmp_path, 1)
    r = run_cli(paths, "--host-geometry")
    assert r.returncode != 0
    assert "floor.fit" in r.stderr


@pytest.fixture(scope="module")
def rs1(tmp_path_factory):
    obj = rs1_obj(*sketch_box())
    out = {}
    for prec in (0, 1):
        paths = project(tmp_path_factory.mktemp("p%d" % prec), prec, obj)
        r = run_cli(paths, "--fp64-ir")
        assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
        o = Path(paths["config"]).parent.parent / "output"
        out[prec] = (paths, np.fromfile(o / "IR0_0.f64.bin", np.float64), np.fromfile(o / "IR0_0.bin", np.float32))
    return out


@pytest.mark.timeout(900)
def test_rs1_setup_matches_the_project(rs1):
    paths, ir64, ir32 = rs1[0]
    wl = P.load_project(**paths)
    # poles 2 5; four plate tiles with real multi-pole fits; absorbing walls (R_inf 0)
    assert wl.materials.K == 2 and wl.materials.L == 5
    names = wl.scene.group_names
    A = wl.materials.A
    for q, g in enumerate(names):
        if g.startswith("C-RockFon"):
            assert np.count_nonzero(A[q]) == 2 and np.count_nonzero(wl.materials.B[q]) == 5
            assert 0.5 < wl.materials.Rinf[q] < 0.7
        else:
            assert wl.materials.Rinf[q] == 0.0 and not A[q].any()
    assert tuple(wl.params.pml_keys) == (1.0, 0.0, 1.0, 1.0, 1.0, 1.0)
    assert wl.src_opt == 0 and ir64.size == wl.nsteps and wl.nsteps > 2000
    assert np.array_equal(ir32, ir64.astype(np.float32))


@pytest.mark.timeout(900)
def test_rs1_fp64_matches_oracle(rs1):
    """The FP64 run of the project (through the library, as the CLI runs it) against the
    FP64 oracle over the first 200 steps: every field bit for bit within 1e-10, and the
    CLI's FP64 IR file equals the library run's IR."""
    from oracle import oracle as O
    paths, ir64, _ = rs1[0]
    wl = P.load_project(**paths)
    sv = W.make_solver(wl, "fp64")
    ir, _, _, _ = sv.run(ORACLE_STEPS)
    f = sv.get_fields()
    sv.close()
    np.testing.assert_array_equal(ir[0], ir64[:ORACLE_STEPS])
    print("[case] oracle %d steps of %s cells" % (ORACLE_STEPS, wl.dims), flush=True)
    ref = O.run_workload(wl, nsteps=ORACLE_STEPS, want_fields=True)
    for a, b in zip(f, ref["fields"]):
        assert np.abs(b).max() > 0
        assert np.linalg.norm(a - b) <= 1e-10 * np.linalg.norm(b)
    np.testing.assert_allclose(ir, ref["ir"], rtol=0, atol=1e-10 * max(np.abs(ref["ir"]).max(), 1e-30))


def test_rs1_fp32_full_length_within_1e5_of_fp64(rs1):
    _, ir64, _ = rs1[0]
    _, ir32_64, _ = rs1[1]
    assert np.abs(ir64).max() > 1e-3
    err = np.linalg.norm(ir32_64 - ir64) / np.linalg.norm(ir64)
    print("BRAS RS1 FP32 vs FP64 over %d steps: IR rel L2 %.2e" % (ir64.size, err))
    assert err <= 1e-5
