"""fases3D command-line drop-in: project layout, file formats, error exits and, on a GPU,
the IR files against the FP64 oracle run on the same project.

The reference's executable (3D-solver/core/main3D.cu) takes
`<scene.obj> <config.txt> <src.txt> <rec.txt> [poles.txt]`, writes
output/IR<r>_<s>.bin (float32 per step), VEL<r>_<s>.bin, header.txt, pre_header.txt,
config/Vol.bin, config/p_ref.bin, and prints "NN%" progress lines read by common/gui.py.
"""
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from fases_amd import project as P
from fases_amd import workloads as W
from fases_amd.solver import write_fit

ROOT = Path(__file__).resolve().parent.parent
CLI = ROOT / "fases_amd" / "fases3D"

WALLS = ("west", "east", "south", "north", "floor", "ceiling")


def box_obj(lx, ly, lz):
    v = [(0, 0, 0), (lx, 0, 0), (lx, ly, 0), (0, ly, 0), (0, 0, lz), (lx, 0, lz), (lx, ly, lz), (0, ly, lz)]
    faces = {"west": (4, 8, 5, 1), "east": (2, 6, 7, 3), "south": (1, 5, 6, 2), "north": (3, 7, 8, 4),
             "floor": (1, 2, 3, 4), "ceiling": (5, 8, 7, 6)}
    out = ["o Box", ""] + ["v %r %r %r" % p for p in v] + [""]
    for g in WALLS:
        out += ["g " + g, "f %d %d %d %d" % faces[g], ""]
    return "\n".join(out)


def make_project(tmp, *, precision=0, t_max=0.006, src_opt=1, K=2, L=1, seed=11, pml=(0, 0, 0, 0, 0, 0),
                 rec_rows=((0.52, 0.31, 0.22), (0.21, 0.44, 0.3)), extra=()):
    cfg = [0.02, t_max, 1000.0, 20.0, 0.5, 12, *pml, 0, 0, 0, src_opt, 0.0 if src_opt != 1 else 0.0, 0]
    if src_opt == 1:
        cfg[16] = 900.0                       # Ricker peak frequency (config line 17)
    cfg = [str(v) for v in cfg] + ["@precision %d" % precision] + list(extra)
    rng = np.random.default_rng(seed)
    dt = float(W.time_step(0.5, np.float32(0.02), 343.0))
    fits = {}
    for g in WALLS:
        d = W.draw_material(rng, K, L, dt)
        p = Path(tmp) / ("fit_" + g)
        write_fit(str(p), d["Rinf"], d["A"], d["zeta"], d["B"], d["C"], d["alpha"], d["beta"])
        fits[g] = p.read_text()
    paths = P.write_project(Path(tmp) / "proj", box_obj(0.8, 0.6, 0.5), cfg,
                            [(0.3, 0.25, 0.24, 80.0, 500.0)], list(rec_rows), poles=(K, L), fits=fits)
    return paths


def run_cli(paths, *args, timeout=600):
    cmd = [str(CLI), paths["scene"], paths["config"], paths["src"], paths["rec"], *args]
    return subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)


@pytest.fixture(scope="module", autouse=True)
def _cli_built():
    if not CLI.exists():
        pytest.skip("fases3D not built (python -m fases_amd.build)")


def test_estimate_prints_mem_line_and_volume(tmp_path):
    paths = make_project(tmp_path)
    r = run_cli(paths, "--estimate")
    assert r.returncode == 0, r.stderr
    mem = [ln for ln in r.stdout.splitlines() if ln.startswith("[MEM]")]
    assert len(mem) == 1
    kv = dict(tok.split("=") for tok in mem[0].split()[1:])
    wl = P.load_project(**paths)
    assert int(kv["voxels"]) == int(np.prod(wl.dims))
    vol = np.fromfile(Path(paths["config"]).parent / "Vol.bin", np.float32)
    n_air = int((wl.scene.S > 0).sum())
    h = np.float32(0.02)
    assert vol.size == 1 and vol[0] == ((np.float32(n_air) * h) * h) * h
    # the box is 0.8 x 0.6 x 0.5 m; the air volume sits within a cell layer of it
    assert abs(vol[0] - 0.24) < 0.02


def test_missing_poles_file_fails(tmp_path):
    paths = make_project(tmp_path)
    (Path(paths["config"]).parent / "poles.txt").unlink()
    r = run_cli(paths)
    assert r.returncode != 0
    assert "poles" in r.stderr


def test_missing_fit_file_fails(tmp_path):
    paths = make_project(tmp_path)
    (Path(paths["config"]).parent.parent / "input" / "fit" / "floor.fit").unlink()
    r = run_cli(paths)
    assert r.returncode != 0
    assert "floor.fit" in r.stderr


def test_malformed_fit_file_fails(tmp_path):
    paths = make_project(tmp_path)
    fit = Path(paths["config"]).parent.parent / "input" / "fit" / "east.fit"
    fit.write_text("# R_inf\n0.9\n# K L\n2 1\n# A_k zeta_k\n1.0\n")
    r = run_cli(paths)
    assert r.returncode != 0
    assert "east.fit" in r.stderr


def test_receiver_outside_domain_stops(tmp_path):
    paths = make_project(tmp_path, rec_rows=((0.4, 0.3, 0.2), (5.0, 0.3, 0.2)))
    r = run_cli(paths)
    assert r.returncode != 0
    assert "Receiver(s) outside domain" in r.stdout


def test_geometrical_acoustics_flags_rejected(tmp_path):
    paths = make_project(tmp_path)
    r = run_cli(paths, "--ga-only")
    assert r.returncode != 0 and "geometrical-acoustics" in r.stderr


def test_load_project_matches_reference_derivations(tmp_path):
    """c = sqrt(1.4 * 287.05 * (T + 273.15)) and dt = CFL h / c in float32; Nt = t_max/dt + 1."""
    paths = make_project(tmp_path)
    wl = P.load_project(**paths)
    f32 = np.float32
    c = f32(np.sqrt(f32(1.4) * f32(287.05) * (f32(20.0) + f32(273.15))))
    dt = f32(f32(0.5) * f32(0.02) / c)
    assert wl.params.c == float(c) and wl.params.dt == float(dt)
    assert wl.nsteps == int(f32(0.006) / dt) + 1
    assert wl.params.precision == "fp64"
    assert len(wl.rec_ijk) == 2 and len(wl.src_ijk) == 1
    assert wl.materials.A.shape == (6, 2) and np.all(wl.materials.Rinf > 0)


@pytest.mark.gpu
@pytest.mark.parametrize("src_opt", [1, 2], ids=["ricker", "sine"])
def test_cli_fp64_ir_files_match_oracle(tmp_path, src_opt):
    from oracle import oracle as O
    paths = make_project(tmp_path, precision=0, src_opt=src_opt)
    print("[case] fases3D fp64 src_opt=%d" % src_opt, flush=True)
    r = run_cli(paths, "--fp64-ir")
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    wl = P.load_project(**paths)
    Nt = wl.nsteps
    ref = O.run_workload(wl)["ir"]
    out = Path(paths["config"]).parent.parent / "output"
    for rr in range(len(wl.rec_ijk)):
        ir32 = np.fromfile(out / ("IR%d_0.bin" % rr), np.float32)
        ir64 = np.fromfile(out / ("IR%d_0.f64.bin" % rr), np.float64)
        vel = np.fromfile(out / ("VEL%d_0.bin" % rr), np.float32)
        assert ir32.size == Nt and ir64.size == Nt and vel.size == 3 * Nt
        rl2 = np.linalg.norm(ir64 - ref[rr]) / np.linalg.norm(ref[rr])
        assert rl2 <= 1e-10, rl2
        # float32 file = the FP64 IR rounded once
        assert np.array_equal(ir32, ir64.astype(np.float32))
        assert np.abs(ir64).max() > 0
    hdr = (out / "header.txt").read_text().split()
    assert [int(v) for v in hdr[:4]] == [*wl.dims, Nt]
    pre = (out / "pre_header.txt").read_text().split()
    assert [int(v) for v in pre[:3]] == list(wl.dims)
    # one "NN%" line per 5 % tick of n = 0 .. Nt-1, as functions3D.cuh:1431-1439 prints them
    ticks = [ln for ln in r.stdout.splitlines() if ln.endswith("%") and ln[:-1].isdigit()]
    step = max(Nt // 20, 1)
    assert ticks == ["%d%%" % int(100.0 * n / Nt + 0.5) for n in range(Nt) if n % step == 0]
    if src_opt == 1:
        pref = np.fromfile(Path(paths["config"]).parent / "p_ref.bin", np.float64)
        assert pref.size == 2 * Nt + 1
        # p_ref is scaled to the source's Lp,1m: its peak is the 80 dB target pressure
        assert abs(np.abs(pref).max() - 2e-5 * 10 ** (80 / 20)) < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("extra", [(), ("@source_patch 20",)])
def test_cli_fp32_default_precision_within_tolerance(tmp_path, extra):
    """FP32 CLI run against the FP64 oracle; also with a wider FP64 source region set from
    config.txt (@source_patch)."""
    from oracle import oracle as O
    paths = make_project(tmp_path, precision=1, extra=extra)
    print("[case] fases3D fp32 %s" % (extra,), flush=True)
    r = run_cli(paths, "--fp64-ir")
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    wl = P.load_project(**paths)
    ref = O.run_workload(wl)["ir"]
    out = Path(paths["config"]).parent.parent / "output"
    for rr in range(len(wl.rec_ijk)):
        ir64 = np.fromfile(out / ("IR%d_0.f64.bin" % rr), np.float64)
        rl2 = np.linalg.norm(ir64 - ref[rr]) / np.linalg.norm(ref[rr])
        assert rl2 <= 1e-5, rl2


@pytest.mark.gpu
def test_cli_keep_other_sources(tmp_path):
    """Stale IR files of other source indices survive with --keep-other-sources."""
    paths = make_project(tmp_path, precision=1, t_max=0.001)
    out = Path(paths["config"]).parent.parent / "output"
    (out / "IR0_5.bin").write_bytes(b"\0" * 8)
    (out / "IR1_0.bin").write_bytes(b"\0" * 8)
    print("[case] fases3D keep-other-sources", flush=True)
    r = run_cli(paths, "--keep-other-sources")
    assert r.returncode == 0, r.stderr
    assert (out / "IR0_5.bin").exists()
    assert (out / "IR1_0.bin").stat().st_size > 8
    r = run_cli(paths)
    assert r.returncode == 0, r.stderr
    assert not (out / "IR0_5.bin").exists()


@pytest.mark.gpu
def test_cli_animation_frames_match_oracle_fields(tmp_path):
    """Config lines 19-22 (slice planes at x, y, z metres; 60 frames requested): a frame
    of pressureX/Y/Z.bin and of the six unit energy-flow files after every anim_step-th
    step (main3D.cu:1322-1388), header pbin = frames written. The planes of the frame
    after step n are the FP64 fields after n + 1 steps rounded to float32 (the FP64 path
    equals the oracle), and the flow direction is sign(p) (u, v, w) / (|(u, v, w)| + 1e-20)."""
    from oracle import oracle as O
    paths = make_project(tmp_path, precision=0, extra=("0.4", "0.3", "0.2", "60"))
    print("[case] fases3D animation frames", flush=True)
    r = run_cli(paths)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    wl = P.load_project(**paths)
    Nx, Ny, Nz = wl.dims
    Nt = wl.nsteps
    out = Path(paths["config"]).parent.parent / "output"
    hdr = (out / "header.txt").read_text().split()
    pbin, anim_step = int(hdr[7]), int(hdr[8])
    assert anim_step == max(Nt // min(60, Nt), 1)
    assert pbin == len(range(0, Nt, anim_step))
    f32 = np.float32
    i0, j0, k0 = [int(f32(x) / f32(0.02)) + 3 for x in (0.4, 0.3, 0.2)]
    shapes = {"X": (Nz, Ny), "Y": (Nz, Nx), "Z": (Ny, Nx)}
    frames = {}
    for ax, shp in shapes.items():
        a = np.fromfile(out / ("pressure%s.bin" % ax), np.float32)
        assert a.size == pbin * shp[0] * shp[1]
        frames[ax] = a.reshape(pbin, *shp)
    comps = {"X": ("Iy_X", "Iz_X"), "Y": ("Ix_Y", "Iz_Y"), "Z": ("Ix_Z", "Iy_Z")}
    dirs = {}
    for ax, names in comps.items():
        dirs[ax] = [np.fromfile(out / (n + ".bin"), np.float32).reshape(frames[ax].shape) for n in names]
    # frame q follows step n = q * anim_step, i.e. n + 1 steps from the start
    q = pbin // 2
    n_steps = q * anim_step + 1
    f = O.run_workload(wl, nsteps=n_steps, want_fields=True)["fields"]
    p, u, v, w = [np.asarray(x).reshape(Nz, Ny, Nx).astype(np.float32) for x in f]
    planes = {"X": (lambda a: a[:, :, i0]), "Y": (lambda a: a[:, j0, :]), "Z": (lambda a: a[k0, :, :])}
    inplane = {"X": (v, w), "Y": (u, w), "Z": (u, v)}
    for ax, cut in planes.items():
        assert np.array_equal(frames[ax][q], cut(p)), ax
        pp, uu, vv, ww = cut(p), cut(u), cut(v), cut(w)
        sp = np.where(pp >= 0, np.float32(1), np.float32(-1))
        inv = sp / (np.sqrt(uu * uu + vv * vv + ww * ww) + np.float32(1e-20))
        for got, comp in zip(dirs[ax], inplane[ax]):
            np.testing.assert_allclose(got[q], cut(comp) * inv, rtol=1e-6, atol=0)
    assert np.abs(frames["Z"][q]).max() > 0


@pytest.mark.gpu
def test_cli_gpu_and_host_classification_give_identical_runs(tmp_path):
    """The CLI classifies the scene on the GPU by default; --host-geometry on the host.
    Same S and oid, so the IR files are byte-identical."""
    paths = make_project(tmp_path, precision=1, t_max=0.002)
    out = Path(paths["config"]).parent.parent / "output"
    print("[case] fases3D gpu vs host classification", flush=True)
    r = run_cli(paths)
    assert r.returncode == 0, r.stderr
    gpu = [(out / ("IR%d_0.bin" % q)).read_bytes() for q in range(2)]
    r = run_cli(paths, "--host-geometry")
    assert r.returncode == 0, r.stderr
    host = [(out / ("IR%d_0.bin" % q)).read_bytes() for q in range(2)]
    assert gpu == host and len(gpu[0]) > 0

