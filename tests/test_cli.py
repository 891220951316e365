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


OCT = (63.0, 125.0, 250.0, 500.0, 1000.0, 2000.0, 4000.0, 8000.0)


def set_config_line(paths, index, value):
    """config.txt line index + 1 (conf_arr[index]) set to value."""
    cfg = Path(paths["config"])
    lines = cfg.read_text().splitlines()
    lines[index] = str(value)
    cfg.write_text("\n".join(lines) + "\n")


def ga_project(tmp, *, order=2, rays=0, device=None, extra=()):
    """make_project with ISM order (config line 13), ray count (line 15), per-wall .mat
    absorption and .sctr scattering files; returns (paths, alpha [6, 8], sctr [6])."""
    ext = tuple(extra) + (("@device %d" % device,) if device is not None else ())
    paths = make_project(tmp, precision=1, t_max=0.05, extra=ext)
    set_config_line(paths, 12, order)
    set_config_line(paths, 14, rays)
    root = Path(paths["config"]).parent.parent
    rng = np.random.default_rng(5)
    alpha = rng.uniform(0.05, 0.5, (6, 8)).astype(np.float32)
    sctr = np.full(6, 0.01, np.float32)
    (root / "input" / "mat").mkdir(parents=True, exist_ok=True)
    (root / "input" / "sctr").mkdir(parents=True, exist_ok=True)
    for q, g in enumerate(WALLS):
        (root / "input" / "mat" / (g + ".mat")).write_text("# alpha 63..8k\n" + " ".join("%.9g" % a for a in alpha[q]) + "\n")
    for q, g in enumerate(WALLS[:3]):              # the rest keep the 0.01 default
        sctr[q] = np.float32(0.1 * (q + 1))
        (root / "input" / "sctr" / (g + ".sctr")).write_text("# scattering\n%.9g  # global\n" % sctr[q])
    return paths, alpha, sctr


def ga_reference_run(paths, alpha_by_wall, sctr_by_wall, order, rays, device):
    """The library's fases_ga_run on the project's scene, sources and receivers, with the
    tables the CLI derives: .mat absorption and .sctr scattering per scene group, air
    attenuation at the config's temperature, 2e-5 10^(Lp / 20) source pressure."""
    import ctypes as C
    from fases_amd import _lib
    from fases_amd import ga as GA
    from fases_amd.solver import Scene
    wl = P.load_project(**paths)
    sc = Scene.from_obj(paths["scene"], np.float32(0.02))
    idx = [WALLS.index(n) for n in sc.group_names]
    alpha, sctr = alpha_by_wall[idx], sctr_by_wall[idx]
    m_air = np.zeros(8, np.float32)
    oct_ = np.array(OCT, np.float32)
    _lib.check(_lib.load().fases_air_attenuation(C.c_float(np.float32(20.0) + np.float32(273.15)), C.c_float(0.0),
                                                C.c_float(1.1), oct_.ctypes.data_as(C.POINTER(C.c_float)), 8,
                                                m_air.ctypes.data_as(C.POINTER(C.c_float))))
    src = np.array([[0.3, 0.25, 0.24]], np.float32)
    lp = np.float32(80.0)
    src_p = np.array([np.float32(2e-5 * float(np.float32(10.0 ** (lp / np.float32(20.0)))))], np.float32)
    rec = np.array([(0.52, 0.31, 0.22), (0.21, 0.44, 0.3)], np.float32)
    Nt = int(np.float32(0.05) / np.float32(1.0 / 44100.0))
    g = GA.run(sc, src, rec, GA.GAParams(device=device, ism_order=order, n_rays=rays, c=wl.params.c, Nt=Nt),
               src_pressure=src_p, alpha=alpha, sctr=sctr, m_air=m_air)
    return g, Nt


def test_ga_only_writes_band_files_and_keeps_wave_results(tmp_path):
    """--ga-only on the host (@device -1: image sources without the ray tracer): HIR / HRT
    per receiver and octave band above the 1 kHz wave band (2, 4, 8 kHz), ARR per receiver;
    stale HIR files go, the wave-based IR files stay. The files hold what fases_ga_run of
    the same scene and tables gives (its synthesis is pinned on its own)."""
    paths, alpha, sctr = ga_project(tmp_path, order=2, device=-1)
    out = Path(paths["config"]).parent.parent / "output"
    (out / "IR0_0.bin").write_bytes(b"\1" * 8)
    (out / "HIR0_0_63.bin").write_bytes(b"\1" * 8)
    (out / "header.txt").write_text("kept")
    r = run_cli(paths, "--ga-only")
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert (out / "IR0_0.bin").read_bytes() == b"\1" * 8
    assert (out / "header.txt").read_text() == "kept"
    assert not (out / "HIR0_0_63.bin").exists()
    assert "[GA] image-source paths cover orders 0 .. 2; specular ray paths start at order 3" in r.stdout
    g, Nt = ga_reference_run(paths, alpha, sctr, 2, 0, -1)
    written = sorted(p.name for p in out.glob("HIR*.bin"))
    assert written == sorted("HIR%d_0_%d.bin" % (q, f) for q in range(2) for f in (2000, 4000, 8000))
    for b in (5, 6, 7):
        hf, rt = g.band_irs(b)
        for q in range(2):
            got = np.fromfile(out / ("HIR%d_0_%d.bin" % (q, OCT[b])), np.float32)
            got_rt = np.fromfile(out / ("HRT%d_0_%d.bin" % (q, OCT[b])), np.float32)
            assert got.size == Nt and np.array_equal(got, hf[q])
            assert np.array_equal(got_rt, rt[q]) and not got_rt.any()
            assert np.abs(got).max() > 0
    for q in range(2):
        raw = np.fromfile(out / ("ARR%d_0.bin" % q), np.uint8)
        n = int(raw[:4].view(np.int32)[0])
        arr = raw[4:].view(np.float32).reshape(-1, 12)
        assert arr.shape[0] == n and np.array_equal(arr, g.arrivals(q))
        assert n == 1 + len(g.images(0, q))


def test_low_frequency_config_runs_only_the_geometrical_part(tmp_path):
    """A config frequency below the 63 Hz band's lower edge runs no wave-based part: no
    IR / header files. The seam is still that band's upper edge (89 Hz), so the bands
    from 125 Hz up are written (main3D.cu:141-150, 1602)."""
    paths, alpha, sctr = ga_project(tmp_path, order=1, device=-1)
    set_config_line(paths, 2, 40.0)
    out = Path(paths["config"]).parent.parent / "output"
    r = run_cli(paths)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert not list(out.glob("IR*.bin")) and not (out / "header.txt").exists()
    assert sorted(int(p.stem.split("_")[2]) for p in out.glob("HIR0_0_*.bin")) == [int(f) for f in OCT[1:]]


@pytest.mark.skipif(not __import__("oracle.oracle", fromlist=["x"]).ref_available(),
                    reason="reference checker not built (oracle/_ref)")
def test_dump_rays_file_equals_the_reference(tmp_path):
    """--dump-rays N M writes output/ray_paths.txt and exits: the reference's rt_dump_paths
    file for the same scene and sources, line for line."""
    from oracle import oracle as O
    paths = make_project(tmp_path)
    src = Path(paths["src"])
    src.write_text(src.read_text() + "0.6 0.4 0.3 75 500\n")
    out = Path(paths["config"]).parent.parent / "output"
    r = run_cli(paths, "--dump-rays", "64", "4")
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "[OK] output/ray_paths.txt: 64 rays per source, at most 4 reflections each" in r.stdout
    assert not list(out.glob("IR*.bin"))
    ref = tmp_path / "ref_paths.txt"
    O.ref_dump_rays(paths["scene"], np.float32(0.02), [(0.3, 0.25, 0.24), (0.6, 0.4, 0.3)], 64, 4, str(ref))
    mine = (out / "ray_paths.txt").read_text().splitlines()
    theirs = ref.read_text().splitlines()
    assert mine == theirs
    assert sum(1 for ln in mine if not ln.startswith("#")) >= 2 * 64


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



@pytest.mark.gpu
def test_cli_runs_wave_part_then_geometrical_part(tmp_path):
    """Full CLI run on the GPU: wave-based IR files, then image sources (GPU validity) and
    the ray tracer. Arrival lists are deterministic and equal the library's; the ray-traced
    HRT energy per band matches the library's run within the Monte-Carlo spread of two
    runs (float atomics sum in another order)."""
    paths, alpha, sctr = ga_project(tmp_path, order=2, rays=4000)
    out = Path(paths["config"]).parent.parent / "output"
    print("[case] fases3D wave + GA", flush=True)
    r = run_cli(paths)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert (out / "IR0_0.bin").exists() and (out / "header.txt").exists()
    assert "[OK] HIR / HRT / ARR files written" in r.stdout
    g, Nt = ga_reference_run(paths, alpha, sctr, 2, 4000, 0)
    for q in range(2):
        raw = np.fromfile(out / ("ARR%d_0.bin" % q), np.uint8)
        arr = raw[4:].view(np.float32).reshape(-1, 12)
        assert np.array_equal(arr, g.arrivals(q))
    for b in (5, 6, 7):
        _, rt = g.band_irs(b)
        for q in range(2):
            got = np.fromfile(out / ("HRT%d_0_%d.bin" % (q, OCT[b])), np.float32)
            assert got.size == Nt and np.isfinite(got).all()
            e_cli, e_lib = float((got.astype(np.float64) ** 2).sum()), float((rt[q].astype(np.float64) ** 2).sum())
            assert e_lib > 0 and abs(e_cli - e_lib) <= 0.05 * e_lib, (b, q, e_cli, e_lib)


def test_scene_cache_written_reused_and_invalidated(tmp_path):
    """output/S.bin, oid.bin and scene_cache.txt (main3D.cu:334-428): a real run writes the
    classification and the key (scene path, scene file time, h, tol = 3h) with the grid;
    the next run reads S and oid from the cache (an edit of S.bin shows up in Vol.bin);
    touching the scene invalidates it; --estimate reads but never writes it."""
    import os
    from fases_amd.solver import Scene
    paths, _, _ = ga_project(tmp_path, order=0, device=-1)
    out = Path(paths["config"]).parent.parent / "output"
    r = run_cli(paths, "--estimate")
    assert r.returncode == 0, r.stderr
    assert not (out / "scene_cache.txt").exists() and not (out / "S.bin").exists()
    r = run_cli(paths, "--ga-only")
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "[scene] classifying the scene" in r.stdout
    sc = Scene.from_obj(paths["scene"], np.float32(0.02))
    Nx, Ny, Nz = sc.dims
    S = np.fromfile(out / "S.bin", np.float32)
    oid = np.fromfile(out / "oid.bin", np.uint8)
    assert np.array_equal(S.view(np.uint32), sc.S.ravel().view(np.uint32)) and np.array_equal(oid, sc.oid.ravel())
    key, dims = (out / "scene_cache.txt").read_text().splitlines()
    assert key.startswith(str(Path(paths["scene"]).resolve()) + " ") and key.endswith(" 0.02 0.06")
    assert [int(v) for v in dims.split()[:3]] == [Nx, Ny, Nz]
    # reuse: make a layer of air cells solid in the cached S; the volume follows the cache
    S2 = S.copy()
    air = np.flatnonzero(S2 > 0)[:500]
    S2[air] = -1.0
    S2.tofile(out / "S.bin")
    r = run_cli(paths, "--estimate")
    assert r.returncode == 0 and "read from output/S.bin" in r.stdout, r.stdout[-2000:]
    h = np.float32(0.02)
    vol = np.fromfile(Path(paths["config"]).parent / "Vol.bin", np.float32)[0]
    assert vol == ((np.float32((S2 > 0).sum()) * h) * h) * h
    # a newer scene file invalidates the key
    st = os.stat(paths["scene"])
    os.utime(paths["scene"], ns=(st.st_atime_ns, st.st_mtime_ns + 10 ** 9))
    r = run_cli(paths, "--estimate")
    assert r.returncode == 0 and "[scene] classifying the scene" in r.stdout
