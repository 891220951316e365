"""A reference project directory (config/, input/scene|src|rec|fit, output/) loaded into a
Workload, with the same derivations as the fases3D command-line drop-in
(3D-solver/core/main3D.cu:101-1124): float32 h, t_max, T, CFL; c = sqrt(1.4 * 287.05 * T),
rho = 1.225 unless overridden by '@c' / '@rho' lines; dt = CFL h / c and Nt = t_max/dt + 1
in float32; sources scaled to their Lp,1m by the analytic 1 m peak; the shared waveform
stored as float32."""
from __future__ import annotations

import math
import os
from pathlib import Path

import numpy as np

from . import _lib
from .solver import Materials, Scene, SolverParams, read_config, read_fit_files, read_poles, read_receivers
from .solver import read_sources, source_pref, source_waveform
from .workloads import Workload

OCTAVES = (63.0, 125.0, 250.0, 500.0, 1000.0, 2000.0, 4000.0, 8000.0)


def extension_keys(config: str) -> dict:
    keys = {}
    for line in open(config):
        line = line.split("#", 1)[0].split()
        if len(line) >= 2 and line[0].startswith("@"):
            try:
                keys[line[0][1:]] = float(line[1])
            except ValueError:
                pass
    return keys


def load_project(scene: str, config: str, src: str, rec: str) -> Workload:
    f32 = np.float32
    conf = read_config(config)
    ext = extension_keys(config)
    h, t_max, f_in = f32(conf[0]), f32(conf[1]), f32(conf[2])
    f_x = f_in
    for oc in OCTAVES:
        hi = f32(oc) * f32(1.41421356)
        if f_in <= hi:
            f_x = hi
            break
    f_cut = f32(2.0) * f32(f_x)
    temp = f32(conf[3]) + f32(273.15)
    c = f32(np.sqrt(f32(f32(1.4) * f32(287.05)) * temp))
    rho = f32(1.225)
    if "c" in ext:
        c = f32(ext["c"])
    if "rho" in ext:
        rho = f32(ext["rho"])
    dt = f32(f32(f32(conf[4]) * h) / c)
    Nt = int(f32(t_max / dt)) + 1
    src_opt = int(conf[15]) if conf.size > 15 else 0
    glob_f = float(conf[16]) if conf.size > 16 else (25.0 if src_opt == 0 else 0.5 * float(f_cut))
    prec = {0: "fp64", 1: "fp32", 2: "fp16"}[int(ext.get("precision", 1))]

    sc = Scene.from_obj(scene, float(h))
    cfg_dir = Path(config).resolve().parent
    root = cfg_dir.parent if (Path(config).name == "config.txt" and cfg_dir.name == "config") else Path(os.getcwd())
    K, L = read_poles(str(cfg_dir / "poles.txt"))
    paths = []
    for g in sc.group_names:
        if "ounding" in g or "oundary" in g:
            paths.append(None)
        else:
            paths.append(str(root / "input" / "fit" / (g + ".fit")))
    mats = read_fit_files(paths, K, L) if paths else Materials.rigid(1, 0.0)

    cells, lp, f0, _ = read_sources(src, float(h), sc.dims, float(f32(0.5) * f_cut))
    sval = np.array([2e-5 * math.pow(10.0, float(v) / 20.0) for v in lp], np.float64)
    u_s = None
    A_s = 6.0 * float(h) * float(h)
    if src_opt == _lib.FASES_SRC_SINE:
        freq = [float(v) for v in f0]
        sval = np.array([s / (float(rho) * A_s * fr / 2.0) for s, fr in zip(sval, freq)], np.float64)
    else:
        freq = None
        u_s = source_waveform(src_opt, float(dt), glob_f, float(f_cut), Nt)
        _, peak = source_pref(u_s, Nt, float(dt), float(h), float(rho))
        sval = sval / peak
        u_s = u_s.astype(np.float32).astype(np.float64)
    rcells, _ = read_receivers(rec, float(h), sc.dims)
    Nx, Ny, _ = sc.dims
    ijk = lambda cidx: (int(cidx % Nx), int((cidx // Nx) % Ny), int(cidx // (Nx * Ny)))
    prm = SolverParams(h=float(h), dt=float(dt), c=float(c), rho=float(rho), precision=prec,
                       pml_keys=tuple(float(v) for v in conf[6:12]),
                       pml_cells=int(ext.get("pml_cells", 20)), pml_sigma_max=float(ext.get("pml_sigma_max", 0.0)),
                       source_patch=int(ext.get("source_patch", 0)))
    wl = Workload(name=Path(scene).stem, room=tuple(d - 6 for d in sc.dims), h=float(h), nsteps=Nt, params=prm,
                  materials=mats, rec_ijk=[ijk(r) for r in rcells], src_ijk=[ijk(s) for s in cells],
                  src_sval=list(sval), src_opt=src_opt, u_s=u_s, src_freq=freq)
    wl._scene = sc
    return wl


def write_project(root: Path, obj_text: str, config_lines: list, src_rows: list, rec_rows: list, poles=(1, 1),
                  fits: dict | None = None) -> dict:
    """Creates a reference-layout project directory; returns the four CLI paths."""
    root = Path(root)
    for d in ("config", "input/scene", "input/src", "input/rec", "input/fit", "output"):
        (root / d).mkdir(parents=True, exist_ok=True)
    (root / "input/scene/scene.obj").write_text(obj_text)
    (root / "config/config.txt").write_text("\n".join(str(v) for v in config_lines) + "\n")
    (root / "config/poles.txt").write_text("%d %d\n" % poles)
    (root / "input/src/src.txt").write_text("# x y z Lp [f0]\n" + "\n".join(" ".join(str(v) for v in r) for r in src_rows) + "\n")
    (root / "input/rec/rec.txt").write_text("# x y z\n" + "\n".join(" ".join(str(v) for v in r) for r in rec_rows) + "\n")
    for name, text in (fits or {}).items():
        (root / "input/fit" / (name + ".fit")).write_text(text)
    return {"scene": str(root / "input/scene/scene.obj"), "config": str(root / "config/config.txt"),
            "src": str(root / "input/src/src.txt"), "rec": str(root / "input/rec/rec.txt")}
