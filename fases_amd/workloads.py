"""Synthetic shoebox workloads of the benchmark suite (rooms, materials, sources, receivers).

Every room is an axis-aligned shoebox meshed exactly like an OBJ with six wall groups
(west/east/south/north/floor/ceiling = x-/x+/y-/y+/z-/z+), classified by the same
SDF/OID pipeline as OBJ scenes. Material i drives wall group i.

Seeded ADE materials: real poles -zeta_k and complex poles -alpha_l +/- j beta_l with
zeta, alpha ~ U[200, 8000] rad/s and beta ~ U[500, 20000] rad/s. Residues start at
U[-1, 1] times the pole magnitude and are then scaled by the largest factor (times 0.9)
that keeps |R(f)| in [0.5, 0.95] on a 2048-point grid from 0 to the Nyquist rate
1/(2 dt), around R_inf ~ U[0.6, 0.85].
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import dataclasses

import numpy as np

from . import _lib
from .solver import GHOST, Materials, Scene, SolverParams, source_pref, source_waveform

LP_1PA = 20.0 * np.log10(1.0 / 2e-5)   # Lp,1m that the reference maps to 1 Pa


def f32(x) -> float:
    return float(np.float32(x))


def write_tessellated_room(path, X: float, Y: float, Z: float, n: int) -> int:
    """OBJ of an X x Y x Z [m] shoebox whose six walls (groups west, east, south, north,
    floor, ceiling) are each split into n x n quads of two triangles: a mesh of 12 n^2
    triangles with shared vertices, for classification at mesh sizes of real scenes.
    Returns the triangle count."""
    V, idx = [], {}

    def vid(p):
        key = tuple(np.round(p, 12))
        if key not in idx:
            idx[key] = len(V)
            V.append(p)
        return idx[key]
    planes = [("west", (0, 0, 0), (0, Y, 0), (0, 0, Z)), ("east", (X, 0, 0), (0, 0, Z), (0, Y, 0)),
              ("south", (0, 0, 0), (0, 0, Z), (X, 0, 0)), ("north", (0, Y, 0), (X, 0, 0), (0, 0, Z)),
              ("floor", (0, 0, 0), (X, 0, 0), (0, Y, 0)), ("ceiling", (0, 0, Z), (0, Y, 0), (X, 0, 0))]
    faces = []
    for g, o, u, w in planes:
        o, u, w = (np.array(v, float) for v in (o, u, w))
        tris = []
        for a in range(n):
            for b in range(n):
                q = [vid(o + u * (a + da) / n + w * (b + db) / n) for da, db in ((0, 0), (1, 0), (1, 1), (0, 1))]
                tris += [(q[0], q[1], q[2]), (q[0], q[2], q[3])]
        faces.append((g, tris))
    with open(path, "w") as f:
        f.write("o room\n")
        for v in V:
            f.write("v %.9g %.9g %.9g\n" % tuple(float(a) for a in v))
        for g, tris in faces:
            f.write("g %s\n" % g)
            for t in tris:
                f.write("f %d %d %d\n" % tuple(i + 1 for i in t))
    return sum(len(t) for _, t in faces)


def time_step(cfl: float, h: float, c: float) -> float:
    """dt = CFL*h/c evaluated in float32 like the reference (main3D.cu:729-730)."""
    return float(np.float32(np.float32(cfl) * np.float32(h)) / np.float32(c))


def draw_material(rng: np.random.Generator, K: int, L: int, dt: float) -> dict:
    zeta = rng.uniform(200.0, 8000.0, K)
    alpha = rng.uniform(200.0, 8000.0, L)
    beta = rng.uniform(500.0, 20000.0, L)
    A = rng.uniform(-1.0, 1.0, K) * zeta
    B = rng.uniform(-1.0, 1.0, L) * np.hypot(alpha, beta)
    C = rng.uniform(-1.0, 1.0, L) * np.hypot(alpha, beta)
    Rinf0 = rng.uniform(0.6, 0.85)
    f = np.linspace(0.0, 0.5 / dt, 2048)
    s = 2j * np.pi * f
    G = np.zeros_like(s)
    for k in range(K):
        G += A[k] / (s + zeta[k])
    for l in range(L):
        G += (B[l] * (s + alpha[l]) + C[l] * beta[l]) / ((s + alpha[l]) ** 2 + beta[l] ** 2)

    def ok(scale):
        m = np.abs(Rinf0 + scale * G)
        return m.min() >= 0.5 and m.max() <= 0.95

    lo, hi = 0.0, 1.0
    if ok(hi):
        lo = hi
    else:
        for _ in range(60):
            mid = 0.5 * (lo + hi)
            lo, hi = (mid, hi) if ok(mid) else (lo, mid)
    sc = 0.9 * lo
    return dict(Rinf=Rinf0, A=A * sc, zeta=zeta, B=B * sc, C=C * sc, alpha=alpha, beta=beta)


def materials_from(mats: list, K: int, L: int) -> Materials:
    m = Materials.empty(len(mats), K, L)
    for i, d in enumerate(mats):
        if d is None:
            continue
        m.Rinf[i] = d["Rinf"]
        m.A[i, : len(d["A"])] = d["A"]; m.zeta[i, : len(d["zeta"])] = d["zeta"]
        m.B[i, : len(d["B"])] = d["B"]; m.C[i, : len(d["C"])] = d["C"]
        m.alpha[i, : len(d["alpha"])] = d["alpha"]; m.beta[i, : len(d["beta"])] = d["beta"]
    return m


def interior_cell(rng: np.random.Generator, room: tuple, margin: int, x_hi_margin: int = 0) -> tuple:
    """Random padded-grid cell at least `margin` cells from every wall (and `x_hi_margin`
    from the +x wall, to keep clear of a PML there)."""
    nx, ny, nz = room
    i = int(rng.integers(margin, nx - max(margin, x_hi_margin)))
    j = int(rng.integers(margin, ny - margin))
    k = int(rng.integers(margin, nz - margin))
    return (i + GHOST, j + GHOST, k + GHOST)


@dataclass
class Workload:
    name: str
    room: tuple
    h: float
    nsteps: int
    params: SolverParams
    materials: Materials
    rec_ijk: list
    src_ijk: list = field(default_factory=list)
    src_sval: list = field(default_factory=list)
    src_opt: int = _lib.FASES_SRC_RICKER
    u_s: Optional[np.ndarray] = None
    src_freq: Optional[list] = None    # per-source frequency of sine sources [Hz]
    ic: Optional[dict] = None          # {"center": (i,j,k), "sigma": cells, "amp": Pa}
    _scene: Optional[Scene] = None

    @property
    def scene(self) -> Scene:
        if self._scene is None:
            self._scene = Scene.shoebox(*self.room, self.h)
        return self._scene

    @property
    def dims(self) -> tuple:
        """Padded grid dims as the classification built them (room + 6, or + 7 on an axis
        where float rounding of the extent adds a ghost cell, as in the reference)."""
        return self.scene.dims

    @property
    def room_cells(self) -> int:
        return int(np.prod(self.room))

    def cell(self, ijk) -> int:
        Nx, Ny, _ = self.dims
        i, j, k = ijk
        return int(i + Nx * (j + Ny * k))

    def initial_pressure(self) -> Optional[np.ndarray]:
        """Gaussian pressure pulse on air cells (S > 0); None without an initial condition."""
        if self.ic is None:
            return None
        Nx, Ny, Nz = self.dims
        i0, j0, k0 = self.ic["center"]
        sig = float(self.ic["sigma"])
        k, j, i = np.meshgrid(np.arange(Nz), np.arange(Ny), np.arange(Nx), indexing="ij")
        r2 = (i - i0) ** 2 + (j - j0) ** 2 + (k - k0) ** 2
        p = self.ic["amp"] * np.exp(-r2 / (2.0 * sig * sig))
        p = p.ravel()
        p[self.scene.S <= 0] = 0.0
        return p

    def with_precision(self, precision: str) -> "Workload":
        import dataclasses
        prm = dataclasses.replace(self.params, precision=precision)
        return dataclasses.replace(self, params=prm)


def _ricker_source(wl_dt: float, h: float, rho: float, f_peak: float, nsteps: int):
    u_s = source_waveform(_lib.FASES_SRC_RICKER, wl_dt, f_peak, 2.0 * f_peak, nsteps)
    _, peak = source_pref(u_s, nsteps, wl_dt, h, rho)
    return u_s, 1.0 / peak   # 1 Pa at 1 m


def rigid_box(room=(64, 48, 40), h=0.05, nsteps=600, ic_center=(20, 18, 15), rec=(45, 30, 25)) -> Workload:
    """Rigid shoebox with an initial Gaussian pressure pulse (sigma = 2 cells, 1 Pa)."""
    h = f32(h)
    c, rho = 343.0, 1.2
    dt = time_step(0.5, h, c)
    prm = SolverParams(h=h, dt=dt, c=c, rho=rho)
    return Workload(
        name="rigid_box_%dx%dx%d" % room, room=tuple(room), h=h, nsteps=nsteps, params=prm,
        materials=Materials.rigid(6, 1.0), rec_ijk=[tuple(r + GHOST for r in rec)],
        ic={"center": tuple(v + GHOST for v in ic_center), "sigma": 2.0, "amp": 1.0})


def ade_room(room, h, nsteps, seed, n_rec, n_types=None, K=4, L=2, pml_x_hi=True, f_peak=2000.0,
             precision="fp64", pml_cells=16) -> Workload:
    """Shoebox with ADE walls, an optional 16-cell cPML on +x (sigma_max = 3c/dx,
    kappa_max 5, alpha_max 20 pi) whose wall is non-reflecting (R = 0), a Ricker
    point source (1 Pa at 1 m) and seeded receivers."""
    h = f32(h)
    c, rho = 343.0, 1.2
    dt = time_step(0.5, h, c)
    rng = np.random.default_rng(seed)
    walls = 6
    if n_types is None:
        mats = [draw_material(rng, K, L, dt) for _ in range(walls)]
    else:
        types = [draw_material(rng, K, L, dt) for _ in range(n_types)]
        pick = rng.integers(0, n_types, walls)
        mats = [types[int(t)] for t in pick]
    keys = (0, 0, 0, 0, 0, 0)
    if pml_x_hi:
        mats[1] = None          # east wall (x+) is the PML face: R = 0
        keys = (0, 0, 1, 0, 0, 0)
    m = materials_from(mats, K, L)
    prm = SolverParams(h=h, dt=dt, c=c, rho=rho, precision=precision, pml_keys=keys, pml_cells=pml_cells,
                       pml_sigma_max=3.0 * c / h)
    margin = 8
    xm = pml_cells + margin if pml_x_hi else margin
    src = interior_cell(rng, room, margin, xm)
    recs = [interior_cell(rng, room, margin, xm) for _ in range(n_rec)]
    u_s, sval = _ricker_source(dt, h, rho, f_peak, nsteps)
    return Workload(name="ade_room_%dx%dx%d" % tuple(room), room=tuple(room), h=h, nsteps=nsteps, params=prm,
                    materials=m, rec_ijk=recs, src_ijk=[src], src_sval=[sval], u_s=u_s)


def batch_jobs(n: int = 32, seed: int = 1003, h: float = 0.02) -> list:
    """Independent rooms of the batched benchmark: room cells U[96..256] x U[96..192] x
    U[96..160], 1-3 wall materials, a Ricker source with peak U[500, 4000] Hz, U[1500, 4000]
    steps, 2 receivers, FP32 storage, all drawn from one seed."""
    rng = np.random.default_rng(seed)
    jobs = []
    for q in range(n):
        room = (int(rng.integers(96, 257)), int(rng.integers(96, 193)), int(rng.integers(96, 161)))
        n_types = int(rng.integers(1, 4))
        f_peak = float(rng.uniform(500.0, 4000.0))
        steps = int(rng.integers(1500, 4001))
        sub = int(rng.integers(0, 2 ** 31 - 1))
        wl = ade_room(room=room, h=h, nsteps=steps, seed=sub, n_rec=2, n_types=n_types, pml_x_hi=False,
                      f_peak=f_peak, precision="fp32")
        wl.name = "job%02d_%dx%dx%d" % ((q,) + room)
        jobs.append(wl)
    return jobs


def config(n: int, **over) -> Workload:
    """Workload of benchmark configuration n (0..2, 4)."""
    if n == 0:
        return rigid_box(**over)
    if n == 1:
        kw = dict(room=(256, 192, 128), h=0.02, nsteps=4000, seed=1001, n_rec=4)
        kw.update(over)
        return ade_room(**kw)
    if n == 2:
        kw = dict(room=(512, 384, 256), h=0.01, nsteps=8000, seed=1002, n_rec=8, n_types=3)
        kw.update(over)
        return ade_room(**kw)
    if n == 4:
        # the gradient workload: FP64 radius 32 around the source under FP32 storage (the FP32
        # drive gradient of this room is 3-14 % off FP64 with the default 12, 8e-4 with 32)
        kw = dict(room=(192, 160, 128), h=0.02, nsteps=3000, seed=1004, n_rec=8, n_types=3, K=4, L=1,
                  pml_x_hi=False)
        kw.update(over)
        wl = ade_room(**kw)
        return dataclasses.replace(wl, params=dataclasses.replace(wl.params, source_patch=32))
    raise ValueError("unknown configuration %r" % n)


def make_solver(wl: Workload, precision: Optional[str] = None):
    """Solver context loaded with the workload (geometry, materials, sources, receivers, IC)."""
    from .solver import Solver
    import dataclasses
    prm = wl.params if precision is None else dataclasses.replace(wl.params, precision=precision)
    sv = Solver(prm)
    sc = wl.scene
    sv.set_geometry(sc.S, sc.oid, sc.dims)
    sv.set_materials(wl.materials)
    if wl.src_ijk:
        sv.set_sources([wl.cell(s) for s in wl.src_ijk], wl.src_sval, wl.src_opt, wl.u_s, wl.src_freq)
    sv.set_receivers([wl.cell(r) for r in wl.rec_ijk])
    p0 = wl.initial_pressure()
    sv.set_fields(p0)
    return sv
