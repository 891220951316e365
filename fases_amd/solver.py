"""Python host API over the C ABI: scenes, project formats and the HIP solver context.

Mirrors the reference's solver inputs (3D-solver/core/main3D.cu): the padded grid
with signed distance S and material id oid (parsing3D.h), ADE materials laid out
[m*K + k] (main3D.cu:575-668, filterABC.h:31-92), point sources with a shared
waveform (main3D.cu:853-1124) and receiver cells (main3D.cu:919-953).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check, load

GHOST = 3  # padding cells per side (parsing3D.h:128)


def _ptr(a: Optional[np.ndarray], ctype):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


# ----------------------------------------------------------------------------------
# Formats
# ----------------------------------------------------------------------------------

def read_config(path: str) -> np.ndarray:
    """config.txt -> float32 values, first float of each line (main3D.cu:114-128)."""
    lib = load()
    n = C.c_int(0)
    check(lib.fases_read_config(path.encode(), None, 0, C.byref(n)))
    out = np.zeros(max(n.value, 1), np.float32)
    check(lib.fases_read_config(path.encode(), _ptr(out, C.c_float), out.size, C.byref(n)))
    return out[: n.value]


def read_poles(path: str) -> tuple[int, int]:
    lib = load()
    K, L = C.c_int(0), C.c_int(0)
    check(lib.fases_read_poles(path.encode(), C.byref(K), C.byref(L)))
    return K.value, L.value


@dataclass
class Materials:
    """ADE admittance per material (vector fit): R(s) = R_inf + sum_k A_k/(s + zeta_k)
    + sum_l (B_l (s + alpha_l) + C_l beta_l)/((s + alpha_l)^2 + beta_l^2)."""
    Rinf: np.ndarray
    A: np.ndarray      # [nMat, K]
    zeta: np.ndarray
    B: np.ndarray      # [nMat, L]
    C: np.ndarray
    alpha: np.ndarray
    beta: np.ndarray
    names: list = field(default_factory=list)

    @property
    def n(self) -> int:
        return int(self.Rinf.shape[0])

    @property
    def K(self) -> int:
        return int(self.A.shape[1])

    @property
    def L(self) -> int:
        return int(self.B.shape[1])

    @staticmethod
    def empty(n: int, K: int, L: int) -> "Materials":
        z = lambda k: np.zeros((n, k))
        return Materials(np.zeros(n), z(K), z(K), z(L), z(L), z(L), z(L))

    @staticmethod
    def rigid(n: int, Rinf: float = 1.0) -> "Materials":
        m = Materials.empty(n, 1, 1)
        m.Rinf[:] = Rinf
        return m

    def response(self, f: np.ndarray) -> np.ndarray:
        """Continuous R(j 2 pi f) per material, shape [nMat, len(f)]."""
        s = 2j * np.pi * np.asarray(f, float)[None, :]
        R = self.Rinf[:, None] + 0j * s
        for k in range(self.K):
            R = R + self.A[:, k, None] / (s + self.zeta[:, k, None])
        for l in range(self.L):
            a, b = self.alpha[:, l, None], self.beta[:, l, None]
            R = R + (self.B[:, l, None] * (s + a) + self.C[:, l, None] * b) / ((s + a) ** 2 + b ** 2)
        return R


def read_fit_files(paths: Sequence[Optional[str]], K: int, L: int) -> Materials:
    """One .fit per material slot (None = skipped slot with R_inf = 0, as the reference
    does for BoundingBox layers, main3D.cu:624-634)."""
    lib = load()
    n = len(paths)
    m = Materials.empty(n, K, L)
    arrs = [m.Rinf, m.A, m.zeta, m.B, m.C, m.alpha, m.beta]
    for i, p in enumerate(paths):
        if p is None:
            continue
        check(lib.fases_read_fit(p.encode(), i, K, L, *[_ptr(a, C.c_double) for a in arrs]))
    return m


def write_fit(path: str, Rinf: float, A, zeta, B, C_, alpha, beta) -> None:
    """Writes a .fit file in the section grammar read_fit_file parses."""
    with open(path, "w") as f:
        f.write("# R_inf\n%.17g\n# K L\n%d %d\n# A_k zeta_k\n" % (Rinf, len(A), len(B)))
        for a, z in zip(A, zeta):
            f.write("%.17g %.17g\n" % (a, z))
        f.write("# B_l C_l alpha_l beta_l\n")
        for b, c, al, be in zip(B, C_, alpha, beta):
            f.write("%.17g %.17g %.17g %.17g\n" % (b, c, al, be))


def read_sources(path: str, h: float, dims, default_f0: float):
    lib = load()
    n = C.c_int(0)
    check(lib.fases_read_sources(path.encode(), h, *dims, default_f0, 0, None, None, None, None, C.byref(n)))
    cnt = max(n.value, 1)
    cell = np.zeros(cnt, np.int64); lp = np.zeros(cnt, np.float32); f0 = np.zeros(cnt, np.float32)
    xyz = np.zeros(3 * cnt, np.float32)
    check(lib.fases_read_sources(path.encode(), h, *dims, default_f0, cnt, _ptr(cell, C.c_int64),
                                 _ptr(lp, C.c_float), _ptr(f0, C.c_float), _ptr(xyz, C.c_float), C.byref(n)))
    k = n.value
    return cell[:k], lp[:k], f0[:k], xyz[: 3 * k].reshape(-1, 3)


def read_receivers(path: str, h: float, dims):
    lib = load()
    n = C.c_int(0)
    check(lib.fases_read_receivers(path.encode(), h, *dims, 0, None, None, C.byref(n)))
    cnt = max(n.value, 1)
    cell = np.zeros(cnt, np.int64); xyz = np.zeros(3 * cnt, np.float32)
    check(lib.fases_read_receivers(path.encode(), h, *dims, cnt, _ptr(cell, C.c_int64), _ptr(xyz, C.c_float),
                                   C.byref(n)))
    k = n.value
    return cell[:k], xyz[: 3 * k].reshape(-1, 3)


def source_waveform(src_opt: int, dt: float, f: float, f_cutoff: float, Nt: int) -> np.ndarray:
    """u_s[0..2Nt] at dt/2 (PCS or Ricker, main3D.cu:963-1058)."""
    u = np.zeros(2 * Nt + 1)
    check(load().fases_source_waveform(src_opt, dt, f, f_cutoff, Nt, _ptr(u, C.c_double)))
    return u


def source_pref(u_s: np.ndarray, Nt: int, dt: float, h: float, rho: float):
    """(p_ref[0..2Nt], peak) of the 1 m free-field pressure per unit drive (main3D.cu:1079-1094)."""
    u_s = _f64(u_s)
    pr = np.zeros(2 * Nt + 1)
    peak = C.c_double(0)
    check(load().fases_source_pref(_ptr(u_s, C.c_double), Nt, dt, h, rho, _ptr(pr, C.c_double), C.byref(peak)))
    return pr, peak.value


def write_f32(path: str, x: np.ndarray) -> None:
    x = _f64(x).ravel()
    check(load().fases_write_f32(path.encode(), _ptr(x, C.c_double), x.size))


def hbm_copy_gbs(device: int = 0, nbytes: int = 2 << 30, reps: int = 5) -> float:
    """Read + write bandwidth [GB/s] of a streaming copy of nbytes on the device
    (fases_hbm_copy): the measured HBM reference for the roofline figures."""
    g = C.c_double()
    check(load().fases_hbm_copy(device, nbytes, reps, C.byref(g)))
    return g.value


def cell_index(i: int, j: int, k: int, dims) -> int:
    Nx, Ny, _ = dims
    return int(i + Nx * (j + Ny * k))


# ----------------------------------------------------------------------------------
# Scene
# ----------------------------------------------------------------------------------

class Scene:
    """Padded grid + signed distance S (float32) + material id oid (uint8)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        lib = load()
        nx, ny, nz, ng = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        L3 = (C.c_float * 3)()
        check(lib.fases_scene_info(handle, C.byref(nx), C.byref(ny), C.byref(nz), C.byref(ng), L3))
        self.dims = (nx.value, ny.value, nz.value)
        self.extent = tuple(float(v) for v in L3)
        buf = C.create_string_buffer(256)
        self.group_names = []
        for g in range(ng.value):
            check(lib.fases_scene_group_name(handle, g, buf, 256))
            self.group_names.append(buf.value.decode())
        N = int(np.prod(self.dims))
        self.S = np.zeros(N, np.float32)
        self.oid = np.zeros(N, np.uint8)
        check(lib.fases_scene_get(handle, _ptr(self.S, C.c_float), _ptr(self.oid, C.c_uint8)))

    @classmethod
    def from_obj(cls, path: str, h: float, device: int | None = None) -> "Scene":
        """device None: classified on the host; an index: on that GPU (same S and oid)."""
        hd = C.c_void_p()
        dev = -1 if device is None else int(device)
        check(load().fases_scene_load_obj_dev(path.encode(), h, dev, C.byref(hd)))
        return cls(hd)

    @classmethod
    def shoebox(cls, nx: int, ny: int, nz: int, h: float, device: int | None = None) -> "Scene":
        hd = C.c_void_p()
        dev = -1 if device is None else int(device)
        check(load().fases_scene_shoebox_dev(nx, ny, nz, h, dev, C.byref(hd)))
        return cls(hd)

    @property
    def n_cells(self) -> int:
        return int(np.prod(self.dims))

    def __del__(self):
        try:
            if self._h:
                load().fases_scene_free(self._h)
                self._h = None
        except Exception:
            pass


def boundary_map(oid: np.ndarray):
    oid = np.ascontiguousarray(oid, np.uint8)
    bmap = np.zeros(oid.size, np.int32)
    nB = C.c_int64(0)
    check(load().fases_boundary_map(_ptr(oid, C.c_uint8), oid.size, _ptr(bmap, C.c_int32), C.byref(nB)))
    return bmap, nB.value


# ----------------------------------------------------------------------------------
# Solver context
# ----------------------------------------------------------------------------------

@dataclass
class SolverParams:
    h: float
    dt: float
    c: float = 343.0
    rho: float = 1.225
    precision: str = "fp64"
    device: int = 0
    tol: float = 0.0                    # 0 -> 3h
    pml_keys: tuple = (0, 0, 0, 0, 0, 0)  # U D R L F B
    pml_cells: int = 20
    pml_sigma_max: float = 0.0          # 0 -> reference profile
    pml_kappa_max: float = 5.0
    pml_alpha_max: float = float(np.pi * 20.0)
    fp16_rescale_every: int = 100
    use_graph: bool = False             # reserved, ignored (fases_run launches stage by stage)
    kernel_variant: int = 0             # 0 auto, 1 per-cell (reference order), 2 tiled
    tile_z: int = 0                     # tiled kernel z planes per block; 0 = autotune
    adjoint_variant: int = 0            # flux adjoint: 0 tiled, 1 per cell, 2 tiled in XCD-aware order
    tile_rows: int = 0                  # tiled kernel waves per block (8 or 16); 0 = autotune
    source_patch: int = 0               # FP64 radius around sources under FP32/FP16 storage; 0 = 12
    tile_cols: int = 0                  # tiled kernel columns per block (64 or 62); 0 = autotune

    def to_c(self) -> _lib.fases_params:
        p = _lib.fases_params()
        check(load().fases_params_default(C.byref(p)))
        p.device = self.device
        p.precision = _lib.PRECISIONS[self.precision]
        p.h, p.dt, p.c, p.rho, p.tol = self.h, self.dt, self.c, self.rho, self.tol
        for i, k in enumerate(self.pml_keys):
            p.pml_keys[i] = float(k)
        p.pml_cells = self.pml_cells
        p.pml_sigma_max = self.pml_sigma_max
        p.pml_kappa_max = self.pml_kappa_max
        p.pml_alpha_max = self.pml_alpha_max
        p.fp16_rescale_every = self.fp16_rescale_every
        p.use_graph = int(self.use_graph)
        p.kernel_variant = int(self.kernel_variant)
        p.tile_z = int(self.tile_z)
        p.adjoint_variant = int(self.adjoint_variant)
        p.tile_rows = int(self.tile_rows)
        p.source_patch = int(self.source_patch)
        p.tile_cols = int(self.tile_cols)
        return p


class Solver:
    """One HIP solver context (one device, one stream)."""

    def __init__(self, params: SolverParams):
        self.params = params
        self._lib = load()
        self._ctx = C.c_void_p()
        cp = params.to_c()
        check(self._lib.fases_create(C.byref(cp), C.byref(self._ctx)))
        self.dims = None
        self.n_rec = 0

    def _chk(self, rc):
        check(rc, self._ctx)

    def close(self):
        if self._ctx:
            self._lib.fases_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_geometry(self, S: np.ndarray, oid: np.ndarray, dims) -> None:
        S = np.ascontiguousarray(S, np.float32).ravel()
        oid = np.ascontiguousarray(oid, np.uint8).ravel()
        self._chk(self._lib.fases_set_geometry(self._ctx, _ptr(S, C.c_float), _ptr(oid, C.c_uint8), *dims))
        self.dims = tuple(dims)

    def set_materials(self, m: Materials) -> None:
        arrs = [_f64(a) for a in (m.Rinf, m.A, m.zeta, m.B, m.C, m.alpha, m.beta)]
        self._chk(self._lib.fases_set_materials(self._ctx, m.n, m.K, m.L, *[_ptr(a, C.c_double) for a in arrs]))
        self.mat_shape = (m.n, m.K, m.L)

    def set_sources(self, cells, sval, src_opt: int = _lib.FASES_SRC_RICKER, u_s=None, freq=None) -> None:
        cells = np.ascontiguousarray(cells, np.int64)
        sval = _f64(sval)
        u = None if u_s is None else _f64(u_s)
        fr = None if freq is None else _f64(freq)
        self._chk(self._lib.fases_set_sources(self._ctx, cells.size, _ptr(cells, C.c_int64), _ptr(sval, C.c_double),
                                              src_opt, _ptr(u, C.c_double), 0 if u is None else u.size,
                                              _ptr(fr, C.c_double)))
        self.n_src = len(set(int(c) for c in cells)) if src_opt != _lib.FASES_SRC_SINE else cells.size
        self.n_samp = 0 if u is None else u.size

    def set_receivers(self, cells) -> None:
        cells = np.ascontiguousarray(cells, np.int64)
        self._chk(self._lib.fases_set_receivers(self._ctx, cells.size, _ptr(cells, C.c_int64)))
        self.n_rec = cells.size

    def set_fields(self, p=None, u=None, v=None, w=None) -> None:
        arrs = [None if a is None else _f64(a).ravel() for a in (p, u, v, w)]
        self._chk(self._lib.fases_set_fields(self._ctx, *[_ptr(a, C.c_double) for a in arrs]))

    def get_fields(self):
        N = int(np.prod(self.dims))
        out = [np.zeros(N) for _ in range(4)]
        self._chk(self._lib.fases_get_fields(self._ctx, *[_ptr(a, C.c_double) for a in out]))
        return out

    def get_slices(self, i0: int, j0: int, k0: int):
        """Animation planes of the current p, u, v, w: a [4, Ny*Nz + Nx*Nz + Nx*Ny] float32
        array, each row [YZ at x = i0 | XZ at y = j0 | XY at z = k0] (fases_get_slices)."""
        Nx, Ny, Nz = self.dims
        out = np.zeros((4, Ny * Nz + Nx * Nz + Nx * Ny), np.float32)
        self._chk(self._lib.fases_get_slices(self._ctx, int(i0), int(j0), int(k0), _ptr(out, C.c_float)))
        return out

    def reset(self) -> None:
        self._chk(self._lib.fases_reset(self._ctx))

    def run(self, nsteps: int, want_vel: bool = False, kernel_timing: bool = False):
        """Returns (ir [n_rec, nsteps], vel [n_rec, nsteps, 3] or None, elapsed_ms, kernel_ms)."""
        ir = np.zeros((self.n_rec, nsteps))
        vel = np.zeros((self.n_rec, nsteps, 3)) if want_vel else None
        el = C.c_double(0)
        km = C.c_double(0)
        self._chk(self._lib.fases_run(self._ctx, nsteps, _ptr(ir, C.c_double), _ptr(vel, C.c_double), C.byref(el),
                                      C.byref(km) if kernel_timing else None))
        return ir, vel, el.value, (km.value if kernel_timing else None)

    def info(self) -> dict:
        a, b, c, d = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        self._chk(self._lib.fases_info(self._ctx, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return {"n_cells": a.value, "n_bnd": b.value, "n_upd": c.value, "dev_bytes": d.value}

    def adjoint_forward(self, nsteps: int, checkpoint_every: int = 64) -> np.ndarray:
        """Forward run storing checkpoints for adjoint(); returns ir [n_rec, nsteps]."""
        ir = np.zeros((self.n_rec, nsteps))
        self._chk(self._lib.fases_adjoint_forward(self._ctx, nsteps, checkpoint_every, _ptr(ir, C.c_double)))
        self._adj_steps = nsteps
        return ir

    def adjoint(self, dj_dir: np.ndarray, want_us: bool = False):
        """Reverse sweep for dJ/d ir; returns (grad_mat [nMat, 1+2K+4L], grad_sval [n_src],
        grad_us [n_samp] or None). Gradient layout per material: R_inf, A[K], zeta[K], B[L],
        C[L], alpha[L], beta[L]. The field state is undefined afterwards."""
        steps = getattr(self, "_adj_steps", None)     # None: the library reports the missing forward
        dj = _f64(dj_dir).reshape(self.n_rec, -1 if steps is None else steps)
        n, K, L = self.mat_shape
        gm = np.zeros((n, 1 + 2 * K + 4 * L))
        gs = np.zeros(max(getattr(self, "n_src", 0), 1))
        gu = np.zeros(self.n_samp) if (want_us and getattr(self, "n_samp", 0)) else None
        self._chk(self._lib.fases_adjoint(self._ctx, _ptr(dj, C.c_double), _ptr(gm, C.c_double), _ptr(gs, C.c_double),
                                          _ptr(gu, C.c_double)))
        return gm, gs[: getattr(self, "n_src", 0)], gu

    def pml_tables(self, axis: int) -> dict:
        """cPML tables of axis 0 (x), 1 (y), 2 (z) as the kernels use them: band index (-1
        outside), kc = 1/kappa - 1, b and a of the psi update, one entry per padded cell."""
        n = self.dims[axis]
        band = np.zeros(n, np.int32)
        kc, b, a = np.zeros(n), np.zeros(n), np.zeros(n)
        self._chk(self._lib.fases_pml_tables(self._ctx, axis, _ptr(band, C.c_int32), _ptr(kc, C.c_double),
                                             _ptr(b, C.c_double), _ptr(a, C.c_double)))
        return {"band": band, "kc": kc, "b": b, "a": a}

    def tile_config(self) -> tuple:
        """(columns, rows, z planes) per block of the tiled flux kernel (z = 0 before tuning)."""
        x, y, z = C.c_int(), C.c_int(), C.c_int()
        self._chk(self._lib.fases_tile_config(self._ctx, C.byref(x), C.byref(y), C.byref(z)))
        return x.value, y.value, z.value


def batch_run(solvers: Sequence[Solver], nsteps: Sequence[int]):
    """Advances the solvers together (fases_batch_run); returns ([ir_j [n_rec_j, nsteps_j]], elapsed_ms)."""
    n = len(solvers)
    ctxs = (C.c_void_p * n)(*[s._ctx.value for s in solvers])
    ns = np.ascontiguousarray(nsteps, np.int64)
    irs = [np.zeros((s.n_rec, int(k))) for s, k in zip(solvers, ns)]
    ptrs = (C.POINTER(C.c_double) * n)(*[_ptr(a, C.c_double) for a in irs])
    el = C.c_double(0)
    rc = load().fases_batch_run(ctxs, n, _ptr(ns, C.c_int64), ptrs, C.byref(el))
    check(rc, solvers[0]._ctx)
    return irs, el.value
