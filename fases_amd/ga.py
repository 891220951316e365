"""Geometrical acoustics of a scene (fases_ga_*): image sources up to an order, the ray
tracer above it, the band impulse responses (HIR/HRT) and arrival lists (ARR) of the
hybrid IR.

Replaces the reference's ism_solve (3D-solver/include/ism3D.h:665), rt_build /
rt_trace_late (3D-solver/include/raytrace3D.h:229, 335) and the GA part of
3D-solver/core/main3D.cu (:1479-1805)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check, load

OCTAVES = (63.0, 125.0, 250.0, 500.0, 1000.0, 2000.0, 4000.0, 8000.0)
BANDS = len(OCTAVES)


def _f32(a) -> Optional[np.ndarray]:
    return None if a is None else np.ascontiguousarray(a, np.float32).ravel()


def _p(a, t=C.c_float):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


@dataclass
class GAParams:
    device: int = 0            # < 0: image sources on the host, no ray tracer
    ism_order: int = 3
    n_rays: int = 10000
    rt_max_order: int = 100000
    recv_radius: float = 0.12
    c: float = 343.0
    fs: float = 44100.0
    Nt: int = 0

    def to_c(self) -> _lib.fases_ga_params:
        p = _lib.fases_ga_params()
        check(load().fases_ga_params_default(C.byref(p)))
        p.device, p.ism_order, p.n_rays, p.rt_max_order = self.device, self.ism_order, self.n_rays, self.rt_max_order
        p.recv_radius, p.c, p.fs, p.Nt = self.recv_radius, self.c, self.fs, self.Nt
        return p


@dataclass
class Image:
    distance: float
    doa: np.ndarray
    amp: float
    group: np.ndarray       # per bounce, source -> receiver
    cos_inc: np.ndarray
    angle: np.ndarray


class GA:
    """Results of one fases_ga_run (freed with the object)."""

    def __init__(self, scene, params: GAParams, src_xyz, rec_xyz, src_pressure=None, alpha=None, sctr=None,
                 m_air=None):
        self._lib = load()
        self.params = params
        src = _f32(src_xyz)
        rec = _f32(rec_xyz)
        self.n_src, self.n_rec = src.size // 3, rec.size // 3
        arrs = [_f32(a) for a in (src_pressure, alpha, sctr, m_air)]
        self._h = C.c_void_p()
        cp = params.to_c()
        check(self._lib.fases_ga_run(scene._h, C.byref(cp), self.n_src, _p(src), _p(arrs[0]), self.n_rec, _p(rec),
                                     _p(arrs[1]), _p(arrs[2]), _p(arrs[3]), C.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.fases_ga_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        o, nn, npth = C.c_int(), C.c_int64(), C.c_int64()
        a, b, k = C.c_double(), C.c_double(), C.c_double()
        check(self._lib.fases_ga_info(self._h, C.byref(o), C.byref(nn), C.byref(npth), C.byref(a), C.byref(b),
                                      C.byref(k)))
        return {"order_used": o.value, "nodes": nn.value, "paths": npth.value, "ism_ms": a.value, "rt_ms": b.value,
                "rt_kernel_ms": k.value}

    def direct(self, s: int, r: int) -> float:
        d = C.c_float()
        check(self._lib.fases_ga_direct(self._h, s, r, C.byref(d)))
        return d.value

    def images(self, s: int, r: int) -> list:
        n, nb = C.c_int(), C.c_int()
        L = self._lib
        check(L.fases_ga_images(self._h, s, r, C.byref(n), C.byref(nb), None, None, None, None, None, None, None))
        k, m = n.value, nb.value
        dist, doa, amp = np.zeros(k, np.float32), np.zeros(3 * k, np.float32), np.zeros(k, np.float32)
        order = np.zeros(k, np.int32)
        grp, cs, ang = np.zeros(m, np.int32), np.zeros(m, np.float32), np.zeros(m, np.float32)
        check(L.fases_ga_images(self._h, s, r, C.byref(n), C.byref(nb), _p(dist), _p(doa), _p(amp),
                                _p(order, C.c_int), _p(grp, C.c_int), _p(cs), _p(ang)))
        out, o = [], 0
        for q in range(k):
            e = o + int(order[q])
            out.append(Image(float(dist[q]), doa[3 * q:3 * q + 3].copy(), float(amp[q]), grp[o:e].copy(), cs[o:e].copy(),
                             ang[o:e].copy()))
            o = e
        return out

    def energy(self):
        """(e_spec, e_diff), each [8, n_rec, Nt]."""
        n = BANDS * self.n_rec * self.params.Nt
        es, ed = np.zeros(n, np.float32), np.zeros(n, np.float32)
        check(self._lib.fases_ga_energy(self._h, _p(es), _p(ed)))
        shp = (BANDS, self.n_rec, self.params.Nt)
        return es.reshape(shp), ed.reshape(shp)

    def band_irs(self, band: int):
        """(p_hf [n_rec, Nt], p_rt [n_rec, Nt]) of octave band `band` (HIR / HRT)."""
        n = self.n_rec * self.params.Nt
        hf, rt = np.zeros(n, np.float32), np.zeros(n, np.float32)
        check(self._lib.fases_ga_band_irs(self._h, band, _p(hf), _p(rt)))
        return hf.reshape(self.n_rec, -1), rt.reshape(self.n_rec, -1)

    def arrivals(self, r: int) -> np.ndarray:
        """[n, 12] arrivals at receiver r: delay, direction [3], gains per band [8] (ARR)."""
        n = C.c_int()
        check(self._lib.fases_ga_arrivals(self._h, r, C.byref(n), None))
        buf = np.zeros(12 * n.value, np.float32)
        check(self._lib.fases_ga_arrivals(self._h, r, C.byref(n), _p(buf)))
        return buf.reshape(-1, 12)


def run(scene, src_xyz, rec_xyz, params: GAParams, src_pressure=None, alpha=None, sctr=None, m_air=None) -> GA:
    return GA(scene, params, src_xyz, rec_xyz, src_pressure, alpha, sctr, m_air)


def band_pass_octaves(f_x: float) -> Sequence[int]:
    """Octave bands the geometrical part writes: those whose upper edge lies above the
    wave-based seam f_x (main3D.cu:1600-1603)."""
    return [b for b, f in enumerate(OCTAVES) if np.float32(f) * np.float32(1.41421356) > np.float32(f_x)]
