"""ORACLE -- test infrastructure only. ctypes loader of the FP64 CPU oracle
(oracle/_build/libfases_oracle.so, built from oracle/fvm_oracle.cpp) and of the
reference geometry checker (oracle/_ref/libref_geom.so)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "_build" / "libfases_oracle.so"
REF_SO = HERE / "_ref" / "libref_geom.so"

_P = C.POINTER
_dbl = _P(C.c_double)


class oracle_cfg(C.Structure):
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int),
                ("h", C.c_double), ("dt", C.c_double), ("c", C.c_double), ("rho", C.c_double),
                ("tol", C.c_double), ("pml_keys", C.c_float * 6), ("pml_cells", C.c_int),
                ("pml_sigma_max", C.c_double), ("pml_kappa_max", C.c_double), ("pml_alpha_max", C.c_double)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            raise ImportError(f"{ORACLE_SO} missing: run `python -m fases_amd.build`")
        L = C.CDLL(str(ORACLE_SO))
        L.oracle_run.restype = C.c_int
        L.oracle_run.argtypes = [_P(oracle_cfg), _P(C.c_float), _P(C.c_uint8), C.c_int, C.c_int, C.c_int] + \
            [_dbl] * 7 + [C.c_int, _P(C.c_int64), _dbl, C.c_int, _dbl, _dbl, C.c_int, _P(C.c_int64)] + \
            [_dbl] * 4 + [C.c_int64, C.c_int64] + [_dbl] * 6
        L.oracle_weno5.restype = C.c_double
        L.oracle_weno5.argtypes = [_dbl, C.c_int]
        L.oracle_godunov.restype = None
        L.oracle_godunov.argtypes = [C.c_double] * 6 + [_dbl, _dbl]
        L.oracle_ade_series.restype = None
        L.oracle_ade_series.argtypes = [C.c_int, C.c_int, C.c_double] + [_dbl] * 6 + [C.c_double, _dbl, C.c_int64,
                                                                                    _dbl]
        L.oracle_pml_profile.restype = C.c_int
        L.oracle_pml_profile.argtypes = [_P(oracle_cfg), C.c_int, C.c_int, C.c_int, _dbl]
        _lib = L
    return _lib


def _p(a, t=C.c_double):
    return None if a is None else a.ctypes.data_as(_P(t))


def run(S, oid, dims, params, materials, nsteps, src_cells=(), sval=(), src_opt=1, u_s=None, freq=None,
        rec_cells=(), ic=None, step0=0, want_fields=False, want_vel=False):
    """Runs the oracle; returns dict(ir [n_rec, nsteps], vel, fields (p,u,v,w))."""
    cf = oracle_cfg()
    cf.Nx, cf.Ny, cf.Nz = dims
    cf.h, cf.dt, cf.c, cf.rho = params.h, params.dt, params.c, params.rho
    cf.tol = params.tol if params.tol > 0 else 3.0 * params.h
    for i, k in enumerate(params.pml_keys):
        cf.pml_keys[i] = float(k)
    cf.pml_cells = params.pml_cells
    cf.pml_sigma_max = params.pml_sigma_max
    cf.pml_kappa_max = params.pml_kappa_max
    cf.pml_alpha_max = params.pml_alpha_max
    S = np.ascontiguousarray(S, np.float32)
    oid = np.ascontiguousarray(oid, np.uint8)
    m = materials
    arrs = [np.ascontiguousarray(a, np.float64) for a in (m.Rinf, m.A, m.zeta, m.B, m.C, m.alpha, m.beta)]
    sc = np.ascontiguousarray(src_cells, np.int64)
    sv = np.ascontiguousarray(sval, np.float64)
    us = None if u_s is None else np.ascontiguousarray(u_s, np.float64)
    fr = None if freq is None else np.ascontiguousarray(freq, np.float64)
    rc = np.ascontiguousarray(rec_cells, np.int64)
    N = int(np.prod(dims))
    ics = [None] * 4
    if ic is not None:
        ics = [None if a is None else np.ascontiguousarray(a, np.float64).ravel() for a in ic]
    ir = np.zeros((rc.size, nsteps))
    vel = np.zeros((rc.size, nsteps, 3)) if want_vel else None
    outs = [np.zeros(N) for _ in range(4)] if want_fields else [None] * 4
    r = lib().oracle_run(C.byref(cf), _p(S, C.c_float), _p(oid, C.c_uint8), m.n, m.K, m.L,
                         *[_p(a) for a in arrs], sc.size, _p(sc, C.c_int64), _p(sv), src_opt, _p(us), _p(fr),
                         rc.size, _p(rc, C.c_int64), *[_p(a) for a in ics], step0, nsteps, _p(ir), _p(vel),
                         *[_p(a) for a in outs])
    if r != 0:
        raise RuntimeError("oracle_run failed: %d" % r)
    return {"ir": ir, "vel": vel, "fields": outs if want_fields else None}


def _cfg(dims, params):
    cf = oracle_cfg()
    cf.Nx, cf.Ny, cf.Nz = dims
    cf.h, cf.dt, cf.c, cf.rho = params.h, params.dt, params.c, params.rho
    cf.tol = params.tol if params.tol > 0 else 3.0 * params.h
    for i, k in enumerate(params.pml_keys):
        cf.pml_keys[i] = float(k)
    cf.pml_cells = params.pml_cells
    cf.pml_sigma_max = params.pml_sigma_max
    cf.pml_kappa_max = params.pml_kappa_max
    cf.pml_alpha_max = params.pml_alpha_max
    return cf


def pml_profile(dims, params, axis: int):
    """The oracle's cPML profile along one axis: (in_band [N] bool, prof [N, 5] = sigma,
    kappa, alpha, b, a per padded index)."""
    cf = _cfg(dims, params)
    n = dims[axis]
    inb = np.zeros(n, bool)
    prof = np.zeros((n, 5))
    out = np.zeros(5)
    for i in range(n):
        inb[i] = bool(lib().oracle_pml_profile(C.byref(cf), axis, n, i, _p(out)))
        prof[i] = out
    return inb, prof


def run_workload(wl, nsteps=None, want_fields=False, want_vel=False):
    n = wl.nsteps if nsteps is None else nsteps
    sc = wl.scene
    return run(sc.S, sc.oid, sc.dims, wl.params, wl.materials, n,
               src_cells=[wl.cell(s) for s in wl.src_ijk], sval=wl.src_sval, src_opt=wl.src_opt, u_s=wl.u_s,
               freq=wl.src_freq, rec_cells=[wl.cell(r) for r in wl.rec_ijk], ic=(wl.initial_pressure(), None, None, None),
               want_fields=want_fields, want_vel=want_vel)


def weno5(u5, idx: int) -> float:
    u5 = np.ascontiguousarray(u5, np.float64)
    return lib().oracle_weno5(_p(u5), idx)


def godunov(pM, pP, uM, uP, rho, c):
    a, b = C.c_double(), C.c_double()
    lib().oracle_godunov(pM, pP, uM, uP, rho, c, C.byref(a), C.byref(b))
    return a.value, b.value


def ade_series(Rinf, A, zeta, B, Cc, alpha, beta, dt, w3):
    A = np.ascontiguousarray(A if len(A) else [0.0], np.float64)
    zeta = np.ascontiguousarray(zeta if len(zeta) else [0.0], np.float64)
    K = len(A)
    B = np.ascontiguousarray(B if len(B) else [0.0], np.float64)
    Cc = np.ascontiguousarray(Cc if len(Cc) else [0.0], np.float64)
    alpha = np.ascontiguousarray(alpha if len(alpha) else [0.0], np.float64)
    beta = np.ascontiguousarray(beta if len(beta) else [0.0], np.float64)
    L = len(B)
    w3 = np.ascontiguousarray(w3, np.float64)
    n = w3.size // 3
    out = np.zeros(3 * n)
    lib().oracle_ade_series(K, L, Rinf, _p(A), _p(zeta), _p(B), _p(Cc), _p(alpha), _p(beta), dt, _p(w3), n, _p(out))
    return out.reshape(n, 3)


# ---------------------------------------------------------------------------------
# reference geometry (compiled from the reference's own sources where present)
# ---------------------------------------------------------------------------------

_ref = None


def ref_available() -> bool:
    return REF_SO.exists()


def ref_lib() -> C.CDLL:
    global _ref
    if _ref is None:
        L = C.CDLL(str(REF_SO))
        L.ref_scene.restype = C.c_int
        L.ref_scene.argtypes = [C.c_char_p, C.c_float, _P(C.c_int), _P(_P(C.c_float)), _P(_P(C.c_uint8))]
        L.ref_free.restype = None
        L.ref_free.argtypes = [C.c_void_p]
        L.ref_boundary_map.restype = C.c_int
        L.ref_boundary_map.argtypes = [_P(C.c_uint8), C.c_int, _P(C.c_int)]
        L.ref_read_fit.restype = None
        L.ref_read_fit.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int] + [_P(C.c_float)] * 7
        L.ref_group_areas.restype = C.c_int
        L.ref_group_areas.argtypes = [C.c_char_p, C.c_float, _P(C.c_float), C.c_int, _P(C.c_float)]
        L.ref_ism.restype = C.c_int
        L.ref_ism.argtypes = [C.c_char_p, C.c_float, _P(C.c_float), C.c_int, _P(C.c_float), C.c_int, C.c_int,
                              C.c_int, C.c_int, _P(C.c_float), _P(C.c_int), _P(C.c_float), _P(C.c_float),
                              _P(C.c_float), _P(C.c_int), _P(C.c_int), _P(C.c_float), _P(C.c_float)]
        L.ref_rt.restype = C.c_int
        L.ref_rt.argtypes = [C.c_char_p, C.c_float, _P(C.c_float), C.c_int, _P(C.c_float), _P(C.c_float), C.c_int,
                             _P(C.c_float), _P(C.c_float), _P(C.c_float), C.c_float, C.c_float, C.c_int, C.c_int,
                             C.c_int, C.c_float, C.c_int, _P(C.c_float), _P(C.c_float)]
        L.ref_r_oblique.restype = C.c_float
        L.ref_r_oblique.argtypes = [C.c_float, C.c_float]
        L.ref_rt_sign.restype = C.c_float
        L.ref_rt_sign.argtypes = [C.c_uint]
        L.ref_dump_rays.restype = C.c_int
        L.ref_dump_rays.argtypes = [C.c_char_p, C.c_float, _P(C.c_float), C.c_int, C.c_int, C.c_int, C.c_char_p]
        L.ref_t60.restype = C.c_int
        L.ref_t60.argtypes = [C.c_char_p, C.c_float, C.c_char_p, C.c_float, C.c_float, _P(C.c_float),
                              _P(C.c_float), _P(C.c_float), C.c_int]
        _ref = L
    return _ref


def ref_scene(obj_path: str, h: float):
    L = ref_lib()
    dims = (C.c_int * 3)()
    Sp = _P(C.c_float)()
    Op = _P(C.c_uint8)()
    r = L.ref_scene(obj_path.encode(), h, dims, C.byref(Sp), C.byref(Op))
    if r != 0:
        raise RuntimeError("reference scene build failed: %d" % r)
    d = tuple(dims)
    N = int(np.prod(d))
    S = np.ctypeslib.as_array(Sp, shape=(N,)).copy()
    oid = np.ctypeslib.as_array(Op, shape=(N,)).copy()
    L.ref_free(C.cast(Sp, C.c_void_p))
    L.ref_free(C.cast(Op, C.c_void_p))
    return d, S, oid


def ref_boundary_map(oid):
    oid = np.ascontiguousarray(oid, np.uint8).copy()
    bmap = np.zeros(oid.size, np.int32)
    nB = ref_lib().ref_boundary_map(_p(oid, C.c_uint8), oid.size, _p(bmap, C.c_int))
    return bmap, nB


def ref_read_fit(path, m, K, L, nMat):
    arrs = [np.zeros(nMat, np.float32)] + [np.zeros(nMat * K, np.float32)] * 0
    R = np.zeros(nMat, np.float32)
    A = np.zeros(nMat * K, np.float32); Z = np.zeros(nMat * K, np.float32)
    B = np.zeros(nMat * L, np.float32); Cc = np.zeros(nMat * L, np.float32)
    al = np.zeros(nMat * L, np.float32); be = np.zeros(nMat * L, np.float32)
    ref_lib().ref_read_fit(path.encode(), m, K, L, *[_p(a, C.c_float) for a in (R, A, Z, B, Cc, al, be)])
    del arrs
    return R, A, Z, B, Cc, al, be


def ref_group_areas(obj_path: str, h: float):
    """The reference's mesh_group_areas: (areas per group, total) of the placed mesh."""
    A = np.zeros(256, np.float32)
    tot = C.c_float()
    n = ref_lib().ref_group_areas(obj_path.encode(), h, _p(A, C.c_float), A.size, C.byref(tot))
    if n < 0:
        raise RuntimeError("reference OBJ parse failed")
    return A[:n].copy(), np.float32(tot.value)


def ref_t60(obj_path: str, h: float, mat_dir: str, V: float, c: float, m_air):
    """The reference's estimate_T60s on its own areas and .mat loading: (T60[8], alpha[n, 8])."""
    m = np.ascontiguousarray(m_air, np.float32)
    T = np.zeros(8, np.float32)
    al = np.zeros(256 * 8, np.float32)
    n = ref_lib().ref_t60(obj_path.encode(), h, mat_dir.encode(), V, c, _p(m, C.c_float), _p(T, C.c_float),
                          _p(al, C.c_float), 256)
    if n < 0:
        raise RuntimeError("reference OBJ parse failed")
    return T, al[:8 * n].reshape(n, 8).copy()



def ref_ism(obj_path: str, h: float, src, rec, order: int, cap_img: int = 200000, cap_b: int = 2000000):
    """The reference's ism_solve: (order_used, direct [ns, nr], images[s][r] = list of
    (distance, doa[3], amp, groups, cos, angles))."""
    src = np.ascontiguousarray(src, np.float32).ravel()
    rec = np.ascontiguousarray(rec, np.float32).ravel()
    ns, nr = src.size // 3, rec.size // 3
    direct = np.zeros(ns * nr, np.float32)
    n_img = np.zeros(ns * nr, np.int32)
    dist, amp = np.zeros(cap_img, np.float32), np.zeros(cap_img, np.float32)
    doa = np.zeros(3 * cap_img, np.float32)
    order_out = np.zeros(cap_img, np.int32)
    grp, cs, ang = np.zeros(cap_b, np.int32), np.zeros(cap_b, np.float32), np.zeros(cap_b, np.float32)
    f = lambda a: _p(a, C.c_float)  # noqa: E731
    i = lambda a: _p(a, C.c_int)  # noqa: E731
    used = ref_lib().ref_ism(obj_path.encode(), h, f(src), ns, f(rec), nr, order, cap_img, cap_b, f(direct), i(n_img),
                             f(dist), f(doa), f(amp), i(order_out), i(grp), f(cs), f(ang))
    if used < 0:
        raise RuntimeError("reference ism_solve failed: %d" % used)
    images, q, b = [], 0, 0
    for s in range(ns):
        row = []
        for r in range(nr):
            lst = []
            for _ in range(int(n_img[s * nr + r])):
                m = int(order_out[q])
                lst.append((float(dist[q]), doa[3 * q:3 * q + 3].copy(), float(amp[q]), grp[b:b + m].copy(),
                            cs[b:b + m].copy(), ang[b:b + m].copy()))
                q += 1
                b += m
            row.append(lst)
        images.append(row)
    return used, direct.reshape(ns, nr), images


def ref_rt(obj_path: str, h: float, src, src_p, rec, alpha_gb, m_air, sctr_g, c, fs, Nt, n_rays, max_order, radius,
           ism_order):
    """The reference's rt_build + rt_trace_late: (e_spec, e_diff), each [8, nr, Nt]."""
    src = np.ascontiguousarray(src, np.float32).ravel()
    rec = np.ascontiguousarray(rec, np.float32).ravel()
    ns, nr = src.size // 3, rec.size // 3
    arr = [np.ascontiguousarray(a, np.float32).ravel() for a in (src_p, alpha_gb, m_air, sctr_g)]
    es, ed = np.zeros(8 * nr * Nt, np.float32), np.zeros(8 * nr * Nt, np.float32)
    f = lambda a: _p(a, C.c_float)  # noqa: E731
    rc = ref_lib().ref_rt(obj_path.encode(), h, f(src), ns, f(arr[0]), f(rec), nr, f(arr[1]), f(arr[2]), f(arr[3]), c,
                          fs, Nt, n_rays, max_order, radius, ism_order, f(es), f(ed))
    if rc != 0:
        raise RuntimeError("reference rt failed: %d" % rc)
    return es.reshape(8, nr, Nt), ed.reshape(8, nr, Nt)


def ref_dump_rays(obj_path: str, h: float, src, n_rays: int, max_order: int, out_path: str) -> None:
    """The reference's rt_dump_paths (--dump-rays) written to out_path."""
    src = np.ascontiguousarray(src, np.float32).ravel()
    rc = ref_lib().ref_dump_rays(obj_path.encode(), h, _p(src, C.c_float), src.size // 3, n_rays, max_order,
                                 out_path.encode())
    if rc != 0:
        raise RuntimeError("reference dump_rays failed: %d" % rc)
