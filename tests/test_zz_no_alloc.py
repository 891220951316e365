"""fases_get_slices_noalloc (static device staging array, no device allocation) against
fases_get_slices on the same state, called through the C ABI: two calls in a row, the
second on a smaller grid, must give identical planes; a grid whose planes exceed the
static array is rejected before any GPU work."""
import ctypes as C
import sys

import numpy as np
import pytest

from fases_amd import _lib
from fases_amd import workloads as W

pytestmark = pytest.mark.gpu


def announce(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def planes(sv, fn, i0, j0, k0):
    Nx, Ny, Nz = sv.dims
    out = np.full((4, Ny * Nz + Nx * Nz + Nx * Ny), np.nan, np.float32)
    rc = getattr(_lib.load(), fn)(sv._ctx, i0, j0, k0, out.ctypes.data_as(C.POINTER(C.c_float)))
    return rc, out


@pytest.mark.parametrize("prec", ["fp32", "fp64", "fp16"])
def test_noalloc_slices_equal_allocating_slices(prec):
    cases = [((60, 50, 40), 40, (20, 17, 30)), ((30, 26, 22), 25, (9, 12, 7))]     # second grid smaller
    for room, steps, ijk in cases:
        announce("%s room %s, %d steps, planes %s" % (prec, room, steps, ijk))
        wl = W.ade_room(room=room, h=0.02, nsteps=steps, seed=7, n_rec=1, K=2, L=1, pml_cells=5, precision=prec)
        sv = W.make_solver(wl, prec)
        try:
            sv.run(steps)
            bytes0 = sv.info()["dev_bytes"]
            rc1, a = planes(sv, "fases_get_slices_noalloc", *ijk)
            assert rc1 == 0, _lib.load().fases_last_error(sv._ctx)
            assert sv.info()["dev_bytes"] == bytes0          # nothing allocated by the context
            rc2, b = planes(sv, "fases_get_slices", *ijk)
            assert rc2 == 0
            assert np.isfinite(a).all()
            assert np.array_equal(a, b)
            assert np.abs(a[0]).max() > 0
        finally:
            sv.close()


def test_noalloc_slices_reject_planes_above_the_static_buffer():
    """730 x 730 x 7 padded cells: 4 (2 * 730 * 7 + 730^2) floats > 2^21."""
    from fases_amd.solver import Solver, SolverParams
    dims = (730, 730, 7)
    N = int(np.prod(dims))
    announce("noalloc size limit %s" % (dims,))
    sv = Solver(SolverParams(h=0.02, dt=2.9e-5, precision="fp32"))
    try:
        sv.set_geometry(np.ones(N, np.float32), np.full(N, 0xFF, np.uint8), dims)
        rc, out = planes(sv, "fases_get_slices_noalloc", 1, 1, 1)
        assert rc == _lib.FASES_E_ARG
        assert b"static staging buffer" in _lib.load().fases_last_error(sv._ctx)
        assert np.isnan(out).all()
        rc, out = planes(sv, "fases_get_slices", 1, 1, 1)
        assert rc == 0 and np.array_equal(out, np.zeros_like(out))
    finally:
        sv.close()
