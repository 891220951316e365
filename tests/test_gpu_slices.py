"""fases_get_slices (animation planes) against fases_get_fields of the same context, for
FP64, FP32 and FP16 storage, and its range check."""
import numpy as np
import pytest

from fases_amd import workloads as W
from fases_amd._lib import FasesError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp16"])
def test_slices_are_the_field_planes(precision):
    wl = W.rigid_box(room=(30, 22, 18), h=0.05, nsteps=40, ic_center=(12, 9, 7), rec=(20, 15, 10))
    print("[slices]", precision, flush=True)
    sv = W.make_solver(wl, precision)
    sv.run(25)
    Nx, Ny, Nz = sv.dims
    i0, j0, k0 = 14, 11, 9
    sl = sv.get_slices(i0, j0, k0)
    fields = [np.asarray(f).reshape(Nz, Ny, Nx) for f in sv.get_fields()]
    sv.close()
    nyz, nxz = Ny * Nz, Nx * Nz
    for q, f in enumerate(fields):
        # the stored value times its power-of-two scale, in float32 (exact for FP32 and
        # FP16 storage, one rounding for FP64)
        f32 = f.astype(np.float32)
        assert np.array_equal(sl[q, :nyz], f32[:, :, i0].ravel())
        assert np.array_equal(sl[q, nyz:nyz + nxz], f32[:, j0, :].ravel())
        assert np.array_equal(sl[q, nyz + nxz:], f32[k0, :, :].ravel())
    assert np.abs(sl[0]).max() > 0


def test_slice_plane_outside_grid_is_an_error():
    wl = W.rigid_box(room=(16, 12, 10), h=0.05, nsteps=4, ic_center=(8, 6, 5), rec=(10, 8, 6))
    sv = W.make_solver(wl, "fp32")
    Nx, Ny, Nz = sv.dims
    with pytest.raises(FasesError):
        sv.get_slices(Nx, 0, 0)
    with pytest.raises(FasesError):
        sv.get_slices(0, -1, 0)
    sv.close()
