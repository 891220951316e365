"""Wall materials of the seeded ADE workloads (256x192x128 seed 1001, 512x384x256 seed 1002,
192x160x128 seed 1004): the recorded .fit fixtures (tests/golden/make_workload_fits.py),
the |R| band the draw targets, and discrete passivity of the ADE the kernel runs.

Passivity follows the reference fitter's check (fminconFIT.py:307-325, discrete_R at
:358-388): the SSP-RK3 ADE's steady-state response |R_d(f)| must stay below 1 up to
1/(2 dt), since a wall with |R_d| > 1 pumps energy at that frequency. R_d comes from
ade_discrete_response (tests/test_oracle.py), which matches the oracle's stage-by-stage
ADE recursion to 1e-10."""
import pathlib

import numpy as np
import pytest

from fases_amd import workloads as W
from fases_amd.solver import read_fit_files

from .test_oracle import ade_discrete_response

FITS = pathlib.Path(__file__).resolve().parent / "golden" / "workload_fits"
CONFIGS = {"shoebox_256x192x128_seed1001": 1, "shoebox_512x384x256_seed1002": 2, "shoebox_192x160x128_seed1004": 4}


def walls(m):
    """(slot, material dict) of the slots that hold a material."""
    out = []
    for i in range(m.n):
        if m.Rinf[i] == 0.0 and not m.A[i].any() and not m.B[i].any():
            continue
        out.append((i, dict(Rinf=m.Rinf[i], A=m.A[i], zeta=m.zeta[i], B=m.B[i], C=m.C[i], alpha=m.alpha[i],
                            beta=m.beta[i])))
    return out


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_recorded_fits_equal_the_seeded_draw(name):
    wl = W.config(CONFIGS[name])
    m = wl.materials
    paths = [str(FITS / name / ("wall%d.fit" % i)) if (FITS / name / ("wall%d.fit" % i)).exists() else None
             for i in range(m.n)]
    assert sum(p is not None for p in paths) == len(walls(m))
    r = read_fit_files(paths, m.K, m.L)
    for key in ("Rinf", "A", "zeta", "B", "C", "alpha", "beta"):
        assert np.array_equal(getattr(r, key), getattr(m, key)), key


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_reflection_band_and_discrete_passivity(name):
    wl = W.config(CONFIGS[name])
    dt = wl.params.dt
    m = wl.materials
    fgrid = np.linspace(0.0, 0.5 / dt, 2048)            # the draw's own grid
    with np.errstate(invalid="ignore", divide="ignore"):      # the empty (cPML) slot divides 0 by 0
        R = np.abs(m.response(fgrid))
    fd = np.logspace(np.log10(5.0), np.log10(0.999 * 0.5 / dt), 1024)
    worst = 0.0
    for i, d in walls(m):
        assert R[i].min() >= 0.5 - 1e-12 and R[i].max() <= 0.95 + 1e-12, (i, R[i].min(), R[i].max())
        gd = np.abs(ade_discrete_response(d, dt, fd))
        worst = max(worst, gd.max())
        assert gd.max() < 1.0, (i, gd.max(), fd[np.argmax(gd)])
    print("%s: max |R_d| up to 1/(2 dt) = %.4f" % (name, worst))
