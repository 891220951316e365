"""cPML profile pins: the grading of sigma, kappa and alpha across the band
(3D-solver/core/functions3D.cuh:692-719) and the psi-update coefficients b, a
(functions3D.cuh:669-673), for the oracle and for the tables the HIP kernels use.

The expected values are written out at depth fractions x = 1/4, 1/2, 3/4, 1 of a 4-cell
band (x = 1 on the outermost updated cell, padded index 3 or N-4):
    sigma / sigma_max = x^2           0.0625, 0.25, 0.5625, 1
    kappa = 1 + (5 - 1) x^2           1.25,   2,    3.25,   5
    alpha / alpha_max = 0.1 + 0.9 (1 - x)   0.775, 0.55, 0.325, 0.1
A linear grading (m = 1) or an unstretched band (kappa = 1) fails every row but x = 1;
a plain PML (alpha_min = 0) fails x = 1. sigma_max defaults to 0.5 c (m + 1) ln(1/R0) / (d h),
R0 = 1e-7 (functions3D.cuh:691,705)."""
import math

import numpy as np
import pytest

from fases_amd.solver import SolverParams
from oracle import oracle as O

D = 4
H = 0.02
C0 = 343.0
DT = 0.5 * H / C0
ALPHA_MAX = 20.0 * math.pi
SIGMA_MAX = 0.5 * C0 * 3.0 * math.log(1e7) / (D * H)
# depth fraction -> (sigma / sigma_max, kappa, alpha / alpha_max)
EXPECTED = {0.25: (0.0625, 1.25, 0.775), 0.5: (0.25, 2.0, 0.55), 0.75: (0.5625, 3.25, 0.325),
            1.0: (1.0, 5.0, 0.1)}
DIMS = (20, 17, 15)


def params(keys):
    return SolverParams(h=H, dt=DT, c=C0, rho=1.2, pml_keys=keys, pml_cells=D, pml_kappa_max=5.0,
                        pml_alpha_max=ALPHA_MAX)


def expected_axis(n, lo, hi):
    """{padded index: (sigma, kappa, alpha)} of the enabled bands of an n-cell axis."""
    out = {}
    for q in range(D):                 # q cells in from the outermost updated cell
        x = (D - q) / D
        s, k, a = EXPECTED[x]
        if lo:
            out[3 + q] = (s * SIGMA_MAX, k, a * ALPHA_MAX)
        if hi:
            out[n - 4 - q] = (s * SIGMA_MAX, k, a * ALPHA_MAX)
    return out


def sigma_alpha_from_coefficients(kappa, b, a):
    """Inverts b = exp(-(alpha + sigma/kappa) dt), a = sigma/(kappa (sigma + kappa alpha)) (b - 1):
    with g = -ln(b)/dt = alpha + sigma/kappa, sigma + kappa alpha = kappa g, so
    sigma = a kappa^2 g / (b - 1) and alpha = g - sigma/kappa."""
    g = -np.log(b) / DT
    sig = a * kappa ** 2 * g / (b - 1.0)
    return sig, g - sig / kappa


# axis -> (key index of the low face, of the high face) in U D R L F B
AXIS_KEYS = {0: (3, 2), 1: (5, 4), 2: (1, 0)}
CASES = [(1, 0, 0, 0, 0, 0), (0, 0, 1, 0, 0, 0), (0, 1, 0, 1, 0, 1), (1, 1, 1, 1, 1, 1), (0, 0, 0, 0, 1, 0)]


def check_axis(n, lo, hi, in_band, sig, kap, alp):
    exp = expected_axis(n, lo, hi)
    assert sorted(np.flatnonzero(in_band).tolist()) == sorted(exp)
    for i, (s, k, a) in exp.items():
        assert sig[i] == pytest.approx(s, rel=1e-12, abs=1e-9), (i, sig[i], s)
        assert kap[i] == pytest.approx(k, rel=1e-12), (i, kap[i], k)
        assert alp[i] == pytest.approx(a, rel=1e-9), (i, alp[i], a)


@pytest.mark.parametrize("keys", CASES)
def test_oracle_pml_profile_matches_reference_grading(keys):
    prm = params(keys)
    for axis, (klo, khi) in AXIS_KEYS.items():
        inb, prof = O.pml_profile(DIMS, prm, axis)
        sig, kap, alp, b, a = prof.T
        check_axis(DIMS[axis], keys[klo] > 0, keys[khi] > 0, inb, sig, kap, alp)
        # the psi coefficients carry the same profile
        idx = np.flatnonzero(inb)
        if idx.size:
            s2, a2 = sigma_alpha_from_coefficients(kap[idx], b[idx], a[idx])
            np.testing.assert_allclose(s2, sig[idx], rtol=1e-6)
            np.testing.assert_allclose(a2, alp[idx], rtol=1e-6)


def test_profile_pin_rejects_linear_grading_and_unstretched_band():
    """The checker itself: a linear sigma/kappa grading, kappa = 1 and alpha_min = 0 all fail."""
    n = DIMS[0]
    exp = expected_axis(n, True, False)
    idx = sorted(exp)
    x = np.array([(3 + D - i) / D for i in idx])
    good = (SIGMA_MAX * x ** 2, 1 + 4 * x ** 2, ALPHA_MAX * (0.1 + 0.9 * (1 - x)))
    planted = {
        "m = 1": (SIGMA_MAX * x, 1 + 4 * x, good[2]),
        "kappa = 1": (good[0], np.ones_like(x), good[2]),
        "alpha_min = 0": (good[0], good[1], ALPHA_MAX * (1 - x)),
    }
    inb = np.zeros(n, bool)
    inb[idx] = True

    def arr(v):
        out = np.zeros(n)
        out[idx] = v
        return out

    check_axis(n, True, False, inb, *[arr(v) for v in good])
    for name, vals in planted.items():
        with pytest.raises(AssertionError):
            check_axis(n, True, False, inb, *[arr(v) for v in vals])


@pytest.mark.gpu
@pytest.mark.parametrize("keys", CASES)
def test_hip_pml_tables_match_reference_grading(keys):
    """The tables the HIP kernels read (band index, 1/kappa - 1, b, a) carry the same
    profile, and equal the oracle's coefficients."""
    from fases_amd.solver import Scene, Solver
    scene = Scene.shoebox(*(d - 6 for d in DIMS), H)
    dims = scene.dims
    prm = params(keys)
    sv = Solver(prm)
    try:
        sv.set_geometry(scene.S, scene.oid, scene.dims)
        for axis, (klo, khi) in AXIS_KEYS.items():
            t = sv.pml_tables(axis)
            inb = t["band"] >= 0
            kap = np.where(inb, 1.0 / (t["kc"] + 1.0), 1.0)
            sig = np.zeros(dims[axis])
            alp = np.zeros(dims[axis])
            idx = np.flatnonzero(inb)
            if idx.size:
                sig[idx], alp[idx] = sigma_alpha_from_coefficients(kap[idx], t["b"][idx], t["a"][idx])
            # the inversion loses a few digits: compare it at 1e-6, the kappa table exactly
            exp = expected_axis(dims[axis], keys[klo] > 0, keys[khi] > 0)
            assert sorted(idx.tolist()) == sorted(exp)
            for i, (s, k, a) in exp.items():
                assert kap[i] == pytest.approx(k, rel=1e-13)
                assert sig[i] == pytest.approx(s, rel=1e-6)
                assert alp[i] == pytest.approx(a, rel=1e-6)
            # tile index of the band: [0, d) low side, [d, 2d) high side
            for i in idx:
                assert t["band"][i] == (i - 3 if i < 3 + D else i - (dims[axis] - 3 - D) + D)
            _, prof = O.pml_profile(dims, prm, axis)
            np.testing.assert_allclose(t["b"][idx], prof[idx, 3], rtol=1e-14)
            np.testing.assert_allclose(t["a"][idx], prof[idx, 4], rtol=1e-13)
            np.testing.assert_allclose(t["kc"][idx], 1.0 / prof[idx, 1] - 1.0, rtol=1e-14)
    finally:
        sv.close()
