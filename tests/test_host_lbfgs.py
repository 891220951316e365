"""Host logic of the L-BFGS recovery driver (fases_amd/autograd.py) on a mock problem:
a linear "room" whose run overflows (non-finite IR) or fails (FasesError) away from the
start point, the way an ADE wall pushed past |R| = 1 does. No GPU."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from fases_amd import autograd as AG
from fases_amd._lib import FasesError


class LinearRoom:
    """ir = M @ [theta, sval]; the run blows up when a parameter leaves [0.5, 2] x start."""
    K, L = 2, 1

    def __init__(self, mode, seed=0):
        rng = np.random.default_rng(seed)
        self.n = 6 * (1 + 2 * self.K + 4 * self.L)
        self.M = rng.standard_normal((40, self.n + 1))
        self.mode = mode
        self.calls = 0

    def forward(self, theta, sval):
        self.calls += 1
        x = np.concatenate([theta.ravel(), sval])
        if self.mode != "none" and (np.any(np.abs(x) > 2.0) or np.any(np.abs(x) < 0.5)):
            if self.mode == "raise":
                raise FasesError(-7, "mock: non-finite field")
            return np.full((4, 10), np.inf)
        return (self.M @ x).reshape(4, 10)

    def backward(self, dj):
        g = self.M.T @ dj.ravel()
        return g[:-1].reshape(6, -1), g[-1:]


@pytest.mark.parametrize("mode", ["none", "inf", "raise"])
def test_lbfgs_recovers_linear_room_and_survives_blowups(mode):
    p = LinearRoom(mode)
    th0 = np.ones((6, 1 + 2 * p.K + 4 * p.L))
    s0 = np.ones(1)
    mask = AG.free_mask(p.K, p.L, 6)
    truth_th = th0 * np.exp(0.1 * np.sin(np.arange(th0.size)).reshape(th0.shape) * mask)
    truth_s = s0 * 0.9
    tgt = p.forward(truth_th, truth_s)
    th, sv, hist = AG.lbfgs_recover(p, tgt, th0, s0, iterations=25)
    assert np.all(np.isfinite(th)) and np.all(np.isfinite(sv))
    assert all(np.isfinite(h) for h in hist)
    # the misfit of accepted iterates never goes up, and ends far below the start
    assert all(b <= a * (1 + 1e-12) for a, b in zip(hist, hist[1:]))
    assert hist[-1] < 1e-3 * hist[0]


@pytest.mark.parametrize("mode", ["inf", "raise"])
def test_lbfgs_backs_off_from_blowups_toward_an_unreachable_target(mode):
    """The target lies beyond the stable region (every parameter at 3, runs fail above
    1.5): the line search has to bracket failed trials and still make progress."""

    class Capped(LinearRoom):
        def forward(self, theta, sval):
            self.calls += 1
            x = np.concatenate([theta.ravel(), sval])
            if np.any(x > 1.5):
                if mode == "raise":
                    raise FasesError(-7, "mock: non-finite field")
                return np.full((4, 10), np.inf)
            return (self.M @ x).reshape(4, 10)

    p = Capped(mode)
    th0 = np.ones((6, 1 + 2 * p.K + 4 * p.L))
    s0 = np.ones(1)
    tgt = (p.M @ np.full(p.n + 1, 3.0)).reshape(4, 10)
    th, sv, hist = AG.lbfgs_recover(p, tgt, th0, s0, iterations=12)
    assert np.all(np.isfinite(th)) and np.all(np.isfinite(sv))
    assert th.max() <= 1.5 and sv.max() <= 1.5
    assert all(b <= a for a, b in zip(hist, hist[1:]))
    assert hist[-1] < 0.8 * hist[0]
    assert p.calls <= 1 + 12 * 9


def test_lbfgs_does_not_rerun_the_accepted_point():
    """Each step() call starts with a closure call at the point the previous line search
    accepted; that point was already run, so an iteration costs only its trials."""
    p = LinearRoom("none")
    th0 = np.ones((6, 1 + 2 * p.K + 4 * p.L))
    s0 = np.ones(1)
    tgt = p.forward(th0 * 1.05, s0)
    p.calls = 0
    AG.lbfgs_recover(p, tgt, th0, s0, iterations=10, warm=0, max_ls=8)
    # the first point once, then the trials: on a quadratic the unit step is accepted
    # at the first trial (two runs per iteration if the accepted point were run again)
    assert p.calls <= 1 + 10 + 3


def test_lbfgs_iterates_on_a_tiny_misfit():
    """The misfit scale is the caller's (Pa^2 sums of IR differences can be ~1e-10): the
    driver must take the same steps whatever the scale, and not stop on torch's absolute
    gradient / change / curvature thresholds."""
    th0 = np.ones((6, 1 + 2 * LinearRoom.K + 4 * LinearRoom.L))
    s0 = np.ones(1)
    hists = []
    for scale in (1.0, 1e-6):
        p = LinearRoom("none")
        p.M *= scale
        tgt = p.forward(th0 * 1.05, s0 * 0.97)
        th, sv, hist = AG.lbfgs_recover(p, tgt, th0, s0, iterations=15)
        hists.append(np.array(hist) / scale ** 2)
    assert hists[0][-1] < 0.01 * hists[0][0]
    np.testing.assert_allclose(hists[1], hists[0], rtol=1e-6)
