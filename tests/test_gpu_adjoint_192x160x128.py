"""The differentiation room at full size: 192x160x128 cells, dx 0.02 m, six walls of three
material types (4 real poles + 1 complex pair each, seed 1004), a Ricker source, 8
receivers, 3000 steps, checkpoints every 64 steps. Target IRs come from the seeded
(held-out) materials; the gradient of J = sum (ir - target)^2 is taken at the recovery
problem's start point: every R_inf, residue and pole real part scaled by a log-uniform
factor in [1/1.2, 1.2] (seed 1004), the drive 20 % high.

Checks: the FP64 forward is deterministic (J twice, bit-equal); the FP64 adjoint against
central differences of the FP64 forward (the oracle's arithmetic bit for bit) with the
relative step 1e-4; the FP32 adjoint against the FP64 adjoint for every parameter.

FD bars. At this size J is rough at fine scales: the WENO-Z weights switch sharply where
the field is near the eps level (the tails ahead of and behind the wave fill most of the
room for hundreds of steps), so central differences do not converge as the step shrinks.
Measured on this problem: the drive FD scatters by 5e-4 over relative steps 1e-2 ... 1e-4
(and is meaningless below 1e-6); the R_inf FD approaches the adjoint as the step shrinks
(6.4e-3 at 1e-2, 1.5e-3 at 3e-4, 2.2e-4 at 1e-4). The bars are set from that scatter:
2e-3 for the drive, 1e-3 for R_inf. Residues and poles (the entry of each of A, zeta, B,
C, alpha, beta with the largest gradient): their gradients are 1e-5 to 1e-6 of the R_inf
one, so below a relative step of ~1e-3 the FD is the roughness of J (~1e-5 of J): for
zeta[5][3] the FD / adjoint ratio is 0.9996, 1.0048, 1.0185, 0.95, 0.18, 1.62 at steps
1e-2 ... 3e-5, while the smooth zeta[0][0] agrees to 1e-5 at every step; so these use the
step 1e-2 and the bar 1e-2. In the smooth regime (tests/test_gpu_adjoint.py, a 30x26x22
room) the same adjoint matches FD to 1e-4 and better.

FP32 storage: the workload keeps a radius-32 FP64 region around the source
(fases_params.source_patch). With the default radius 12 the FP32 drive gradient of this
room is 3 % to 14 % off FP64 on this misfit (33 % to 64 % on J = sum ir^2) over 200 to 3000
steps: it is read off the adjoint at the source cell while the wavelet is on (the first
~40 steps), after the adjoint has passed through the WENO-Z eps-level tails of the early
wave front, and beyond radius 12 the FP32-stored states there differ from FP64's where
the WENO-Z Jacobian is most sensitive to them (FP64 adjoint arithmetic over the same FP32
states was measured just as far off). With radius 32: 8e-4 at 3000 steps (1.1e-3 at 200),
materials 2.7e-5 of max |g|. Bars: drive 2e-3, materials 1e-4."""
import copy
import sys

import numpy as np
import pytest

from fases_amd import workloads as W
from fases_amd.autograd import free_mask, pack, unpack

pytestmark = pytest.mark.gpu

H_FD = 1e-4
# residues and poles: relative FD step and bar (see the module docstring)
H_FD_POLES = 1e-2
FD_BAR_POLES = 1e-2


def announce(msg):
    print("[c4] " + msg, flush=True)
    sys.stdout.flush()


class Problem:
    def __init__(self, prec, wl=None):
        self.wl = W.config(4) if wl is None else wl
        wl = self.wl
        self.sv = W.make_solver(wl, prec)
        self.target = self.sv.run(wl.nsteps)[0]
        truth = pack(wl.materials)
        rng = np.random.default_rng(1004)
        mask = free_mask(wl.materials.K, wl.materials.L, truth.shape[0])
        self.m0 = unpack(truth * np.exp(rng.uniform(-np.log(1.2), np.log(1.2), truth.shape) * mask),
                         wl.materials.K, wl.materials.L)
        self.s0 = float(wl.src_sval[0]) * 1.2

    def _load(self, m, sval):
        wl = self.wl
        self.sv.set_materials(m)
        self.sv.set_sources([wl.cell(s) for s in wl.src_ijk], [sval], wl.src_opt, wl.u_s)
        self.sv.set_fields(wl.initial_pressure())

    def J(self, m=None, sval=None):
        self._load(self.m0 if m is None else m, self.s0 if sval is None else sval)
        r = self.sv.run(self.wl.nsteps)[0]
        return float(((r - self.target) ** 2).sum())

    def grad(self):
        self._load(self.m0, self.s0)
        ir = self.sv.adjoint_forward(self.wl.nsteps, 64)
        gm, gs, _ = self.sv.adjoint(2.0 * (ir - self.target))
        return gm, gs[0]


def index(m, name, k):
    K, L = m.K, m.L
    base = {"Rinf": 0, "A": 1, "zeta": 1 + K, "B": 1 + 2 * K, "C": 1 + 2 * K + L, "alpha": 1 + 2 * K + 2 * L,
            "beta": 1 + 2 * K + 3 * L}[name]
    return base + (0 if k is None else k)


@pytest.fixture(scope="module")
def p64():
    announce("fp64 target and start point")
    p = Problem("fp64")
    yield p
    p.sv.close()


@pytest.fixture(scope="module")
def g64(p64):
    announce("fp64 adjoint, 3000 steps")
    return p64.grad()


def test_fp64_adjoint_matches_central_fd(p64, g64):
    gm, gs = g64
    announce("determinism of the FP64 forward")
    j0 = p64.J()
    assert p64.J() == j0
    h = H_FD * p64.s0
    announce("FD of the drive amplitude")
    fd = (p64.J(sval=p64.s0 + h) - p64.J(sval=p64.s0 - h)) / (2 * h)
    print("[c4] drive: fd %.10e adjoint %.10e (rel %.2e)" % (fd, gs, abs(gs / fd - 1)), flush=True)
    assert abs(gs - fd) <= 2e-3 * abs(fd)
    announce("FD of R_inf of wall 0")
    v0 = p64.m0.Rinf[0]
    hh = H_FD * abs(v0)
    vals = []
    for sg in (1, -1):
        m2 = copy.deepcopy(p64.m0)
        m2.Rinf[0] = v0 + sg * hh
        vals.append(p64.J(m=m2))
    fdm = (vals[0] - vals[1]) / (2 * hh)
    g = gm[0, index(p64.m0, "Rinf", None)]
    print("[c4] Rinf[0]: fd %.10e adjoint %.10e (rel %.2e)" % (fdm, g, abs(g / fdm - 1)), flush=True)
    assert abs(g - fdm) <= 1e-3 * abs(fdm)


@pytest.mark.parametrize("name", ["A", "zeta", "B", "C", "alpha", "beta"])
def test_fp64_adjoint_pole_parameters_match_central_fd(p64, g64, name):
    # the entry of this kind with the largest adjoint gradient (over walls and poles)
    gm, _ = g64
    m0 = p64.m0
    arr = getattr(m0, name)
    cols = [index(m0, name, k) for k in range(arr.shape[1])]
    sub = np.abs(gm[:, cols])
    w, k = np.unravel_index(int(np.argmax(sub)), sub.shape)
    g = gm[w, cols[k]]
    v0 = arr[w, k]
    hh = H_FD_POLES * abs(v0)
    announce("FD of %s[%d][%d]" % (name, w, k))
    vals = []
    for sg in (1, -1):
        m2 = copy.deepcopy(m0)
        getattr(m2, name)[w, k] = v0 + sg * hh
        vals.append(p64.J(m=m2))
    fd = (vals[0] - vals[1]) / (2 * hh)
    print("[c4] %s[%d][%d]: fd %.10e adjoint %.10e (rel %.2e)" % (name, w, k, fd, g, abs(g / fd - 1)), flush=True)
    assert abs(g - fd) <= FD_BAR_POLES * abs(fd)


def test_fp32_adjoint_matches_fp64_adjoint(g64):
    announce("fp32 adjoint, 3000 steps")
    p = Problem("fp32")
    gm32, gs32 = p.grad()
    p.sv.close()
    gm, gs = g64
    err = np.abs(gm32 - gm).max() / np.abs(gm).max()
    print("[c4] fp32 vs fp64 gradient: materials %.3e of max |g|, drive %.3e" % (err, abs(gs32 / gs - 1)),
          flush=True)
    assert p.wl.params.source_patch == 32
    assert err <= 1e-4
    assert abs(gs32 - gs) <= 2e-3 * abs(gs)
