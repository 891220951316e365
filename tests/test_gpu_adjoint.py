"""Hand-written adjoint (fases_adjoint_forward / fases_adjoint) against central finite
differences of the FP64 forward (bit-identical to the FP64 oracle), a dot-product test,
checkpoint-interval invariance, FP32 vs FP64 gradients, the torch.autograd.Function
wrapper and a short L-BFGS recovery.

Finite-difference step: WENO-Z weights are very curved where the field is at the level
of eps (the wave front's tail), so central differences with a relative step of 1e-4 carry
errors up to ~1e-3 there; the FD reference here uses 1e-6 (measured to converge), and the
bar is 1e-4 relative."""
import copy
import dataclasses
import sys

import numpy as np
import pytest

from fases_amd import _lib
from fases_amd import workloads as W
from fases_amd._lib import FasesError

pytestmark = pytest.mark.gpu

H_FD = 1e-6
FIELDS = [("Rinf", None), ("A", 0), ("A", 1), ("zeta", 0), ("zeta", 1), ("B", 0), ("C", 0), ("alpha", 0), ("beta", 0)]


def announce(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def room(nsteps=150):
    return W.ade_room(room=(30, 26, 22), h=0.02, nsteps=nsteps, seed=21, n_rec=3, n_types=3, K=2, L=1, pml_x_hi=False)


def perturbed(m, seed=5, rel=0.2):
    rng = np.random.default_rng(seed)
    m2 = copy.deepcopy(m)
    for a in (m2.Rinf, m2.A, m2.zeta, m2.B, m2.C, m2.alpha):
        a *= np.exp(rng.uniform(-rel, rel, a.shape))
    return m2


def pml_room(nsteps=150):
    """The FD room with cPML bands on three faces, one per axis, low and high faces
    (keys U D R L F B: z+, x+ and y-), 5 cells deep."""
    wl = room(nsteps)
    return dataclasses.replace(wl, params=dataclasses.replace(wl.params, pml_keys=(1, 0, 1, 0, 0, 1), pml_cells=5))


class Setup:
    def __init__(self, prec="fp64", nsteps=150, wl=None):
        self.wl = room(nsteps) if wl is None else wl
        self.sv = W.make_solver(self.wl, prec)
        self.sv.set_materials(perturbed(self.wl.materials))
        self.sv.set_fields(self.wl.initial_pressure())
        self.d, _, _, _ = self.sv.run(self.wl.nsteps)
        self.sv.set_materials(self.wl.materials)

    def J(self, m=None, sval=None, u_s=None):
        wl = self.wl
        self.sv.set_materials(wl.materials if m is None else m)
        self.sv.set_sources([wl.cell(s) for s in wl.src_ijk], wl.src_sval if sval is None else sval, wl.src_opt,
                            wl.u_s if u_s is None else u_s)
        self.sv.set_fields(wl.initial_pressure())
        r, _, _, _ = self.sv.run(wl.nsteps)
        return float(((r - self.d) ** 2).sum()), r

    def grad(self, every=64, dj=None):
        wl = self.wl
        self.sv.set_materials(wl.materials)
        self.sv.set_sources([wl.cell(s) for s in wl.src_ijk], wl.src_sval, wl.src_opt, wl.u_s)
        self.sv.set_fields(wl.initial_pressure())
        ir = self.sv.adjoint_forward(wl.nsteps, every)
        return self.sv.adjoint(2 * (ir - self.d) if dj is None else dj, want_us=True), ir


@pytest.fixture(scope="module")
def s64():
    s = Setup("fp64")
    yield s
    s.sv.close()


def test_material_gradients_match_central_fd(s64):
    announce("adjoint vs FD, 3 materials x 9 parameters")
    (gm, gs, _), _ = s64.grad()
    fd = np.zeros_like(gm)
    for mi in range(3):
        for pi, (name, k) in enumerate(FIELDS):
            base = getattr(s64.wl.materials, name)
            v0 = base[mi] if k is None else base[mi, k]
            h = H_FD * abs(v0)
            vals = []
            for sg in (1, -1):
                m2 = copy.deepcopy(s64.wl.materials)
                arr = getattr(m2, name)
                if k is None:
                    arr[mi] = v0 + sg * h
                else:
                    arr[mi, k] = v0 + sg * h
                vals.append(s64.J(m=m2)[0])
            fd[mi, pi] = (vals[0] - vals[1]) / (2 * h)
    gm, fd = gm[:3], fd[:3]
    err = np.abs(gm - fd)
    print("max rel err %.2e" % (err / np.abs(fd)).max())
    assert np.all(err <= 1e-4 * np.abs(fd) + 1e-7 * np.abs(fd).max())


def test_source_gradients_match_fd(s64):
    announce("adjoint vs FD, source amplitude and waveform samples")
    (gm, gs, gu), _ = s64.grad()
    s0 = s64.wl.src_sval[0]
    h = H_FD * s0
    fd = (s64.J(sval=[s0 + h])[0] - s64.J(sval=[s0 - h])[0]) / (2 * h)
    assert abs(gs[0] - fd) <= 1e-6 * abs(fd)
    # waveform samples (the WENO weights make the IR nonlinear in them, so a small step)
    us = s64.wl.u_s
    k = int(np.argmax(np.abs(gu)))
    for idx in (k, k // 2, 2 * k):
        hs = H_FD * abs(us).max()
        up, dn = us.copy(), us.copy()
        up[idx] += hs
        dn[idx] -= hs
        fd = (s64.J(u_s=up)[0] - s64.J(u_s=dn)[0]) / (2 * hs)
        print("u_s[%d]: fd %.8e adjoint %.8e" % (idx, fd, gu[idx]))
        assert abs(gu[idx] - fd) <= 1e-4 * abs(fd) + 1e-6 * abs(gu).max(), idx


def test_dot_product(s64):
    """<J v, w> = <v, J^T w> with J = d ir / d theta, v a random parameter direction, w a
    random IR weight: the left side from a directional central difference, the right side
    from one adjoint sweep with dJ/d ir = w."""
    announce("dot-product test")
    rng = np.random.default_rng(3)
    wl = s64.wl
    w = rng.standard_normal(s64.d.shape)
    (gm, gs, _), _ = s64.grad(dj=w)
    m = wl.materials
    dirs = {name: rng.standard_normal(getattr(m, name).shape) * np.abs(getattr(m, name)) for name in
            ("Rinf", "A", "zeta", "B", "C", "alpha")}
    dsv = rng.standard_normal() * wl.src_sval[0]

    def shifted(t):
        m2 = copy.deepcopy(m)
        for name, dv in dirs.items():
            getattr(m2, name)[...] += t * dv
        return m2

    lhs_h = []
    for t in (H_FD, 0.1 * H_FD):
        rp = s64.J(m=shifted(t), sval=[wl.src_sval[0] + t * dsv])[1]
        rm = s64.J(m=shifted(-t), sval=[wl.src_sval[0] - t * dsv])[1]
        lhs_h.append(float(((rp - rm) / (2 * t) * w).sum()))
    K, L = m.K, m.L
    rhs = float((gm[:, 0] * dirs["Rinf"]).sum() + (gm[:, 1:1 + K] * dirs["A"]).sum()
                + (gm[:, 1 + K:1 + 2 * K] * dirs["zeta"]).sum() + (gm[:, 1 + 2 * K:1 + 2 * K + L] * dirs["B"]).sum()
                + (gm[:, 1 + 2 * K + L:1 + 2 * K + 2 * L] * dirs["C"]).sum()
                + (gm[:, 1 + 2 * K + 2 * L:1 + 2 * K + 3 * L] * dirs["alpha"]).sum() + gs[0] * dsv)
    print("<Jv,w> = %.10e (step 1e-6), %.10e (step 1e-7); <v,JTw> = %.10e" % (lhs_h[0], lhs_h[1], rhs))
    # the difference quotient itself moves by ~1e-4 between steps (curved WENO weights)
    assert min(abs(v - rhs) for v in lhs_h) <= 1e-4 * abs(rhs)


def test_checkpoint_interval_does_not_change_gradients(s64):
    announce("checkpoint interval 16 vs 64 vs 150")
    (a, sa, _), _ = s64.grad(every=64)
    for ev in (16, 150):
        (b, sb, _), _ = s64.grad(every=ev)
        assert np.array_equal(a, b) and np.array_equal(sa, sb)


def test_fp32_gradients_close_to_fp64(s64):
    announce("fp32 adjoint vs fp64 adjoint")
    (g64, s64g, _), _ = s64.grad()
    s32 = Setup("fp32")
    (g32, s32g, _), _ = s32.grad()
    s32.sv.close()
    for q in range(g64.shape[0]):
        assert np.abs(g32[q] - g64[q]).max() <= 1e-3 * np.abs(g64[q]).max()
    assert abs(s32g[0] - s64g[0]) <= 1e-4 * abs(s64g[0])


def test_autograd_function_and_lbfgs():
    torch = pytest.importorskip("torch")
    from fases_amd.autograd import RoomIR, RoomProblem, lbfgs_recover, pack
    announce("torch.autograd.Function + L-BFGS")
    s = Setup("fp64", nsteps=120)
    prob = RoomProblem(s.sv, s.wl, checkpoint_every=32)
    th = torch.tensor(pack(s.wl.materials), requires_grad=True)
    sv_t = torch.tensor(np.asarray(s.wl.src_sval, float), requires_grad=True)
    ir = RoomIR.apply(th, sv_t, prob)
    J = ((ir - torch.from_numpy(s.d)) ** 2).sum()
    J.backward()
    (gm, gs, _), _ = s.grad()
    assert np.allclose(th.grad.numpy(), gm, rtol=1e-12, atol=0)
    assert np.allclose(sv_t.grad.numpy(), gs, rtol=1e-12, atol=0)
    # a few L-BFGS iterations from the perturbed start towards the target's materials
    target = s.d
    theta, sval, hist = lbfgs_recover(prob, target, pack(s.wl.materials), np.asarray(s.wl.src_sval), iterations=8)
    print("L-BFGS J:", ["%.3e" % v for v in hist])
    assert hist[-1] < 0.3 * hist[0]
    s.sv.close()


@pytest.fixture(scope="module")
def spml():
    s = Setup("fp64", wl=pml_room())
    yield s
    s.sv.close()


def test_pml_gradients_match_fd(spml):
    """cPML on z+, x+ and y-: the adjoint carries the PML part of the flux adjoint and the
    adjoint of the auxiliary states; material and drive gradients against central FD."""
    announce("adjoint with cPML vs FD")
    (gm, gs, _), _ = spml.grad(every=64)
    wl = spml.wl
    s0 = wl.src_sval[0]
    h = H_FD * s0
    fd = (spml.J(sval=[s0 + h])[0] - spml.J(sval=[s0 - h])[0]) / (2 * h)
    print("drive: fd %.10e adjoint %.10e" % (fd, gs[0]))
    assert abs(gs[0] - fd) <= 1e-6 * abs(fd)
    for mi, (name, k) in [(0, ("Rinf", None)), (1, ("A", 0)), (2, ("zeta", 1)), (0, ("B", 0)), (1, ("alpha", 0))]:
        base = getattr(wl.materials, name)
        v0 = base[mi] if k is None else base[mi, k]
        hh = H_FD * abs(v0)
        vals = []
        for sg in (1, -1):
            m2 = copy.deepcopy(wl.materials)
            arr = getattr(m2, name)
            if k is None:
                arr[mi] = v0 + sg * hh
            else:
                arr[mi, k] = v0 + sg * hh
            vals.append(spml.J(m=m2)[0])
        fdm = (vals[0] - vals[1]) / (2 * hh)
        pi = FIELDS.index((name, k))
        print("%s[%d,%s]: fd %.10e adjoint %.10e" % (name, mi, k, fdm, gm[mi, pi]))
        assert abs(gm[mi, pi] - fdm) <= 1e-4 * abs(fdm) + 1e-7 * np.abs(gm[:3]).max()


def test_pml_checkpoint_interval_and_fp32(spml):
    """The psi checkpoints restore the PML states exactly (gradients independent of the
    checkpoint interval); FP32 gradients stay close to FP64 with the PML on."""
    announce("adjoint with cPML: checkpoint interval, fp32 vs fp64")
    (a, sa, _), _ = spml.grad(every=64)
    (b, sb, _), _ = spml.grad(every=7)
    assert np.array_equal(a, b) and np.array_equal(sa, sb)
    s32 = Setup("fp32", wl=pml_room())
    (g32, s32g, _), _ = s32.grad()
    s32.sv.close()
    for q in range(3):
        assert np.abs(g32[q] - a[q]).max() <= 1e-3 * np.abs(a[q]).max()
    assert abs(s32g[0] - sa[0]) <= 1e-4 * abs(sa[0])


def test_adjoint_before_forward_is_an_error():
    announce("adjoint error paths")
    wl = W.ade_room(room=(24, 20, 18), h=0.02, nsteps=10, seed=2, n_rec=1, K=1, L=1, pml_cells=4)
    wl2 = dataclasses.replace(wl, params=dataclasses.replace(wl.params, pml_keys=(0,) * 6))
    sv = W.make_solver(wl2, "fp16")
    with pytest.raises(FasesError, match="adjoint_forward"):
        sv.adjoint(np.zeros((1, 10)))
    sv.close()


def test_fp16_gradients_close_to_fp64(s64):
    """FP16 storage (one power-of-two scale per field, FP32 arithmetic, FP64 ADE states):
    gradients within 5e-3 of the largest FP64 gradient entry for every material (measured
    1.7e-3: a far wall's small entries carry the storage's 2^-11 resolution of the field's
    largest value), drive gradient within 1e-2 (measured 2.5e-3)."""
    announce("fp16 adjoint vs fp64 adjoint")
    (g64, s64g, _), _ = s64.grad()
    s16 = Setup("fp16")
    (g16, s16g, _), _ = s16.grad()
    s16.sv.close()
    gmax = np.abs(g64).max()
    assert np.isfinite(g16).all()
    assert np.abs(g16 - g64).max() <= 5e-3 * gmax
    assert abs(s16g[0] - s64g[0]) <= 1e-2 * abs(s64g[0])


def test_fp16_adjoint_replays_the_rescales():
    """FP16 with a rescale every 7 steps (and at steps 1, 2, 4): the checkpointed forward's IR
    equals the plain run bit for bit, and checkpoint intervals of 64, 16, 5 and 150 steps
    (rescales inside and at the start of segments) give the same gradients bit for bit."""
    announce("fp16 adjoint rescale replay")
    wl = room(150)
    wl = dataclasses.replace(wl, params=dataclasses.replace(wl.params, fp16_rescale_every=7))
    s16 = Setup("fp16", wl=wl)
    (a, sa, _), ir = s16.grad(every=64)
    s16.sv.set_materials(wl.materials)
    s16.sv.set_fields(wl.initial_pressure())
    r, _, _, _ = s16.sv.run(wl.nsteps)
    assert np.array_equal(ir, r)
    for ev in (16, 5, 150):
        (b, sb, _), _ = s16.grad(every=ev)
        assert np.array_equal(a, b) and np.array_equal(sa, sb)
    s16.sv.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32", None)])
def test_tiled_adjoint_equals_per_cell(prec, tol):
    """The tiled flux adjoint (each pair VJP once, gathered by shuffles / LDS / z ring) and
    the per-cell one (every cell re-evaluates its neighbours' pairs) differ only in the
    order of the sums; the room spans several tiles on every axis (tile seams). FP64
    only: 60 steps into a large room the gradient is carried by the wave front's tail,
    where the WENO-Z weights switch sharply, and FP32 rounding moves it by percents
    (the FP32 bar is set on a settled misfit in test_fp32_gradients_close_to_fp64).
    In both precisions the tiled kernel gives the same bits in the XCD-aware block order."""
    announce("tiled vs per-cell adjoint %s" % prec)
    wl = W.ade_room(room=(70, 30, 75), h=0.02, nsteps=60, seed=8, n_rec=4, n_types=2, K=2, L=1, pml_x_hi=False)
    out = []
    for av in (0, 1, 2):
        w2 = dataclasses.replace(wl, params=dataclasses.replace(wl.params, adjoint_variant=av))
        sv = W.make_solver(w2, prec)
        ir = sv.adjoint_forward(wl.nsteps, 16)
        # J = sum ir^2 / 2: a misfit-like weight (random weights on every sample make the
        # gradient a sum of large cancelling terms, which FP32 cannot resolve)
        out.append(sv.adjoint(ir))
        sv.close()
    (ga, sa, _), (gb, sb, _), (gx, sx, _) = out
    # the tile kernel in another block order (XCD-aware, 1-D grid) does the same sums
    assert np.array_equal(ga, gx) and np.array_equal(sa, sx)
    if tol is None:
        return
    scale = np.abs(gb).max(axis=1, keepdims=True)
    print("max rel diff: materials %.3e, drive %.3e" % ((np.abs(ga - gb) / scale).max(),
                                                        (np.abs(sa - sb) / np.abs(sb)).max()))
    assert np.all(np.abs(ga - gb) <= tol * scale)
    assert np.all(np.abs(sa - sb) <= tol * np.abs(sb))


def test_fp32_drive_gradient_near_point_source():
    """FP32 storage: the cells around a point source are advanced in FP64 from FP64
    shadows, and their adjoint is too (without it the FP32 drive gradient of a
    source-dominated misfit is wrong by factors). J = sum ir^2, so dJ/dsval is close to
    2 J / sval (the IR is nearly linear in the drive)."""
    announce("fp32 drive gradient with the FP64 adjoint patch")
    wl = W.ade_room(room=(60, 50, 44), h=0.02, nsteps=300, seed=13, n_rec=4, n_types=2, K=2, L=1, pml_x_hi=False)
    gs = {}
    for prec in ("fp32", "fp64"):
        sv = W.make_solver(wl, prec)
        ir = sv.adjoint_forward(wl.nsteps, 64)
        _, g, _ = sv.adjoint(2 * ir)
        gs[prec] = (g[0], 2 * float((ir ** 2).sum()) / wl.src_sval[0])
        sv.close()
    print("dJ/dsval fp32 %.8e fp64 %.8e (2J/sval %.8e)" % (gs["fp32"][0], gs["fp64"][0], gs["fp64"][1]))
    assert abs(gs["fp32"][0] - gs["fp64"][0]) <= 1e-4 * abs(gs["fp64"][0])
    assert abs(gs["fp64"][0] - gs["fp64"][1]) <= 0.02 * abs(gs["fp64"][1])


def test_fp32_adjoint_stays_finite_where_the_field_is_at_eps():
    """96x80x64 room: early in the run most of the room is at the WENO eps level, where
    the exact adjoint grows ~10^7 over a few tens of steps (FP64 shows it too). The FP32
    weight derivatives must stay bounded there (normalised-weight form) and the gradients
    must follow FP64 once the wave has filled the room."""
    announce("FP32 adjoint at the eps level, 96x80x64")

    def grads(prec, steps, every, djf):
        wl = W.config(4, nsteps=steps, room=(96, 80, 64))
        sv = W.make_solver(wl, prec)
        sv.set_fields(wl.initial_pressure())
        ir = sv.adjoint_forward(steps, every)
        gm, gs, _ = sv.adjoint(djf(ir))
        sv.close()
        return gm, gs

    gm, gs = grads("fp32", 64, 16, np.ones_like)
    assert np.isfinite(gm).all() and np.isfinite(gs).all()
    g64, s64 = grads("fp64", 256, 64, lambda ir: 2.0 * ir)
    g32, s32 = grads("fp32", 256, 64, lambda ir: 2.0 * ir)
    assert np.abs(g32 - g64).max() <= 1e-3 * np.abs(g64).max()
    assert abs(s32[0] - s64[0]) <= 1e-3 * abs(s64[0])
