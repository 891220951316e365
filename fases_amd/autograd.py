"""torch.autograd.Function over the hand-written adjoint of the FVM run, and an L-BFGS
recovery of ADE material parameters from receiver IRs.

The differentiable inputs are the material table (per material: R_inf, A[K], zeta[K],
B[L], C[L], alpha[L], beta[L], the same layout as fases_adjoint's gradient) and the
source drive amplitudes; the output is the receiver pressure IR [n_rec, nsteps].
PyTorch only carries the tensors and the optimizer: the forward run, the checkpointed
reverse sweep and the gradients are the library's (fases_adjoint_forward /
fases_adjoint). This replaces the reference's gradient-free fitting loops
(vectFIT.py / fminconFIT.py) with gradients of the full room simulation.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from ._lib import FasesError
from .solver import Materials

FIELDS = ("Rinf", "A", "zeta", "B", "C", "alpha", "beta")


def pack(m: Materials) -> np.ndarray:
    """Materials -> [nMat, 1 + 2K + 4L] in the adjoint's gradient layout."""
    return np.concatenate([m.Rinf[:, None], m.A, m.zeta, m.B, m.C, m.alpha, m.beta], axis=1)


def unpack(theta: np.ndarray, K: int, L: int) -> Materials:
    t = np.asarray(theta, np.float64)
    o = [0, 1, 1 + K, 1 + 2 * K, 1 + 2 * K + L, 1 + 2 * K + 2 * L, 1 + 2 * K + 3 * L, 1 + 2 * K + 4 * L]
    parts = [t[:, o[q]:o[q + 1]] for q in range(7)]
    return Materials(parts[0][:, 0].copy(), *[p.copy() for p in parts[1:]])


class RoomIR(torch.autograd.Function):
    """ir = RoomIR.apply(theta [nMat, P], sval [n_src], problem). `problem` is a
    RoomProblem (solver, workload, steps, checkpoint interval). backward() runs the
    adjoint for the last forward of that solver, so call it before the next forward."""

    @staticmethod
    def forward(ctx, theta, sval, problem):
        ir = problem.forward(theta.detach().cpu().numpy(), sval.detach().cpu().numpy())
        ctx.problem = problem
        return torch.from_numpy(ir)

    @staticmethod
    def backward(ctx, grad_ir):
        gm, gs = ctx.problem.backward(grad_ir.detach().cpu().numpy())
        return torch.from_numpy(gm), torch.from_numpy(gs), None


class RoomProblem:
    """A loaded solver context and workload with differentiable materials and drives."""

    def __init__(self, solver, workload, nsteps: int | None = None, checkpoint_every: int = 64):
        self.sv = solver
        self.wl = workload
        self.nsteps = workload.nsteps if nsteps is None else nsteps
        self.every = checkpoint_every
        self.K, self.L = workload.materials.K, workload.materials.L
        self.cells = [workload.cell(s) for s in workload.src_ijk]
        # fases_set_sources merges non-sine sources that share a cell (their drives add,
        # in first-appearance order); each source's drive gradient is its merged entry's
        first = {}
        for c in self.cells:
            first.setdefault(c, len(first))
        sine = workload.src_opt == _lib.FASES_SRC_SINE
        self.merged = np.arange(len(self.cells)) if sine else np.array([first[c] for c in self.cells], np.int64)

    def forward(self, theta: np.ndarray, sval: np.ndarray) -> np.ndarray:
        wl = self.wl
        self.sv.set_materials(unpack(theta, self.K, self.L))
        self.sv.set_sources(self.cells, sval, wl.src_opt, wl.u_s, wl.src_freq)
        self.sv.set_fields(wl.initial_pressure())
        return self.sv.adjoint_forward(self.nsteps, self.every)

    def backward(self, dj_dir: np.ndarray):
        gm, gs, _ = self.sv.adjoint(dj_dir)
        return gm, gs[self.merged]


def free_mask(K: int, L: int, n: int) -> np.ndarray:
    """1 for the fitted entries (R_inf, residues A, B, C, pole real parts zeta, alpha),
    0 for the pair frequencies beta, which stay at their initial values."""
    m = np.ones((n, 1 + 2 * K + 4 * L))
    m[:, 1 + 2 * K + 3 * L:] = 0.0
    return m


def lbfgs_recover(problem: RoomProblem, target_ir: np.ndarray, theta0: np.ndarray, sval0: np.ndarray,
                  iterations: int = 50, lr: float = 0.02, warm: int = 3, max_ls: int = 8, log=None):
    """Fits theta = theta0 * exp(u) (multiplicative, so signs and scales are kept; the pair
    frequencies stay fixed) and sval = sval0 * exp(v) to the target IRs by L-BFGS on
    J = sum (ir - target)^2 (normalised by sum target^2 inside the optimiser). Returns
    (theta, sval, history) at the lowest misfit found;
    history[0] is the starting misfit and history[n] the lowest after iteration n.
    Each iteration is one quasi-Newton direction and a strong-Wolfe line search of at
    most `max_ls` trial runs. The first trial is lr times the quasi-Newton step for the
    first `warm` iterations (the first step is then at most lr per log-parameter: a full
    step can push a wall past |R| = 1, where the room becomes unstable), the full step
    afterwards; a trial point whose run fails, or whose misfit or gradient is not
    finite, counts as 1e3 times the first misfit with a zero gradient, so the line
    search brackets it and backs off."""
    th0 = torch.from_numpy(np.asarray(theta0, np.float64))
    s0 = torch.from_numpy(np.asarray(sval0, np.float64))
    mask = torch.from_numpy(free_mask(problem.K, problem.L, th0.shape[0]))
    u = torch.zeros_like(th0, requires_grad=True)
    v = torch.zeros_like(s0, requires_grad=True)
    tgt = torch.from_numpy(np.asarray(target_ir, np.float64))
    # the optimiser sees J / sum(target^2): torch's L-BFGS has absolute thresholds (the
    # curvature pair is kept only if y.s > 1e-10), so the misfit has to be O(1)-scaled
    jscale = float((tgt ** 2).sum()) or 1.0
    # one step() call runs the closure once, then the line search's first trial and at
    # most max_eval - 1 more (torch's default max_eval for max_iter = 1 leaves the line
    # search no trial beyond the first, so a failed trial could never be bracketed)
    # torch's absolute stopping tolerances (gradient 1e-7, change 1e-9) would end the run
    # early on a small misfit; the iteration count is the caller's, so they are off
    opt = torch.optim.LBFGS([u, v], lr=lr, max_iter=1, max_eval=max_ls, history_size=20,
                            tolerance_grad=0.0, tolerance_change=0.0, line_search_fn="strong_wolfe")
    first = []
    runs = [0]
    best = {"J": float("inf"), "u": u.detach().clone(), "v": v.detach().clone()}

    def unstable():
        # a large finite misfit and no slope (a NaN misfit would defeat the line
        # search's comparisons and could be accepted)
        u.grad = torch.zeros_like(u)
        v.grad = torch.zeros_like(v)
        return torch.tensor(1e3 * (first[0] if first else 1.0), dtype=torch.float64)

    # every step() call first evaluates the closure at the current point, which the
    # previous line search has already run: answer it from the evaluations kept here
    seen = []

    def closure():
        opt.zero_grad()
        for (cu, cv, cJ, gu, gv) in seen:
            if torch.equal(cu, u.detach()) and torch.equal(cv, v.detach()):
                u.grad = gu.clone()
                v.grad = gv.clone()
                return torch.tensor(cJ, dtype=torch.float64)
        try:
            ir = RoomIR.apply(th0 * torch.exp(u * mask), s0 * torch.exp(v), problem)
        except FasesError:
            return unstable()
        J = ((ir - tgt) ** 2).sum() / jscale
        if not bool(torch.isfinite(J)):
            # the run overflowed (FP32 storage does not stop on a non-finite value)
            return unstable()
        J.backward()
        if not (bool(torch.isfinite(u.grad).all()) and bool(torch.isfinite(v.grad).all())):
            return unstable()
        Jf = float(J.detach())
        if not first:
            first.append(Jf)
        if Jf < best["J"]:
            best.update(J=Jf, u=u.detach().clone(), v=v.detach().clone())
        seen.append((u.detach().clone(), v.detach().clone(), Jf, u.grad.clone(), v.grad.clone()))
        del seen[:-2 * max_ls]
        runs[0] += 1
        return J

    hist = []
    for it in range(iterations):
        if it == warm:
            for g in opt.param_groups:
                g["lr"] = 1.0
        J = float(opt.step(closure).detach())
        if it == 0:
            hist.append(J * jscale)
        hist.append(best["J"] * jscale)
        if log:
            log("L-BFGS iteration %d: J = %.6e (%d runs)" % (it, best["J"] * jscale, runs[0]))
    with torch.no_grad():
        return (th0 * torch.exp(best["u"] * mask)).numpy(), (s0 * torch.exp(best["v"])).numpy(), hist
