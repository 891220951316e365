"""Repeated set-up calls on one context: the device footprint stays constant, a source
moved between two adjoint runs gives the gradients of a fresh context, and duplicate
source cells get the drive gradient of their merged source."""
import copy
import dataclasses
import sys

import numpy as np
import pytest

from fases_amd import workloads as W

pytestmark = pytest.mark.gpu


def announce(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def room(nsteps=80):
    return W.ade_room(room=(30, 26, 22), h=0.02, nsteps=nsteps, seed=21, n_rec=3, n_types=3, K=2, L=1, pml_x_hi=True,
                      pml_cells=5)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_device_bytes_constant_over_repeated_setup(prec):
    announce("20 x set_materials / set_sources / set_receivers / set_fields / run, %s" % prec)
    wl = room()
    sv = W.make_solver(wl, prec)
    try:
        rng = np.random.default_rng(3)
        sizes = []
        for it in range(20):
            m = copy.deepcopy(wl.materials)
            m.Rinf *= 0.9 + 0.1 * rng.random()
            sv.set_materials(m)
            # two source positions in turn (the FP64 patch around a source differs in size)
            src = wl.src_ijk[0] if it % 2 == 0 else (12, 11, 10)
            sv.set_sources([wl.cell(src)], wl.src_sval, wl.src_opt, wl.u_s)
            sv.set_receivers([wl.cell(r) for r in wl.rec_ijk])
            sv.set_fields(wl.initial_pressure())
            sv.run(30)
            sizes.append(sv.info()["dev_bytes"])
        print("dev_bytes:", sizes[:3], "...", sizes[-1])
        assert sizes[2:] == sizes[:2] * 9, sizes
    finally:
        sv.close()


def test_moved_source_adjoint_equals_fresh_context():
    """Two adjoint runs on one FP32 context with the source moved in between: the second
    gradients equal a fresh context's (the FP64 source-region maps are rebuilt)."""
    announce("adjoint after moving the source, fp32")
    # both positions keep the whole radius-12 FP64 region inside the room, so the region
    # has the same size at both: only the cell positions tell the two set-ups apart
    base = W.ade_room(room=(44, 40, 36), h=0.02, nsteps=100, seed=21, n_rec=3, n_types=3, K=2, L=1, pml_x_hi=False)
    wl = dataclasses.replace(base, src_ijk=[(23, 23, 21)])
    wl2 = dataclasses.replace(base, src_ijk=[(27, 20, 23)])

    def grads(sv, w):
        sv.set_materials(w.materials)
        sv.set_sources([w.cell(s) for s in w.src_ijk], w.src_sval, w.src_opt, w.u_s)
        sv.set_fields(w.initial_pressure())
        ir = sv.adjoint_forward(w.nsteps, 32)
        gm, gs, _ = sv.adjoint(2 * ir)
        return gm, gs

    sv = W.make_solver(wl, "fp32")
    try:
        grads(sv, wl)
        gm1, gs1 = grads(sv, wl2)
    finally:
        sv.close()
    sv2 = W.make_solver(wl2, "fp32")
    try:
        gm2, gs2 = grads(sv2, wl2)
    finally:
        sv2.close()
    np.testing.assert_array_equal(gm1, gm2)
    np.testing.assert_array_equal(gs1, gs2)


def test_duplicate_source_cells_get_the_merged_gradient():
    torch = pytest.importorskip("torch")
    from fases_amd.autograd import RoomIR, RoomProblem, pack
    announce("autograd with two sources in one cell")
    wl = room(60)
    a, b = 0.4 * wl.src_sval[0], 0.6 * wl.src_sval[0]
    wl_dup = dataclasses.replace(wl, src_ijk=[wl.src_ijk[0], wl.src_ijk[0]], src_sval=[a, b])
    out = []
    for w in (wl_dup, wl):
        sv = W.make_solver(w, "fp64")
        try:
            prob = RoomProblem(sv, w, checkpoint_every=16)
            th = torch.tensor(pack(w.materials), requires_grad=True)
            s = torch.tensor(np.asarray(w.src_sval, float), requires_grad=True)
            ir = RoomIR.apply(th, s, prob)
            (ir ** 2).sum().backward()
            out.append((ir.detach().numpy(), th.grad.numpy(), s.grad.numpy()))
        finally:
            sv.close()
    (ir_d, gm_d, gs_d), (ir_1, gm_1, gs_1) = out
    assert gs_d.shape == (2,)
    np.testing.assert_allclose(ir_d, ir_1, rtol=1e-14, atol=1e-300)
    np.testing.assert_allclose(gs_d, [gs_1[0], gs_1[0]], rtol=1e-12)
    np.testing.assert_allclose(gm_d, gm_1, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_second_geometry_on_one_context_equals_fresh_context(prec):
    """set_geometry on a context that has run (sources, FP64 source region, receivers, FP16
    scales all set up) starts it afresh: the run after it equals a fresh context's."""
    announce("set_geometry twice on one context, %s" % prec)
    wl = room(60)
    sc = wl.scene

    def load(sv):
        sv.set_geometry(sc.S, sc.oid, sc.dims)
        sv.set_materials(wl.materials)
        sv.set_sources([wl.cell(s) for s in wl.src_ijk], wl.src_sval, wl.src_opt, wl.u_s)
        sv.set_receivers([wl.cell(r) for r in wl.rec_ijk])
        sv.set_fields(wl.initial_pressure())

    sv = W.make_solver(wl, prec)
    try:
        ir0 = sv.run(wl.nsteps)[0]
        load(sv)
        ir1 = sv.run(wl.nsteps)[0]
    finally:
        sv.close()
    assert np.abs(ir0).max() > 0
    np.testing.assert_array_equal(ir1, ir0)
