"""Physics pins of the FP64 oracle that do not depend on retyping its formulas: rigid-box
mode frequencies, the cPML's reflection of a normally incident plane wave, and the
normal-incidence reflection factor of ADE walls against the material's model R(f).

The GPU FP64 path equals the oracle to <= 1e-10 (tests/test_gpu_parity.py), so these
pins carry over to it.

Method of the wall tests (the one of tests/ADE_benchmark/ADE_TST_3D.py:162-221): a plane
pulse runs down a duct of rigid side walls towards the +x wall; p_d is the same pulse in
a duct three times as long (no reflection inside the window), so p - p_d is the wave
reflected at +x. R_num(f) = FFT(p_wall - p_d) / FFT(p_rigid - p_d).
"""
import numpy as np
import pytest

from fases_amd import workloads as W
from fases_amd.solver import Materials, Scene, SolverParams
from oracle import oracle as O

H = np.float32(0.05)
C0, RHO = 343.0, 1.2
DT = W.time_step(0.5, H, C0)
NSTEPS = 700
RIGID = {"Rinf": 1.0, "A": [0.0], "zeta": [0.0], "B": [0.0], "C": [0.0], "alpha": [0.0], "beta": [0.0]}


def duct(nx, east=None, pml=False, ny=8, nz=8, x0=60, xr=100, sigma=2.0, K=1, L=1):
    """Right-going Gaussian plane pulse (p = Z u) in an nx x ny x nz duct; the +x wall has
    material `east` (rigid when None). Returns the pressure at room cell (xr, ny/2, nz/2)."""
    sc = Scene.shoebox(nx, ny, nz, float(H))
    Nx, Ny, Nz = sc.dims
    walls = [RIGID] * 6
    walls = [dict(w) for w in walls]
    if east is not None:
        walls[1] = east
    K = max(K, max(len(w["A"]) for w in walls))
    L = max(L, max(len(w["B"]) for w in walls))
    m = W.materials_from(walls, K, L)
    keys = (0, 0, 1, 0, 0, 0) if pml else (0,) * 6
    prm = SolverParams(h=float(H), dt=DT, c=C0, rho=RHO, pml_keys=keys, pml_cells=16, pml_sigma_max=3 * C0 / float(H))
    g = np.exp(-((np.arange(Nx) - (x0 + 3)) ** 2) / (2 * sigma * sigma))
    p = np.broadcast_to(g[None, None, :], (Nz, Ny, Nx)).ravel().copy()
    p[sc.S <= 0] = 0.0
    u = p / (RHO * C0)
    rec = [(xr + 3) + Nx * ((ny // 2 + 3) + Ny * (nz // 2 + 3))]
    r = O.run(sc.S, sc.oid, sc.dims, prm, m, NSTEPS, rec_cells=rec, ic=(p, u, None, None))
    return r["ir"][0]


@pytest.fixture(scope="module")
def free_and_rigid():
    return duct(600), duct(200)


def reflection(p_wall, p_d, p_rigid, n=4096):
    Pf = np.fft.rfft(p_rigid - p_d, n)
    f = np.fft.rfftfreq(n, DT)
    band = np.abs(Pf) > 0.05 * np.abs(Pf).max()
    return f, np.fft.rfft(p_wall - p_d, n) / Pf, band


def test_rigid_box_modes_match_effective_room():
    """Spectral peaks of a rigid box IR at the closed-form modes f = c/2 |(l/Lx, m/Ly, n/Lz)|
    (modes3d.py:259-266) of the effective room. Walls are imposed over the 3h boundary band
    (functions3D.cuh:263-271, 1097), so the effective room is the OBJ room minus 3h per
    side; tests/Room_Modes_benchmark/validate_modes.py:27 uses the same (padded grid - 12h)."""
    room, h = (24, 18, 14), 0.1
    wl = W.rigid_box(room=room, h=h, nsteps=4000, ic_center=(3, 4, 2), rec=(20, 15, 11))
    ir = O.run_workload(wl)["ir"][0]
    n = ir.size
    spec = np.abs(np.fft.rfft((ir - ir.mean()) * np.hanning(n), 16 * n))
    f = np.fft.rfftfreq(16 * n, wl.params.dt)
    sel = (f > 40) & (f < 200)
    i = np.nonzero(sel)[0]
    peaks = [k for k in i if spec[k] > spec[k - 1] and spec[k] >= spec[k + 1] and spec[k] > 0.05 * spec[sel].max()]
    Le = [(d - 6) * wl.h for d in room]
    modes = {(l, m, q): 0.5 * wl.params.c * np.sqrt((l / Le[0]) ** 2 + (m / Le[1]) ** 2 + (q / Le[2]) ** 2)
             for l in range(4) for m in range(4) for q in range(4) if l + m + q}
    pf = f[peaks]
    for key in ((1, 0, 0), (0, 1, 0), (1, 1, 0), (2, 0, 0)):
        assert np.min(np.abs(pf - modes[key])) / modes[key] < 0.01, (key, modes[key], pf)
    fm = np.array(sorted(modes.values()))
    for fp in pf:
        assert np.min(np.abs(fm - fp)) / fp < 0.01, fp
    # the OBJ-size room does not fit: its (1,0,0) mode is 25 % lower and absent
    f_obj = 0.5 * wl.params.c / (room[0] * wl.h)
    assert np.min(np.abs(pf - f_obj)) / f_obj > 0.05


def test_cpml_absorbs_normally_incident_plane_wave(free_and_rigid):
    """16-cell cPML (sigma_max = 3c/h, kappa_max 5, alpha_max 20 pi) in front of a rigid +x
    wall: the reflected pulse is below 0.5 % of the incident one (the wall alone returns
    all of it), and the signal before the reflection can arrive is untouched."""
    p_d, p_rigid = free_and_rigid
    p_pml = duct(200, pml=True)
    inc = np.abs(p_d).max()
    assert np.abs(p_pml - p_d).max() / inc < 5e-3
    assert np.abs(p_rigid - p_d).max() / inc > 0.5
    # the PML starts 84 cells past the receiver: before the reflection can be back only the
    # Gaussian's far tail (below 1e-40 at the PML) and the scheme's stencil reach differ
    t_arr = int(2 * 84 * float(H) / C0 / DT) - 40
    assert np.abs(p_pml[:t_arr] - p_d[:t_arr]).max() <= 1e-4 * inc


def designed_material():
    z, al, be = 2 * np.pi * 300, 2 * np.pi * 100, 2 * np.pi * 500
    hy = np.hypot(al, be)
    return {"Rinf": 0.3, "A": [0.5 * z], "zeta": [z], "B": [-0.2 * hy], "C": [0.15 * hy], "alpha": [al], "beta": [be]}


def test_ade_wall_reflection_matches_model(free_and_rigid):
    """R_num(f) of a wall with R(s) = 0.3 + A/(s + zeta) + (B(s + alpha) + C beta)/((s + alpha)^2
    + beta^2) follows the model across |R| = 0.06 .. 0.91: within 0.08 up to 500 Hz (>= 13
    cells per wavelength), 0.2 up to the top of the pulse band (~1 kHz, grid dispersion)."""
    p_d, p_rigid = free_and_rigid
    mat = designed_material()
    f, R, band = reflection(duct(200, east=mat), p_d, p_rigid)
    Rm = W.materials_from([mat], 1, 1).response(f)[0]
    err = np.abs(R - Rm)
    assert err[band & (f <= 500)].max() < 0.08
    assert err[band].max() < 0.2
    # a sign error in the real-pole term or in the complex pair would be far off
    for wrong in ({"A": [-mat["A"][0]]}, {"C": [-mat["C"][0]]}):
        bad = dict(mat, **wrong)
        e_bad = np.abs(R - W.materials_from([bad], 1, 1).response(f)[0])
        assert e_bad[band].max() > 0.3


@pytest.mark.parametrize("rinf", [0.0, -1.0, 0.5, 0.8])
def test_frequency_independent_walls(free_and_rigid, rinf):
    """R_inf only: absorbing (0), pressure-release (-1) and partly reflecting walls. The
    rigid wall is the normalisation, so the deviation grows with |1 - R| (the 3h blend band
    places the impedance wall slightly differently from the rigid one): within
    0.015 + 0.06 |1 - R| up to 500 Hz."""
    p_d, p_rigid = free_and_rigid
    f, R, band = reflection(duct(200, east=dict(RIGID, Rinf=rinf)), p_d, p_rigid)
    assert np.abs(R[band & (f <= 500)] - rinf).max() < 0.015 + 0.06 * abs(1.0 - rinf)


def test_spherical_pulse_matches_closed_form():
    """Free-space propagation against the closed-form 3D solution: an initial Gaussian
    pressure g(r) at rest gives r p(r, t) = [F(r - ct) + F(r + ct)] / 2 with F(s) = s g(s)
    (odd extension). Receivers 16-17 cells away along an axis, a face diagonal and a body
    diagonal of a 72^3 rigid box, read before any wall reflection returns. Phase, 1/r
    spreading and isotropy are all in the waveform error, which falls at high order with
    the pulse width (WENO5 in space, SSP-RK3 in time): measured 1.8e-2 / 4.5e-3 at
    sigma = 3 / 4 cells (about 4.7th order)."""
    n, c = 72, 36
    recs = [(c + 16, c, c), (c + 10, c + 10, c + 10), (c + 14, c + 9, c)]
    errs = {}
    for sigma in (3.0, 4.0):
        dist = [float(np.linalg.norm(np.subtract(r, (c, c, c)))) for r in recs]
        nsteps = int(2 * (max(dist) + 5 * sigma)) + 2
        wl = W.rigid_box(room=(n, n, n), h=0.05, nsteps=nsteps, ic_center=(c, c, c), rec=recs[0])
        wl.ic = dict(wl.ic, sigma=sigma)
        wl.rec_ijk = [tuple(v + 3 for v in r) for r in recs]
        ir = O.run_workload(wl)["ir"]
        t = (np.arange(nsteps) + 1) * wl.params.c * wl.params.dt / wl.h          # cells travelled
        g = lambda s: np.exp(-s * s / (2 * sigma * sigma))
        for k, r in enumerate(dist):
            ex = ((r - t) * g(r - t) + (r + t) * g(r + t)) / (2 * r)
            w = np.abs(ex) > 1e-3 * np.abs(ex).max()
            errs[(sigma, k)] = np.linalg.norm((ir[k] - ex)[w]) / np.linalg.norm(ex[w])
            assert abs(ir[k].max() / ex.max() - 1) < (0.02 if sigma == 3.0 else 0.005), (sigma, k)
    for k in range(len(recs)):
        assert errs[(4.0, k)] < 8e-3, errs
        order = np.log(errs[(3.0, k)] / errs[(4.0, k)]) / np.log(4.0 / 3.0)
        assert order > 3.5, (k, order, errs)
    # isotropy: the axis and diagonal errors stay within a factor 2 of each other
    e4 = [errs[(4.0, k)] for k in range(len(recs))]
    assert max(e4) < 2 * min(e4), e4
