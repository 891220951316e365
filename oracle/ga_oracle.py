"""ORACLE -- test infrastructure only. Restatement in Python of the geometrical-acoustics
IR synthesis of 3D-solver/core/main3D.cu, used by tests/ as the checker of
fases_ga_band_irs / fases_ga_arrivals:

  R_oblique                 parsing3D.h:1178-1186
  band IRs (HIR / HRT)      main3D.cu:1600-1757: windowed-sinc fractional-delay deposit of
                            the direct sound and of every image source, then the ray-traced
                            energy as 4 pi sign(t) sqrt(e_spec + e_diff)
  arrival lists (ARR)       main3D.cu:1759-1801
  rt_sign                   raytrace3D.h:69-72

Float32 / float64 promotions follow the C++ expressions (float operands, double literals);
libm's cos / sin / exp of float arguments are evaluated in double and rounded once to
float32, which is within an ulp of the C library's float versions.

Pinned (tests/test_ga.py): r_oblique and rt_sign bit for bit against the reference's own
R_oblique / rt_sign built from its sources (oracle/_ref); r_oblique's closed forms
(sqrt(1 - alpha) at normal incidence, -1 at grazing, anechoic and rigid limits,
monotone in the angle); the deposit's closed forms (one sample at an integer delay,
mirror-symmetric taps at a half-sample delay, unit sum); the direct-sound gain
p / d exp(-m d), a normally incident image's p amp sqrt(1 - alpha) sqrt(1 - s) / d
exp(-m d), and the ray-traced 4 pi sign sqrt(e). The arrival list is the same gains and
delays and is pinned through them."""
from __future__ import annotations

import math

import numpy as np

F = np.float32
NW = 32
M32 = 0xFFFFFFFF


def rt_sign(s: int) -> float:
    s &= M32
    s ^= s >> 16
    s = (s * 0x7feb352d) & M32
    s ^= s >> 15
    s = (s * 0x846ca68b) & M32
    s ^= s >> 16
    return 1.0 if (s & 1) else -1.0


def r_oblique(alpha, cos_theta) -> np.float32:
    alpha, ct = F(alpha), F(cos_theta)
    if alpha > F(0.99):
        return F(0.0)
    R = np.sqrt(F(1.0 - float(alpha)))
    if float(R) > 0.99999:
        return R
    xi = F((1.0 + float(R)) / (1.0 - float(R)))
    xc = F(xi * ct)
    return F((float(xc) - 1.0) / (float(xc) + 1.0))


def _f(x):            # float libm function of a float32 argument, rounded to float32
    return lambda v: F(x(float(F(v))))


cosf, sinf, expf = _f(math.cos), _f(math.sin), _f(math.exp)


def sinc(x) -> np.float32:
    x = F(x)
    return F(1.0) if x == 0.0 else F(sinf(x) / x)


def _deposit(row, tau, g, Nt):
    tt = int(tau)
    for n0 in range(-NW // 2, NW // 2 + 1):
        nn = tt + n0
        if nn < 0 or nn >= Nt:
            continue
        xx = F(F(nn) - tau)
        term = float(g) * 0.5 * (1.0 + math.cos(2.0 * math.pi * float(xx) / NW)) * float(sinc(F(math.pi * float(xx))))
        row[nn] = F(float(row[nn]) + term)


def image_gain(ps, im, alpha_gb, sctr, ngroups, band, mair):
    """ps * amp * prod_k R_oblique(alpha, cos angle_k) sqrt(1 - s_k) / d * exp(-m d)."""
    d, amp, groups, angles = F(im["distance"]), F(im["amp"]), im["group"], im["angle"]
    P = F(1.0)
    for gk, th in zip(groups, angles):
        al = alpha_gb[gk * 8 + band] if 0 <= gk < ngroups else F(0.0)
        R = r_oblique(al, cosf(th))
        if 0 <= gk < ngroups:
            R = F(R * np.sqrt(F(max(F(0.0), F(F(1.0) - F(sctr[gk]))))))
        P = F(P * R)
    return F(F(F(F(ps) * amp) * P) / d) * expf(F(-F(mair) * d))


def band_irs(direct, images, src_p, e_spec, e_diff, alpha_gb, sctr, ngroups, m_air, c, fs, Nt, band):
    """(p_hf, p_rt) [nr, Nt] of one band. direct [ns, nr]; images[s][r] = list of dicts
    (distance, amp, group, angle); e_spec / e_diff [8, nr, Nt]."""
    ns, nr = direct.shape
    fs, c = F(fs), F(c)
    mair = F(m_air[band])
    hf = np.zeros((nr, Nt), np.float32)
    rt = np.zeros((nr, Nt), np.float32)
    for r in range(nr):
        for s in range(ns):
            d0 = F(direct[s, r])
            if d0 > F(0.0):
                tau0 = F(F(fs * d0) / c)
                if int(tau0) < Nt:
                    _deposit(hf[r], tau0, F(F(F(src_p[s]) / d0) * expf(F(-mair * d0))), Nt)
            for im in images[s][r]:
                d = F(im["distance"])
                tau = F(F(fs * d) / c)
                if int(tau) >= Nt:
                    continue
                _deposit(hf[r], tau, image_gain(src_p[s], im, alpha_gb, sctr, ngroups, band, mair), Nt)
        for t in range(1, Nt):
            e = F(F(0.0) + e_diff[band, r, t])
            e = F(e + e_spec[band, r, t])
            if e <= 0.0:
                continue
            v = 4.0 * math.pi * rt_sign((t * 2654435761 + r * 40503) & M32) * float(np.sqrt(e))
            hf[r, t] = F(float(hf[r, t]) + v)
            rt[r, t] = F(float(rt[r, t]) + v)
    return hf, rt


def arrivals(direct, images, src_p, src_raw, rec_raw, alpha_gb, sctr, ngroups, m_air, c, fs, Nt, r):
    """[n, 12] arrivals at receiver r: delay, direction receiver -> source (raw frame for the
    direct sound, the image direction otherwise), gain per band."""
    ns = direct.shape[0]
    fs, c = F(fs), F(c)
    out = []
    rx, ry, rz = (F(v) for v in rec_raw[r])
    for s in range(ns):
        d0 = F(direct[s, r])
        if d0 > F(0.0) and int(F(F(fs * d0) / c)) < Nt:
            v = [F(F(src_raw[s][0]) - rx), F(F(src_raw[s][1]) - ry), F(F(src_raw[s][2]) - rz)]
            nrm = np.sqrt(F(F(F(v[0] * v[0]) + F(v[1] * v[1])) + F(v[2] * v[2])))
            if nrm > F(1e-12):
                v = [F(a / nrm) for a in v]
            g = [F(F(F(src_p[s]) / d0) * expf(F(-F(m_air[b]) * d0))) for b in range(8)]
            out.append([F(float(d0) / float(c))] + v + g)
        for im in images[s][r]:
            d = F(im["distance"])
            if int(F(F(fs * d) / c)) >= Nt:
                continue
            g = [image_gain(src_p[s], im, alpha_gb, sctr, ngroups, b, m_air[b]) for b in range(8)]
            out.append([F(float(d) / float(c))] + [F(v) for v in im["doa"]] + g)
    return np.array(out, np.float32).reshape(-1, 12)
