// Discrete adjoint (reverse mode) of the FVM step: vector-Jacobian products of the
// WENO5-Z pair, the Godunov flux, the SSP-RK3 stage combination, the immersed boundary
// blend and the ADE recursion, written by hand from the forward formulas.
//
// Forward stage (functions3D.cuh:1086-1327 for the flux stages, :462-565 for the
// boundary): out = alpha U0 + beta (V + dt L(V)) on updated cells, then the source
// add, then on boundary cells out = (1 - chi) out + chi b(probe(V), ADE). Reverse order
// per stage: boundary (scales the output adjoint by 1 - chi on boundary cells, emits
// the probe adjoint and the ADE state adjoints), source gradients, flux.
#pragma once

#include "fvm_kernels.hpp"

namespace fasesdev {

// Reverse mode of weno_pair: given the adjoints mA of A = V^-_{i+1/2} and mB of
// B = V^+_{i-1/2} of the cell with window (a,b,c,d,e), returns the adjoints of the five
// window values in g[0..4]. The forward quantities are recomputed in the same product
// form (A = c + nl/dl, B = c + nr/dr with g_r = (e_r + tau) prod_{s != r} e_s).
template <typename R>
__device__ inline void weno_pair_vjp(R a, R b, R c, R d, R e, R mA, R mB, R g[5]) {
    const R d1 = b - a, d2 = c - b, d3 = d - c, d4 = e - d;
    const R s0 = d4 - d3, s1 = d3 - d2, s2 = d2 - d1;
    const R t0 = fma(R(-3), d3, d4), t1 = d2 + d3, t2 = fma(R(3), d2, -d1);
    const R k = R(13) / R(3);
    const R e0 = fma(k * s0, s0, fma(t0, t0, R(4e-10)));
    const R e1 = fma(k * s1, s1, fma(t1, t1, R(4e-10)));
    const R e2 = fma(k * s2, s2, fma(t2, t2, R(4e-10)));
    const R df = e0 - e2;
    const R tau = fabs(df);
    const R p12 = e1 * e2, p02 = e0 * e2, p01 = e0 * e1;
    const R g0 = (e0 + tau) * p12, g1 = (e1 + tau) * p02, g2 = (e2 + tau) * p01;
    const R a0 = fma(R(0.2), d3, R(-0.05) * d4);
    const R a1 = fma(R(0.1), d2, R(0.2) * d3);
    const R a2 = fma(R(1.0 / 12.0), d2, R(-1.0 / 30.0) * d1);
    const R b0 = fma(R(-1.0 / 12.0), d3, R(1.0 / 30.0) * d4);
    const R b1 = fma(R(-0.2), d2, R(-0.1) * d3);
    const R b2 = fma(R(-0.2), d2, R(0.05) * d1);
    const R nl = fma(g0, a0, fma(g1, a1, g2 * a2));
    const R nr = fma(g0, b0, fma(g1, b1, g2 * b2));
    const R dl = fma(R(0.3), g0, fma(R(0.6), g1, R(0.1) * g2));
    const R dr = fma(R(0.1), g0, fma(R(0.6), g1, R(0.3) * g2));
    const R rl = fast_rcp(dl), rr = fast_rcp(dr);
    // A = c + nl rl, B = c + nr rr
    const R mnl = mA * rl, mnr = mB * rr;
    const R mdl = -mnl * (nl * rl), mdr = -mnr * (nr * rr);
    // g_r enter nl, nr, dl, dr
    const R mg0 = fma(mnl, a0, fma(mnr, b0, fma(R(0.3), mdl, R(0.1) * mdr)));
    const R mg1 = fma(mnl, a1, fma(mnr, b1, R(0.6) * (mdl + mdr)));
    const R mg2 = fma(mnl, a2, fma(mnr, b2, fma(R(0.1), mdl, R(0.3) * mdr)));
    // candidate offsets (linear in the differences)
    const R ma0 = mnl * g0, ma1 = mnl * g1, ma2 = mnl * g2;
    const R mb0 = mnr * g0, mb1 = mnr * g1, mb2 = mnr * g2;
    R md1 = fma(R(-1.0 / 30.0), ma2, R(0.05) * mb2);
    R md2 = fma(R(0.1), ma1, fma(R(1.0 / 12.0), ma2, R(-0.2) * (mb1 + mb2)));
    R md3 = fma(R(0.2), ma0 + ma1, fma(R(-1.0 / 12.0), mb0, R(-0.1) * mb1));
    R md4 = fma(R(-0.05), ma0, R(1.0 / 30.0) * mb0);
    // g_r = (e_r + tau) prod_{s != r} e_s
    const R mtau = fma(mg0, p12, fma(mg1, p02, mg2 * p01));
    R me0 = fma(mg0, p12, fma(mg1, (e1 + tau) * e2, mg2 * ((e2 + tau) * e1)));
    const R me1 = fma(mg0, (e0 + tau) * e2, fma(mg1, p02, mg2 * ((e2 + tau) * e0)));
    R me2 = fma(mg0, (e0 + tau) * e1, fma(mg1, (e1 + tau) * e0, mg2 * p01));
    const R sg = df > R(0) ? R(1) : (df < R(0) ? R(-1) : R(0));
    me0 = fma(sg, mtau, me0);
    me2 = fma(-sg, mtau, me2);
    // e_r = k s_r^2 + t_r^2 + eps
    const R ms0 = R(2) * k * s0 * me0, ms1 = R(2) * k * s1 * me1, ms2 = R(2) * k * s2 * me2;
    const R mt0 = R(2) * t0 * me0, mt1 = R(2) * t1 * me1, mt2 = R(2) * t2 * me2;
    md4 += ms0 + mt0;
    md3 += fma(R(-3), mt0, ms1 - ms0 + mt1);
    md2 += fma(R(3), mt2, ms2 - ms1 + mt1);
    md1 += -ms2 - mt2;
    // differences of the window; c also carries A and B directly
    g[0] = -md1;
    g[1] = md1 - md2;
    g[2] = md2 - md3 + mA + mB;
    g[3] = md3 - md4;
    g[4] = md4;
}

// Adjoints of the two states of one face from the adjoints of its fluxes
// (Fp = kp (uR + uL) - kc (pR - pL), Fu = kr (pR + pL) - kc (uR - uL)).
template <typename R>
__device__ inline void godunov_vjp(R mFp, R mFu, R kp, R kc, R kr, R &mpL, R &mpR, R &muL, R &muR) {
    mpL = fma(kc, mFp, kr * mFu);
    mpR = fma(-kc, mFp, kr * mFu);
    muL = fma(kp, mFp, kc * mFu);
    muR = fma(kp, mFp, -kc * mFu);
}

template <typename R> struct AdjK {
    R kp, kc, kr;       // rho c^2 / 2, c / 2, 1 / (2 rho)
    R bdh;              // beta dt / h of the stage (scales the output adjoint of updated cells)
    R alpha, beta;      // stage combination out = alpha U0 + beta (V + dt L)
};

// Flux-stage adjoint, one thread per cell of [3, N-3)^3:
//   lin  (=|+=) beta Lout + (flux part)       lbase (=|+=) alpha Lout
// Lout is the adjoint of the stage output after the boundary adjoint; only updated
// cells (upd != 0) carry it into the stage. ACC_IN / ACC_BASE select += over =.
template <typename S, typename R, bool ACC_IN, bool ACC_BASE, bool HAS_BASE>
__global__ void __launch_bounds__(256)
k_flux_adj_cell(Grid g, FieldSet<S> V, Scl<R> sc, const R *__restrict__ L0, const R *__restrict__ L1,
                const R *__restrict__ L2, const R *__restrict__ L3, const uint8_t *__restrict__ upd, R *lin0,
                R *lin1, R *lin2, R *lin3, R *lb0, R *lb1, R *lb2, R *lb3, AdjK<R> K) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 3;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 3;
    const int k = blockIdx.z + 3;
    if (i >= g.Nx - 3 || j >= g.Ny - 3 || k >= g.Nz - 3) return;
    const long long C = (long long)i + (long long)g.Nx * j + g.nxy * k;
    const R *Lf[4] = {L0, L1, L2, L3};
    const S *Vf[4] = {V.p, V.u, V.v, V.w};
    // masked, scaled output adjoint of a cell
    auto lt = [&](int f, long long q) -> R { return upd[q] ? K.bdh * Lf[f][q] : R(0); };
    auto vv = [&](int f, long long q) -> R { return Conv<S, R>::ld(Vf[f][q], sc.s[f]); };
    R acc[4];
    const bool me = upd[C] != 0;
#pragma unroll
    for (int f = 0; f < 4; ++f) acc[f] = me ? K.beta * Lf[f][C] : R(0);
    const long long st[3] = {1, (long long)g.Nx, g.nxy};
    const int pos[3] = {i, j, k};
    const int nax[3] = {g.Nx, g.Ny, g.Nz};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int fu = 1 + ax;
        const long long s = st[ax];
        // cells Cq = C - q s, q = -2..2, whose pair reads C at window slot q + 2
#pragma unroll
        for (int q = -2; q <= 2; ++q) {
            const int cpos = pos[ax] - q;
            if (cpos < 2 || cpos > nax[ax] - 3) continue;
            const long long Cq = C - q * s;
            // face adjoints: face Cq+1/2 between Cq and Cq+1, face Cq-1/2 between Cq-1 and Cq
            const R lp0 = lt(0, Cq), lu0 = lt(fu, Cq);
            const R mFpP = lt(0, Cq + s) - lp0, mFuP = lt(fu, Cq + s) - lu0;
            const R mFpM = lp0 - lt(0, Cq - s), mFuM = lu0 - lt(fu, Cq - s);
            R mpL, mpR, muL, muR, tpL, tpR, tuL, tuR;
            godunov_vjp(mFpP, mFuP, K.kp, K.kc, K.kr, mpL, tpR, muL, tuR);   // A states of Cq
            godunov_vjp(mFpM, mFuM, K.kp, K.kc, K.kr, tpL, mpR, tuL, muR);   // B states of Cq
            R gp[5], gu[5];
            weno_pair_vjp(vv(0, Cq - 2 * s), vv(0, Cq - s), vv(0, Cq), vv(0, Cq + s), vv(0, Cq + 2 * s), mpL, mpR, gp);
            weno_pair_vjp(vv(fu, Cq - 2 * s), vv(fu, Cq - s), vv(fu, Cq), vv(fu, Cq + s), vv(fu, Cq + 2 * s), muL,
                          muR, gu);
            acc[0] += gp[q + 2];
            acc[fu] += gu[q + 2];
        }
    }
    R *li[4] = {lin0, lin1, lin2, lin3};
#pragma unroll
    for (int f = 0; f < 4; ++f) li[f][C] = ACC_IN ? li[f][C] + acc[f] : acc[f];
    if (HAS_BASE) {
        R *lb[4] = {lb0, lb1, lb2, lb3};
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const R v = me ? K.alpha * Lf[f][C] : R(0);
            lb[f][C] = ACC_BASE ? lb[f][C] + v : v;
        }
    }
}

// ADE replay of one step on the boundary list: x1 = stage 0 (x0, wout(V0)),
// x2 = stage 1 (x0, x1, wout(V1)), the same arithmetic as the forward boundary kernel.
struct AdeState {
    double *phi, *p1, *p2;    // [nB K], [nB L], [nB L]
};

template <typename S>
__device__ inline double probe_wout(const FieldSet<S> &V, const Grid &g, int base, int mw, double sx, double sy,
                                    double sz, double nx, double ny, double nz, double Z0, const Scl<double> &sc) {
#pragma clang fp contract(off)
    typedef typename std::conditional<std::is_same<S, double>::value, double, float>::type R;
    typedef Conv<S, R> CV;
    double pD, uD, vD, wD;
    if (mw & 0x100) {
        pD = (double)CV::ld(V.p[base], (R)sc.s[0]);
        uD = (double)CV::ld(V.u[base], (R)sc.s[1]);
        vD = (double)CV::ld(V.v[base], (R)sc.s[2]);
        wD = (double)CV::ld(V.w[base], (R)sc.s[3]);
    } else {
        pD = tri_probe(V.p, base, 0, sx, sy, sz, g.Nx, g.nxy, sc.s[0]);
        uD = tri_probe(V.u, base, 0, sx, sy, sz, g.Nx, g.nxy, sc.s[1]);
        vD = tri_probe(V.v, base, 0, sx, sy, sz, g.Nx, g.nxy, sc.s[2]);
        wD = tri_probe(V.w, base, 0, sx, sy, sz, g.Nx, g.nxy, sc.s[3]);
    }
    const double un = uD * nx + vD * ny + wD * nz;
    return pD + Z0 * un;
}

template <typename S>
__global__ void __launch_bounds__(256)
k_ade_replay(Grid g, FieldSet<S> V0, FieldSet<S> V1, Bnd<double> bd, Ade ade, double rho, double c, double dt,
             AdeState x0, AdeState x1, AdeState x2, Scl<double> sc) {
#pragma clang fp contract(off)
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= bd.n) return;
    const int base = bd.base[b], mw = bd.mat[b], m = mw & 0xFF;
    const double *q = bd.geo + 8 * (long long)b;
    const double Z0 = rho * c;
    const double w0 = probe_wout(V0, g, base, mw, q[0], q[1], q[2], q[3], q[4], q[5], Z0, sc);
    const double w1 = probe_wout(V1, g, base, mw, q[0], q[1], q[2], q[3], q[4], q[5], Z0, sc);
    for (int kk = 0; kk < ade.Keff[m]; ++kk) {
        const long long s = (long long)b * ade.K + kk;
        const double z = ade.zeta[m * ade.K + kk];
        const double xa = x0.phi[s];
        double xb = 0.0 * xa + 1.0 * (xa + dt * (w0 - z * xa));
        xb /= 1.0;
        double xc = 3.0 * xa + 1.0 * (xb + dt * (w1 - z * xb));
        xc /= 4.0;
        x1.phi[s] = xb;
        x2.phi[s] = xc;
    }
    for (int l = 0; l < ade.Leff[m]; ++l) {
        const long long s = (long long)b * ade.L + l;
        const double al = ade.alpha[m * ade.L + l], be = ade.beta[m * ade.L + l];
        const double ya = x0.p1[s], za = x0.p2[s];
        double yb = 0.0 * ya + 1.0 * (ya + dt * (w0 - (al * ya + be * za)));
        yb /= 1.0;
        double zb = 0.0 * za + 1.0 * (za + dt * (be * ya - al * za));
        zb /= 1.0;
        double yc = 3.0 * ya + 1.0 * (yb + dt * (w1 - (al * yb + be * zb)));
        yc /= 4.0;
        double zc = 3.0 * za + 1.0 * (zb + dt * (be * yb - al * zb));
        zc /= 4.0;
        x1.p1[s] = yb; x1.p2[s] = zb;
        x2.p1[s] = yc; x2.p2[s] = zc;
    }
}

// Boundary + ADE adjoint of one stage (a, b, c) with state buffers mid = x0 and right
// xr; lnew is the adjoint of the stage's new state (complete), lmid / lr accumulate.
// Lout (adjoint of the stage output) is read and replaced by (1 - chi) Lout on the
// boundary cell; lD[4 b + f] receives the adjoint of the probe values (p, u, v, w);
// G[P b + .] accumulates the material-parameter gradient of the cell, layout
// (Rinf, A[K], zeta[K], B[L], C[L], alpha[L], beta[L]).
template <typename S, typename R>
__global__ void __launch_bounds__(256)
k_boundary_adj(Grid g, FieldSet<S> Vin, Bnd<double> bd, Ade ade, double rho, double c, double dt, double as,
               double bs, double cs, AdeState xmid, AdeState xr, AdeState lnew, AdeState lmid, AdeState lr,
               R *L0, R *L1, R *L2, R *L3, double *__restrict__ lD, double *__restrict__ G, int P,
               Scl<double> sc) {
#pragma clang fp contract(off)
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= bd.n) return;
    const int C = bd.cell[b], base = bd.base[b], mw = bd.mat[b], m = mw & 0xFF;
    const double *q = bd.geo + 8 * (long long)b;
    const double nx = q[3], ny = q[4], nz = q[5], chi = q[6];
    const double Z0 = rho * c;
    const double wout = probe_wout(Vin, g, base, mw, q[0], q[1], q[2], nx, ny, nz, Z0, sc);
    const int K = ade.K, L = ade.L;
    double *Gb = G + (long long)P * b;
    // output adjoint of the cell, then its pre-blend share
    const double lp = (double)L0[C], lu = (double)L1[C], lv = (double)L2[C], lw = (double)L3[C];
    L0[C] = (R)((1.0 - chi) * lp);
    L1[C] = (R)((1.0 - chi) * lu);
    L2[C] = (R)((1.0 - chi) * lv);
    L3[C] = (R)((1.0 - chi) * lw);
    const double lpb = chi * lp, lux = chi * lu, luy = chi * lv, luz = chi * lw;
    const double lunb = lux * nx + luy * ny + luz * nz;
    const double lwin = 0.5 * lpb - lunb / (2.0 * Z0);
    double lwout = 0.5 * lpb + lunb / (2.0 * Z0);
    const double Rinf = ade.Rinf[m];
    lwout += Rinf * lwin;
    Gb[0] += lwin * wout;
    const double f = bs * dt / cs;
    for (int kk = 0; kk < ade.Keff[m]; ++kk) {
        const long long s = (long long)b * K + kk;
        const int mi = m * K + kk;
        const double z = ade.zeta[mi];
        const double x_r = xr.phi[s], x_m = xmid.phi[s];
        double xn = as * x_m + bs * (x_r + dt * (wout - z * x_r));
        xn /= cs;
        const double ln = lnew.phi[s] + ade.A[mi] * lwin;
        Gb[1 + kk] += lwin * xn;
        lmid.phi[s] += (as / cs) * ln;
        lr.phi[s] += (bs / cs) * (1.0 - dt * z) * ln;
        lwout += f * ln;
        Gb[1 + K + kk] += -f * x_r * ln;
    }
    for (int l = 0; l < ade.Leff[m]; ++l) {
        const long long s = (long long)b * L + l;
        const int mi = m * L + l;
        const double al = ade.alpha[mi], be = ade.beta[mi];
        const double y1 = xr.p1[s], y2 = xr.p2[s];
        double y1n = as * xmid.p1[s] + bs * (y1 + dt * (wout - (al * y1 + be * y2)));
        y1n /= cs;
        double y2n = as * xmid.p2[s] + bs * (y2 + dt * (be * y1 - al * y2));
        y2n /= cs;
        const double l1 = lnew.p1[s] + ade.B[mi] * lwin, l2 = lnew.p2[s] + ade.C[mi] * lwin;
        Gb[1 + 2 * K + l] += lwin * y1n;
        Gb[1 + 2 * K + L + l] += lwin * y2n;
        lmid.p1[s] += (as / cs) * l1;
        lmid.p2[s] += (as / cs) * l2;
        lr.p1[s] += (bs / cs) * (l1 * (1.0 - dt * al) + l2 * (dt * be));
        lr.p2[s] += (bs / cs) * (l1 * (-dt * be) + l2 * (1.0 - dt * al));
        lwout += f * l1;
        Gb[1 + 2 * K + 2 * L + l] += f * (-y1 * l1 - y2 * l2);
        Gb[1 + 2 * K + 3 * L + l] += f * (-y2 * l1 + y1 * l2);
    }
    // wout = pD + Z0 un, tangential part of the probe velocity kept
    const double lun = Z0 * lwout - lunb;
    double *o = lD + 4 * (long long)b;
    o[0] = lwout;
    o[1] = lux + lun * nx;
    o[2] = luy + lun * ny;
    o[3] = luz + lun * nz;
}

// Probe adjoint gather: every grid cell a probe interpolates from collects the weighted
// probe adjoints of the boundary cells that read it (CSR built on the host).
template <typename R>
__global__ void k_probe_gather(int nT, const int *__restrict__ tgt, const int *__restrict__ off,
                               const int *__restrict__ eb, const double *__restrict__ ew, const double *__restrict__ lD,
                               R *L0, R *L1, R *L2, R *L3) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int e = off[t]; e < off[t + 1]; ++e) {
        const double *d = lD + 4 * (long long)eb[e];
        const double w = ew[e];
        s0 += w * d[0];
        s1 += w * d[1];
        s2 += w * d[2];
        s3 += w * d[3];
    }
    const int C = tgt[t];
    L0[C] = (R)((double)L0[C] + s0);
    L1[C] = (R)((double)L1[C] + s1);
    L2[C] = (R)((double)L2[C] + s2);
    L3[C] = (R)((double)L3[C] + s3);
}

// Source gradients of one stage: d/d sval_s and d/d u_s[nn] of p[cell_s] += coef sval_s amp.
template <typename R>
__global__ void k_source_adj(const R *__restrict__ Lp, const int *__restrict__ cell, const double *__restrict__ sval,
                             const double *__restrict__ u_s, long long us_stride, int n_src, long long nn, double coef,
                             double *__restrict__ gs, double *__restrict__ gus) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_src) return;
    const double lp = (double)Lp[cell[s]];
    const long long idx = (long long)s * us_stride + nn;
    gs[s] += lp * coef * u_s[idx];
    if (gus) atomicAdd(gus + idx, lp * coef * sval[s]);
}

// IR misfit injection: Lp[rec_r] += dJ/dir[r][n] (sequential, receivers may share a cell).
template <typename R>
__global__ void k_receiver_adj(R *Lp, const int *__restrict__ cell, int n_rec, const double *__restrict__ dj,
                               long long stride, long long n) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int r = 0; r < n_rec; ++r) Lp[cell[r]] = (R)((double)Lp[cell[r]] + dj[r * stride + n]);
}

}  // namespace fasesdev
