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
    // A = c + sum_r W_r a'_r with normalised weights W_r = d_r g_r / dl (a_r = d_r a'_r):
    // every quantity below stays bounded by the data and 1 / eps. (Expanding through
    // nl / dl^2 and the products of e instead forms terms ~ mA / dl ~ mA / e^3 that cancel
    // to the result: at the eps level that overflows and loses the result to rounding.)
    const R Ao = nl * rl, Bo = nr * rr;                      // A - c, B - c
    const R uA0 = g0 * rl, uA1 = g1 * rl, uA2 = g2 * rl;     // dA/da_r = W_r / d_r
    const R uB0 = g0 * rr, uB1 = g1 * rr, uB2 = g2 * rr;
    const R ma0 = mA * uA0, ma1 = mA * uA1, ma2 = mA * uA2;
    const R mb0 = mB * uB0, mb1 = mB * uB1, mb2 = mB * uB2;
    R md1 = fma(R(-1.0 / 30.0), ma2, R(0.05) * mb2);
    R md2 = fma(R(0.1), ma1, fma(R(1.0 / 12.0), ma2, R(-0.2) * (mb1 + mb2)));
    R md3 = fma(R(0.2), ma0 + ma1, fma(R(-1.0 / 12.0), mb0, R(-0.1) * mb1));
    R md4 = fma(R(-0.05), ma0, R(1.0 / 30.0) * mb0);
    // W_r = d_r (1 + tau/e_r) / sum: dA/de_s = -(a'_s - Ao) W_s tau / (e_s (e_s + tau)),
    // dA/dtau = sum_r (a'_r - Ao) W_r / (e_r + tau); Q_r = mA W_r (a'_r - Ao) + (B side)
    const R Q0 = fma(ma0, fma(R(-0.3), Ao, a0), mb0 * fma(R(-0.1), Bo, b0));
    const R Q1 = fma(ma1, fma(R(-0.6), Ao, a1), mb1 * fma(R(-0.6), Bo, b1));
    const R Q2 = fma(ma2, fma(R(-0.1), Ao, a2), mb2 * fma(R(-0.3), Bo, b2));
    const R R0 = fast_rcp(e0 * (e0 + tau)), R1 = fast_rcp(e1 * (e1 + tau)), R2 = fast_rcp(e2 * (e2 + tau));
    const R mtau = fma(Q0, e0 * R0, fma(Q1, e1 * R1, Q2 * (e2 * R2)));
    R me0 = -Q0 * (tau * R0);
    const R me1 = -Q1 * (tau * R1);
    R me2 = -Q2 * (tau * R2);
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

// Flux-part adjoint of one cell C = (i, j, k): the sum over the three axes and the five
// cells Cq whose WENO pair windows contain C of their pair VJPs at C's slot. vv(f, q) is
// the forward value of field f at cell q, lt(f, q) the scaled, masked output adjoint.
template <typename R, typename VA, typename LA>
__device__ inline void flux_adj_sum(const Grid &g, int i, int j, int k, const VA &vv, const LA &lt, const AdjK<R> &K,
                                    R acc[4]) {
    const long long C = (long long)i + (long long)g.Nx * j + g.nxy * k;
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
}

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
    auto lt = [&](int f, long long q) -> R { return upd[q] ? K.bdh * Lf[f][q] : R(0); };
    auto vv = [&](int f, long long q) -> R { return Conv<S, R>::ld(Vf[f][q], sc.s[f]); };
    const bool me = upd[C] != 0;
    R acc[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) acc[f] = me ? K.beta * Lf[f][C] : R(0);
    flux_adj_sum<R>(g, i, j, k, vv, lt, K, acc);
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

// FP64 adjoint of the cells near point sources (reduced-precision storage): the forward
// updates them in FP64 from FP64 shadows, and so does the adjoint. Reads the forward
// states from the stage input's shadow (forward patch map vidx) and the output adjoint
// from its FP64 adjoint shadow (adjoint patch map pidx; stored copies elsewhere); writes
// the FP64 adjoint shadows of the input (and base) set and their stored copies, after
// the grid kernel has run. Two launches: k_flux_adj_patch_part evaluates, once per cell X
// of the patch dilated by two cells along each axis and per axis, both pair VJPs of X
// (all five window slots, p and the axis velocity) into pv; k_flux_adj_patch then adds,
// for every patch cell C, the slot of C in the pairs of C - q s over the axes and q in the
// per-cell order (bit-identical to a single pass that re-evaluates the pairs).
template <typename S>
struct PatchAdjIn {
    FieldSet<S> V;                 // stage input (stored)
    const double *vsh;             // its FP64 shadows, [4][nv]
    const int *vidx;               // grid -> forward patch index
    int nv;
    Scl<double> sc;
};

// pv layout: [(ax * 10 + m) * nD + d], m < 5 the p-pair window slots, m >= 5 the velocity's
// (slot-major, so that the lanes of a wave, consecutive d, read and write consecutive words).
// Latency-bound gathers: held to 96 VGPRs for five waves per SIMD instead of four (98).
template <typename S, typename R>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5)))
k_flux_adj_patch_part(Grid g, PatchAdjIn<S> in, const R *__restrict__ L0, const R *__restrict__ L1,
                      const R *__restrict__ L2, const R *__restrict__ L3, const double *__restrict__ lsh_out,
                      const uint8_t *__restrict__ upd, const int *__restrict__ dcells, int nD,
                      const int *__restrict__ pidx, int np, AdjK<double> K, double *__restrict__ pv) {
    const long long tt = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (tt >= 3LL * nD) return;
    const int d = (int)(tt % nD), ax = (int)(tt / nD);
    const int X = dcells[d];
    const int pos[3] = {X % g.Nx, (X / g.Nx) % g.Ny, (int)(X / g.nxy)};
    const int nax[3] = {g.Nx, g.Ny, g.Nz};
    const long long s = ax == 0 ? 1LL : (ax == 1 ? (long long)g.Nx : g.nxy);
    const int fu = 1 + ax;
    const R *Lf[4] = {L0, L1, L2, L3};
    const S *Vf[4] = {in.V.p, in.V.u, in.V.v, in.V.w};
    auto lraw = [&](int f, long long x) -> double {
        const int w = pidx[x];
        return w >= 0 ? lsh_out[(long long)f * np + w] : (double)Lf[f][x];
    };
    auto lt = [&](int f, long long x) -> double { return upd[x] ? K.bdh * lraw(f, x) : 0.0; };
    auto vv = [&](int f, long long x) -> double {
        const int w = in.vidx[x];
        return w >= 0 ? in.vsh[(long long)f * in.nv + w] : Conv<S, double>::ld(Vf[f][x], in.sc.s[f]);
    };
    double gp[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, gu[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (pos[ax] >= 2 && pos[ax] <= nax[ax] - 3) {
        const long long Cq = X;
        const double lp0 = lt(0, Cq), lu0 = lt(fu, Cq);
        const double mFpP = lt(0, Cq + s) - lp0, mFuP = lt(fu, Cq + s) - lu0;
        const double mFpM = lp0 - lt(0, Cq - s), mFuM = lu0 - lt(fu, Cq - s);
        double mpL, mpR, muL, muR, tpL, tpR, tuL, tuR;
        godunov_vjp(mFpP, mFuP, K.kp, K.kc, K.kr, mpL, tpR, muL, tuR);
        godunov_vjp(mFpM, mFuM, K.kp, K.kc, K.kr, tpL, mpR, tuL, muR);
        weno_pair_vjp(vv(0, Cq - 2 * s), vv(0, Cq - s), vv(0, Cq), vv(0, Cq + s), vv(0, Cq + 2 * s), mpL, mpR, gp);
        weno_pair_vjp(vv(fu, Cq - 2 * s), vv(fu, Cq - s), vv(fu, Cq), vv(fu, Cq + s), vv(fu, Cq + 2 * s), muL, muR,
                      gu);
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        pv[((long long)ax * 10 + m) * nD + d] = gp[m];
        pv[((long long)ax * 10 + 5 + m) * nD + d] = gu[m];
    }
}

template <typename S, typename R, bool ACC_IN, bool ACC_BASE, bool HAS_BASE>
__global__ void __launch_bounds__(64)
k_flux_adj_patch(Grid g, const R *__restrict__ L0, const R *__restrict__ L1, const R *__restrict__ L2,
                 const R *__restrict__ L3, const double *__restrict__ lsh_out, const int *__restrict__ cells,
                 const int *__restrict__ pidx, int np, const int *__restrict__ dmap, int nD,
                 const double *__restrict__ pv, R *lin0, R *lin1, R *lin2, R *lin3, double *lsh_in, R *lb0, R *lb1,
                 R *lb2, R *lb3, double *lsh_base, AdjK<double> K) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    const int C = cells[t];
    const int pos[3] = {C % g.Nx, (C / g.Nx) % g.Ny, (int)(C / g.nxy)};
    const int nax[3] = {g.Nx, g.Ny, g.Nz};
    const R *Lf[4] = {L0, L1, L2, L3};
    auto lraw = [&](int f, long long x) -> double {
        const int w = pidx[x];
        return w >= 0 ? lsh_out[(long long)f * np + w] : (double)Lf[f][x];
    };
    double acc[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) acc[f] = K.beta * lraw(f, C);
#pragma unroll
    for (int combo = 0; combo < 15; ++combo) {       // ax-major, q = -2..2: the per-cell order
        const int ax = combo / 5, q = combo % 5 - 2;
        double cp = 0.0, cu = 0.0;
        const int cpos = pos[ax] - q;
        if (cpos >= 2 && cpos <= nax[ax] - 3) {
            const long long s = ax == 0 ? 1LL : (ax == 1 ? (long long)g.Nx : g.nxy);
            const long long dd = dmap[C - q * s];
            cp = pv[((long long)ax * 10 + q + 2) * nD + dd];
            cu = pv[((long long)ax * 10 + 5 + q + 2) * nD + dd];
        }
        acc[0] += cp;
        acc[1 + ax] += cu;
    }
    R *li[4] = {lin0, lin1, lin2, lin3};
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        double *o = lsh_in + (long long)f * np + t;
        const double v = ACC_IN ? *o + acc[f] : acc[f];
        *o = v;
        li[f][C] = (R)v;
    }
    if (HAS_BASE) {
        R *lb[4] = {lb0, lb1, lb2, lb3};
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            double *o = lsh_base + (long long)f * np + t;
            const double v = (ACC_BASE ? *o : 0.0) + K.alpha * lraw(f, C);
            *o = v;
            lb[f][C] = (R)v;
        }
    }
}

// Tiled flux-stage adjoint. Block = 64 lanes x NW waves; lane l holds column X0 + l and
// wave w row Y0 - 2 + w, marching over planes in z. Every cell's two WENO pair VJPs per
// axis are evaluated once; their five window contributions reach their cells through lane
// shuffles (x), an LDS exchange (y) and a ring of five plane accumulators (z). Outputs:
// all 64 columns x the inner NW - 4 rows, plane k - 2 once the pipeline has passed plane
// k + 2, i.e. two planes behind the march. The x-pair VJPs of the four cells beyond the
// tile's x edges (X0-2, X0-1, X0+64, X0+65) that reach its outputs are evaluated by the
// four y-halo waves (which have no x or z work) and handed to lanes 0, 1, 62, 63 through
// LDS, in the same place of the same sum.
namespace adjtile {
constexpr int TX = 64;       // output columns per block
constexpr int VX = 72;       // V tile columns: X0-4 .. X0+67
constexpr int LXA = 70;      // adjoint tile columns: X0-3 .. X0+66
}  // namespace adjtile

template <typename R, int NW> struct AdjLds {
    R vp[NW + 4][adjtile::VX];   // rows Y0-4 .. Y0+NW-1
    R vu[NW][adjtile::VX];
    R vv[NW + 4][64];
    R lp[NW + 2][adjtile::LXA];  // rows Y0-3 .. Y0+NW-2 (scaled, masked output adjoint)
    R lu[NW][adjtile::LXA];
    R lv[NW + 2][64];
    R gy[2][5][NW][64];          // y-pair contributions of every row: p, v
    // x-edge contributions per output row (p, u): gp4(X0-2), gp3(X0-1), gp4(X0-1),
    // gp0(X0+64), gp1(X0+64), gp0(X0+65)
    R xe[2][NW][6];
    R spare[64];                 // target of the halo slots that carry no item
};

template <typename S, typename R, bool ACC_IN, bool ACC_BASE, bool HAS_BASE, int NW>
__global__ void __launch_bounds__(64 * NW)
k_flux_adj_tiled(int Nx, int Ny, int Nz, FieldSet<S> V, Scl<R> sc, const R *__restrict__ L0,
                 const R *__restrict__ L1, const R *__restrict__ L2, const R *__restrict__ L3,
                 const uint8_t *__restrict__ upd, R *lin0, R *lin1, R *lin2, R *lin3, R *lb0, R *lb1, R *lb2,
                 R *lb3, AdjK<R> K, int tz, int nbx, int nby) {
    using namespace adjtile;
    __shared__ AdjLds<R, NW> T;
    typedef Conv<S, R> CV;
    const int lx = threadIdx.x, wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int tid = wy * 64 + lx;
    // tile of this block: the grid's own (bx, by, bz), or, on a 1-D grid (nbx > 0), the
    // XCD-aware order of the forward kernel (each XCD walks a contiguous tile range)
    int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
    if (nbx > 0) {
        const unsigned t = xcd_tile(blockIdx.x, gridDim.x);
        bx = (int)(t % (unsigned)nbx);
        by = (int)((t / (unsigned)nbx) % (unsigned)nby);
        bz = (int)(t / ((unsigned)nbx * (unsigned)nby));
    }
    const int X0 = 3 + TX * bx, Y0 = 3 + (NW - 4) * by, k0 = 3 + tz * bz;
    const int kend = min(k0 + tz, Nz - 3);
    const long long nxy = (long long)Nx * Ny;
    const int x = X0 + lx, y = Y0 - 2 + wy;
    const int xc = max(0, min(x, Nx - 1)), yc = max(0, min(y, Ny - 1));
    const long long col = xc + (long long)Nx * yc;
    const bool out_row = (wy >= 2 && wy <= NW - 3) && y < Ny - 3;
    const bool out_cell = out_row && x < Nx - 3;
    const bool halo_wave = wy < 2 || wy > NW - 3;
    const R bdh = K.bdh;
    auto plane = [&](int k) -> long long { return nxy * (long long)max(0, min(k, Nz - 1)); };

    // halo items of this thread: items tid, tid + 64 NW, ... (see the tile shapes above):
    // vp rows 0, 1, NW+2, NW+3 and the four side columns of each end of the own rows,
    // vu the same side columns, vv rows 0, 1, NW+2, NW+3; lp rows 0, NW+1 and three side
    // columns of each end of the own rows, lu the same side columns, lv rows 0, NW+1.
    // Each kind's items start on a multiple of 64, so the kind of a slot is uniform over a
    // wave (kind 6: no item); a slot keeps only its LDS target and its offset in a plane, and
    // a padding slot of a kind loads cell 0 of the plane into the spare row.
    constexpr int n0 = 4 * VX + 8 * NW, n1 = 8 * NW, n2 = 4 * 64, n3 = 2 * LXA + 6 * NW, n4 = 6 * NW, n5 = 2 * 64;
    constexpr int e0 = (n0 + 63) / 64 * 64, e1 = e0 + (n1 + 63) / 64 * 64, e2 = e1 + n2,
                  e3 = e2 + (n3 + 63) / 64 * 64, e4 = e3 + (n4 + 63) / 64 * 64, e5 = e4 + n5;
    constexpr int NH = (e5 + 64 * NW - 1) / (64 * NW);
    int h_kind[NH];             // wave-uniform
    R *h_dst[NH];
    unsigned h_off[NH];         // element offset of the item's cell within a plane
#pragma unroll
    for (int hq = 0; hq < NH; ++hq) {
        const int t0 = tid + hq * 64 * NW;
        const int tw = __builtin_amdgcn_readfirstlane(t0);
        const int kind = tw < e0 ? 0 : tw < e1 ? 1 : tw < e2 ? 2 : tw < e3 ? 3 : tw < e4 ? 4 : tw < e5 ? 5 : 6;
        const int rA[4] = {0, 1, NW + 2, NW + 3};
        const int cA[8] = {0, 1, 2, 3, VX - 4, VX - 3, VX - 2, VX - 1};
        const int cL[6] = {0, 1, 2, LXA - 3, LXA - 2, LXA - 1};
        int r = -1, c = 0, t;
        switch (kind) {
        case 0:
            t = t0;
            if (t < 4 * VX) { r = rA[t / VX]; c = t % VX; }
            else if ((t -= 4 * VX) < 8 * NW) { c = cA[t / NW]; r = 2 + t % NW; }
            break;
        case 1: t = t0 - e0; if (t < n1) { c = cA[t / NW]; r = t % NW; } break;
        case 2: t = t0 - e1; r = rA[t / 64]; c = t % 64; break;
        case 3:
            t = t0 - e2;
            if (t < 2 * LXA) { r = (t / LXA) ? NW + 1 : 0; c = t % LXA; }
            else if ((t -= 2 * LXA) < 6 * NW) { c = cL[t / NW]; r = 1 + t % NW; }
            break;
        case 4: t = t0 - e3; if (t < n4) { c = cL[t / NW]; r = t % NW; } break;
        case 5: t = t0 - e4; r = (t / 64) ? NW + 1 : 0; c = t % 64; break;
        default: break;
        }
        h_kind[hq] = kind;
        h_dst[hq] = &T.spare[lx];
        h_off[hq] = 0;
        if (r >= 0) {
            int gx, gy;
            if (kind <= 2) {            // V tiles: vp (r, c) -> (Y0-4+r, X0-4+c); vu rows are wy; vv cols are lanes
                gx = (kind == 2) ? X0 + c : X0 - 4 + c;
                gy = (kind == 1) ? Y0 - 2 + r : Y0 - 4 + r;
            } else {                    // adjoint tiles: lp (r, c) -> (Y0-3+r, X0-3+c)
                gx = (kind == 5) ? X0 + c : X0 - 3 + c;
                gy = (kind == 4) ? Y0 - 2 + r : Y0 - 3 + r;
            }
            gx = max(0, min(gx, Nx - 1));
            gy = max(0, min(gy, Ny - 1));
            h_off[hq] = (unsigned)(gx + Nx * gy);
            h_dst[hq] = kind == 0 ? &T.vp[r][c] : kind == 1 ? &T.vu[r][c] : kind == 2 ? &T.vv[r][c]
                      : kind == 3 ? &T.lp[r][c] : kind == 4 ? &T.lu[r][c] : &T.lv[r][c];
        }
    }

    // Plane-relative access of the own column (buffer descriptors: scalar plane base,
    // 32-bit lane offset; planes clamped to the grid like the halo items).
    const unsigned cS = (unsigned)col * (unsigned)sizeof(S), cR = (unsigned)col * (unsigned)sizeof(R);
    const unsigned pS = (unsigned)nxy * (unsigned)sizeof(S), pR = (unsigned)nxy * (unsigned)sizeof(R);
    auto vld = [&](const S *f, int kk, R scale) __device__ -> R {
        return CV::ld(PlaneIO<S>::ld(plane_rsrc(f + plane(kk), pS), cS), scale);
    };
    auto rld = [&](const R *a, long long pl) __device__ -> R { return PlaneIO<R>::ld(plane_rsrc(a + pl, pR), cR); };
    auto rst = [&](R *a, long long pl, R x) __device__ { PlaneIO<R>::st(plane_rsrc(a + pl, pR), cR, x); };
    auto uld = [&](long long pl) __device__ -> unsigned { return plane_ld_u8(upd + pl, (unsigned)nxy, (unsigned)col); };
    auto ltp = [&](const R *a, long long pl) __device__ -> R { return uld(pl) ? bdh * rld(a, pl) : R(0); };

    // Rings of five slots indexed at compile time (slot of logical entry q at phase PH:
    // (q + PH) % 5), so advancing a plane moves no registers:
    //   zp, zw: V p, w of the own column at planes k-2 .. k+2;
    //   ap, aw: scaled output adjoint at k-1 .. k+1 (three live slots);
    //   za, zb: z contributions to planes k-2 .. k+2 (p, w);
    //   xa:     x/y contributions to planes k-2 .. k (p, u, v; three live slots).
    R zp[5], zw[5], ap[5], aw[5];
    R za[5] = {0, 0, 0, 0, 0}, zb[5] = {0, 0, 0, 0, 0};
    R xa[5][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    {
        const int kb = k0 - 2;
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            zp[q] = vld(V.p, kb - 2 + q, sc.s[0]);
            zw[q] = vld(V.w, kb - 2 + q, sc.s[3]);
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int kq = kb - 1 + q;
            const bool inside = kq >= 0 && kq < Nz;
            ap[q] = inside ? ltp(L0, nxy * (long long)kq) : R(0);
            aw[q] = inside ? ltp(L3, nxy * (long long)kq) : R(0);
        }
        ap[3] = ap[4] = aw[3] = aw[4] = R(0);
    }
    // Plane k's own-column values and halo items, loaded a step ahead (raw, with the update
    // flag where the value is an output adjoint) so that their latency hides behind a whole
    // step of work instead of stalling the LDS fill before the first barrier.
    S pf_vu = S(), pf_vv = S();
    R pf_lu = R(0), pf_lv = R(0);
    unsigned pf_f = 0;
    S pf_hs[NH];
    R pf_hr[NH];
    unsigned pf_hf[NH];
#pragma unroll
    for (int hq = 0; hq < NH; ++hq) { pf_hs[hq] = S(); pf_hr[hq] = R(0); pf_hf[hq] = 0; }
    auto prefetch = [&](int kk) __device__ __attribute__((always_inline)) {
        const long long pk = nxy * (long long)kk;
        pf_vu = PlaneIO<S>::ld(plane_rsrc(V.u + pk, pS), cS);
        pf_vv = PlaneIO<S>::ld(plane_rsrc(V.v + pk, pS), cS);
        pf_f = uld(pk);
        pf_lu = rld(L1, pk);
        pf_lv = rld(L2, pk);
#pragma unroll
        for (int hq = 0; hq < NH; ++hq) {
            const long long q = pk + h_off[hq];
            switch (h_kind[hq]) {             // wave-uniform
            case 0: pf_hs[hq] = V.p[q]; break;
            case 1: pf_hs[hq] = V.u[q]; break;
            case 2: pf_hs[hq] = V.v[q]; break;
            case 3: pf_hr[hq] = L0[q]; pf_hf[hq] = upd[q]; break;
            case 4: pf_hr[hq] = L1[q]; pf_hf[hq] = upd[q]; break;
            case 5: pf_hr[hq] = L2[q]; pf_hf[hq] = upd[q]; break;
            default: break;
            }
        }
    };
    if (k0 < kend) prefetch(k0);
    auto step = [&](int k, auto phase) __device__ __attribute__((always_inline)) {
        constexpr int PH = decltype(phase)::value;
        constexpr int Z0 = PH % 5, Z1 = (PH + 1) % 5, Z2 = (PH + 2) % 5, Z3 = (PH + 3) % 5, Z4 = (PH + 4) % 5;
        const bool xy = (k >= k0 && k < kend);                 // block-uniform
        // the values the completion of plane k - 2 reads, loaded at the start of the step
        const int m = k - 2;
        const bool done = out_cell && m >= k0 && m < kend;
        const long long pm = nxy * (long long)m;
        bool me = false;
        R o0 = 0, o1 = 0, o2 = 0, o3 = 0, i0 = 0, i1 = 0, i2 = 0, i3 = 0, c0 = 0, c1 = 0, c2 = 0, c3 = 0;
        if (done) {
            me = uld(pm) != 0;
            o0 = rld(L0, pm); o1 = rld(L1, pm); o2 = rld(L2, pm); o3 = rld(L3, pm);
            if (ACC_IN) { i0 = rld(lin0, pm); i1 = rld(lin1, pm); i2 = rld(lin2, pm); i3 = rld(lin3, pm); }
            if (HAS_BASE && ACC_BASE) { c0 = rld(lb0, pm); c1 = rld(lb1, pm); c2 = rld(lb2, pm); c3 = rld(lb3, pm); }
        }
        // lead loads for the next plane's windows
        const R np_ = vld(V.p, k + 3, sc.s[0]);
        const R nw_ = vld(V.w, k + 3, sc.s[3]);
        R nap = 0, naw = 0;
        if (k + 2 < Nz) {
            const long long p2 = nxy * (long long)(k + 2);
            nap = ltp(L0, p2);
            naw = ltp(L3, p2);
        }
        if (xy) {
            T.vp[wy + 2][lx + 4] = zp[Z2];
            T.vu[wy][lx + 4] = CV::ld(pf_vu, sc.s[1]);
            T.vv[wy + 2][lx] = CV::ld(pf_vv, sc.s[2]);
            T.lp[wy + 1][lx + 3] = ap[Z1];
            T.lu[wy][lx + 3] = pf_f ? bdh * pf_lu : R(0);
            T.lv[wy + 1][lx] = pf_f ? bdh * pf_lv : R(0);
#pragma unroll
            for (int hq = 0; hq < NH; ++hq) {
                switch (h_kind[hq]) {             // wave-uniform
                case 0: *h_dst[hq] = CV::ld(pf_hs[hq], sc.s[0]); break;
                case 1: *h_dst[hq] = CV::ld(pf_hs[hq], sc.s[1]); break;
                case 2: *h_dst[hq] = CV::ld(pf_hs[hq], sc.s[2]); break;
                case 3: case 4: case 5: *h_dst[hq] = pf_hf[hq] ? bdh * pf_hr[hq] : R(0); break;
                default: break;
                }
            }
            if (k + 1 < kend) prefetch(k + 1);
            __syncthreads();
            // ---- y pairs of every row
            {
                const R a_m = T.lp[wy][lx + 3], a_c = T.lp[wy + 1][lx + 3], a_p = T.lp[wy + 2][lx + 3];
                const R b_m = T.lv[wy][lx], b_c = T.lv[wy + 1][lx], b_p = T.lv[wy + 2][lx];
                R mpL, mpR, mvL, mvR, t0, t1, t2, t3;
                godunov_vjp(a_p - a_c, b_p - b_c, K.kp, K.kc, K.kr, mpL, t0, mvL, t1);
                godunov_vjp(a_c - a_m, b_c - b_m, K.kp, K.kc, K.kr, t2, mpR, t3, mvR);
                R gp[5], gv[5];
                weno_pair_vjp(T.vp[wy][lx + 4], T.vp[wy + 1][lx + 4], T.vp[wy + 2][lx + 4], T.vp[wy + 3][lx + 4],
                              T.vp[wy + 4][lx + 4], mpL, mpR, gp);
                weno_pair_vjp(T.vv[wy][lx], T.vv[wy + 1][lx], T.vv[wy + 2][lx], T.vv[wy + 3][lx], T.vv[wy + 4][lx],
                              mvL, mvR, gv);
#pragma unroll
                for (int q = 0; q < 5; ++q) {
                    T.gy[0][q][wy][lx] = gp[q];
                    T.gy[1][q][wy][lx] = gv[q];
                }
            }
            // ---- x pairs of the output rows (all lanes of those waves); the four y-halo
            // waves evaluate the x pairs of the cells beyond the tile's x edges
            R xp[5] = {0, 0, 0, 0, 0}, xu[5] = {0, 0, 0, 0, 0};   // own g[2], then the neighbours' terms
            if (out_row) {
                const R a_m = T.lp[wy + 1][lx + 2], a_c = T.lp[wy + 1][lx + 3], a_p = T.lp[wy + 1][lx + 4];
                const R b_m = T.lu[wy][lx + 2], b_c = T.lu[wy][lx + 3], b_p = T.lu[wy][lx + 4];
                R mpL, mpR, muL, muR, t0, t1, t2, t3;
                godunov_vjp(a_p - a_c, b_p - b_c, K.kp, K.kc, K.kr, mpL, t0, muL, t1);
                godunov_vjp(a_c - a_m, b_c - b_m, K.kp, K.kc, K.kr, t2, mpR, t3, muR);
                R gp[5], gu[5];
                weno_pair_vjp(T.vp[wy + 2][lx + 2], T.vp[wy + 2][lx + 3], T.vp[wy + 2][lx + 4], T.vp[wy + 2][lx + 5],
                              T.vp[wy + 2][lx + 6], mpL, mpR, gp);
                weno_pair_vjp(T.vu[wy][lx + 2], T.vu[wy][lx + 3], T.vu[wy][lx + 4], T.vu[wy][lx + 5], T.vu[wy][lx + 6],
                              muL, muR, gu);
                xp[0] = gp[2];
                xp[1] = __shfl_up(gp[3], 1, 64);
                xp[2] = __shfl_up(gp[4], 2, 64);
                xp[3] = __shfl_down(gp[1], 1, 64);
                xp[4] = __shfl_down(gp[0], 2, 64);
                xu[0] = gu[2];
                xu[1] = __shfl_up(gu[3], 1, 64);
                xu[2] = __shfl_up(gu[4], 2, 64);
                xu[3] = __shfl_down(gu[1], 1, 64);
                xu[4] = __shfl_down(gu[0], 2, 64);
            } else if (halo_wave) {
                constexpr int NO = NW - 4;                       // output rows
                const int h = wy < 2 ? wy : wy - (NW - 4);      // 0..3
                if (lx < 2 * NO) {
                    // item t: var (p, u), edge cell e (X0-2, X0-1, X0+64, X0+65), output row r
                    const int t = h * 2 * NO + lx;
                    const int var = t / (4 * NO), e = (t % (4 * NO)) / NO, r = t % NO;
                    const int dx = e < 2 ? e - 2 : 62 + e;
                    const int w = r + 2;                         // the output row's wave
                    const R *lpr = &T.lp[w + 1][dx + 2], *lur = &T.lu[w][dx + 2];
                    R mpL, mpR, muL, muR, t0, t1, t2, t3;
                    godunov_vjp(lpr[2] - lpr[1], lur[2] - lur[1], K.kp, K.kc, K.kr, mpL, t0, muL, t1);
                    godunov_vjp(lpr[1] - lpr[0], lur[1] - lur[0], K.kp, K.kc, K.kr, t2, mpR, t3, muR);
                    const R *vw = var ? &T.vu[w][dx + 2] : &T.vp[w + 2][dx + 2];
                    R g[5];
                    weno_pair_vjp(vw[0], vw[1], vw[2], vw[3], vw[4], var ? muL : mpL, var ? muR : mpR, g);
                    R *o = T.xe[var][w];
                    if (e == 0) {
                        o[0] = g[4];
                    } else if (e == 1) {
                        o[1] = g[3];
                        o[2] = g[4];
                    } else if (e == 2) {
                        o[3] = g[0];
                        o[4] = g[1];
                    } else {
                        o[5] = g[0];
                    }
                }
            }
            __syncthreads();
            if (out_row) {
                // the tile edges' terms from the halo waves, then the same sum as inside
                if (lx == 0) { xp[1] = T.xe[0][wy][1]; xp[2] = T.xe[0][wy][0]; xu[1] = T.xe[1][wy][1]; xu[2] = T.xe[1][wy][0]; }
                if (lx == 1) { xp[2] = T.xe[0][wy][2]; xu[2] = T.xe[1][wy][2]; }
                if (lx == 62) { xp[4] = T.xe[0][wy][3]; xu[4] = T.xe[1][wy][3]; }
                if (lx == 63) { xp[3] = T.xe[0][wy][4]; xp[4] = T.xe[0][wy][5]; xu[3] = T.xe[1][wy][4]; xu[4] = T.xe[1][wy][5]; }
                const R cxp = xp[0] + xp[1] + xp[2] + xp[3] + xp[4];
                const R cxu = xu[0] + xu[1] + xu[2] + xu[3] + xu[4];
                R cyp = 0, cyv = 0;
#pragma unroll
                for (int q = 0; q < 5; ++q) {
                    cyp += T.gy[0][q][wy + 2 - q][lx];
                    cyv += T.gy[1][q][wy + 2 - q][lx];
                }
                xa[Z2][0] = cxp + cyp;
                xa[Z2][1] = cxu;
                xa[Z2][2] = cyv;
            }
        }
        // ---- z pair of cell k (output rows)
        if (out_row && k >= 2 && k <= Nz - 3) {
            R mpL, mpR, mwL, mwR, t0, t1, t2, t3;
            godunov_vjp(ap[Z2] - ap[Z1], aw[Z2] - aw[Z1], K.kp, K.kc, K.kr, mpL, t0, mwL, t1);
            godunov_vjp(ap[Z1] - ap[Z0], aw[Z1] - aw[Z0], K.kp, K.kc, K.kr, t2, mpR, t3, mwR);
            R gp[5], gw[5];
            weno_pair_vjp(zp[Z0], zp[Z1], zp[Z2], zp[Z3], zp[Z4], mpL, mpR, gp);
            weno_pair_vjp(zw[Z0], zw[Z1], zw[Z2], zw[Z3], zw[Z4], mwL, mwR, gw);
            za[Z0] += gp[0]; za[Z1] += gp[1]; za[Z2] += gp[2]; za[Z3] += gp[3]; za[Z4] += gp[4];
            zb[Z0] += gw[0]; zb[Z1] += gw[1]; zb[Z2] += gw[2]; zb[Z3] += gw[3]; zb[Z4] += gw[4];
        }
        // ---- plane k - 2 is complete
        if (done) {
            const R v0 = xa[Z0][0] + za[Z0] + (me ? K.beta * o0 : R(0));
            const R v1 = xa[Z0][1] + (me ? K.beta * o1 : R(0));
            const R v2 = xa[Z0][2] + (me ? K.beta * o2 : R(0));
            const R v3 = zb[Z0] + (me ? K.beta * o3 : R(0));
            rst(lin0, pm, ACC_IN ? i0 + v0 : v0);
            rst(lin1, pm, ACC_IN ? i1 + v1 : v1);
            rst(lin2, pm, ACC_IN ? i2 + v2 : v2);
            rst(lin3, pm, ACC_IN ? i3 + v3 : v3);
            if (HAS_BASE) {
                const R w0 = me ? K.alpha * o0 : R(0), w1 = me ? K.alpha * o1 : R(0);
                const R w2 = me ? K.alpha * o2 : R(0), w3 = me ? K.alpha * o3 : R(0);
                rst(lb0, pm, ACC_BASE ? c0 + w0 : w0);
                rst(lb1, pm, ACC_BASE ? c1 + w1 : w1);
                rst(lb2, pm, ACC_BASE ? c2 + w2 : w2);
                rst(lb3, pm, ACC_BASE ? c3 + w3 : w3);
            }
        }
        // ---- advance the rings: logical slot 0 leaves, the new entries take free slots
        za[Z0] = 0; zb[Z0] = 0;                       // becomes plane k+3's (logical 4)
        zp[Z0] = np_; zw[Z0] = nw_;                   // plane k+3
        ap[Z3] = nap; aw[Z3] = naw;                   // plane k+2 (logical 2 next phase)
        xa[Z3][0] = 0; xa[Z3][1] = 0; xa[Z3][2] = 0;  // plane k+1 (logical 2 next phase)
    };
    // five planes per trip (the ring period); the plane range is block-uniform
    for (int k = k0 - 2; k <= kend + 1; k += 5) {
        step(k, std::integral_constant<int, 0>{});
        if (k + 1 > kend + 1) break;
        step(k + 1, std::integral_constant<int, 1>{});
        if (k + 2 > kend + 1) break;
        step(k + 2, std::integral_constant<int, 2>{});
        if (k + 3 > kend + 1) break;
        step(k + 3, std::integral_constant<int, 3>{});
        if (k + 4 > kend + 1) break;
        step(k + 4, std::integral_constant<int, 4>{});
    }
}


// ADE replay of one step on the boundary list: x1 = stage 0 (x0, wout(V0)),
// x2 = stage 1 (x0, x1, wout(V1)), the same arithmetic as the forward boundary kernel.
struct AdeState {
    double *phi, *p1, *p2;    // [K][nB], [L][nB], [L][nB] (pole-major, see ade_at)
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
k_ade_replay(Grid g, FieldSet<S> V0, FieldSet<S> V1, Bnd bd, Ade ade, double rho, double c, double dt,
             AdeState x0, AdeState x1, AdeState x2, Scl<double> sc) {
#pragma clang fp contract(off)
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= bd.n) return;
    const BndCell bc = bnd_cell(g, bd, b);
    const int base = bc.q.base, mw = bc.mw, m = mw & 0xFF;
    const double sx = bc.q.sx, sy = bc.q.sy, sz = bc.q.sz;
    const double nx = bc.nx, ny = bc.ny, nz = bc.nz;
    const double Z0 = rho * c;
    const double w0 = probe_wout(V0, g, base, mw, sx, sy, sz, nx, ny, nz, Z0, sc);
    const double w1 = probe_wout(V1, g, base, mw, sx, sy, sz, nx, ny, nz, Z0, sc);
    for (int kk = 0; kk < ade.Keff[m]; ++kk) {
        const long long s = ade_at(kk, b, bd.n);
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
        const long long s = ade_at(l, b, bd.n);
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
// xr; lnew is the adjoint of the stage's new state (complete), lmid / lr accumulate
// (SET_MID / SET_R: first writer of that buffer in the step, which writes instead of
// adding, so the buffers need no clearing).
// Lout (adjoint of the stage output) is read and replaced by (1 - chi) Lout on the
// boundary cell; lD[4 b + f] receives the adjoint of the probe values (p, u, v, w);
// G[p nB + b] accumulates the material-parameter gradient of the cell, p over
// (Rinf, A[K], zeta[K], B[L], C[L], alpha[L], beta[L]) (entry-major, so the accumulation
// is coalesced). The stage coefficients come as a, b and 1/c.
template <typename S, typename R, bool SET_MID, bool SET_R>
__global__ void __launch_bounds__(256)
k_boundary_adj(Grid g, FieldSet<S> Vin, Bnd bd, Ade ade, double rho, double c, double dt, double as,
               double bs, double ic, AdeState xmid, AdeState xr, AdeState lnew, AdeState lmid, AdeState lr,
               R *L0, R *L1, R *L2, R *L3, double *__restrict__ lD, double *__restrict__ G, int P,
               Scl<double> sc) {
#pragma clang fp contract(off)
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= bd.n) return;
    const BndCell cel = bnd_cell(g, bd, b);
    const int C = cel.C, base = cel.q.base, mw = cel.mw, m = mw & 0xFF;
    const double nx = cel.nx, ny = cel.ny, nz = cel.nz, chi = cel.q.chi;
    const double Z0 = rho * c;
    const double wout = probe_wout(Vin, g, base, mw, cel.q.sx, cel.q.sy, cel.q.sz, nx, ny, nz, Z0, sc);
    const int K = ade.K, L = ade.L;
    const long long nB = bd.n;
    double *Gb = G + b;                                 // entry p at Gb[p nB]
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
    const double ac = as * ic, bc = bs * ic, f = bc * dt;    // a/c, b/c, b dt/c
    for (int kk = 0; kk < ade.Keff[m]; ++kk) {
        const long long s = ade_at(kk, b, bd.n);
        const int mi = m * K + kk;
        const double z = ade.zeta[mi];
        const double x_r = xr.phi[s], x_m = xmid.phi[s];
        const double xn = (as * x_m + bs * (x_r + dt * (wout - z * x_r))) * ic;
        const double ln = lnew.phi[s] + ade.A[mi] * lwin;
        Gb[(1 + kk) * nB] += lwin * xn;
        // stage 0 has mid == r: both terms land in one buffer, mid first
        lmid.phi[s] = (SET_MID ? 0.0 : lmid.phi[s]) + ac * ln;
        lr.phi[s] = (SET_R ? 0.0 : lr.phi[s]) + bc * (1.0 - dt * z) * ln;
        lwout += f * ln;
        Gb[(1 + K + kk) * nB] += -f * x_r * ln;
    }
    for (int l = 0; l < ade.Leff[m]; ++l) {
        const long long s = ade_at(l, b, bd.n);
        const int mi = m * L + l;
        const double al = ade.alpha[mi], be = ade.beta[mi];
        const double y1 = xr.p1[s], y2 = xr.p2[s];
        const double y1n = (as * xmid.p1[s] + bs * (y1 + dt * (wout - (al * y1 + be * y2)))) * ic;
        const double y2n = (as * xmid.p2[s] + bs * (y2 + dt * (be * y1 - al * y2))) * ic;
        const double l1 = lnew.p1[s] + ade.B[mi] * lwin, l2 = lnew.p2[s] + ade.C[mi] * lwin;
        Gb[(1 + 2 * K + l) * nB] += lwin * y1n;
        Gb[(1 + 2 * K + L + l) * nB] += lwin * y2n;
        lmid.p1[s] = (SET_MID ? 0.0 : lmid.p1[s]) + ac * l1;
        lmid.p2[s] = (SET_MID ? 0.0 : lmid.p2[s]) + ac * l2;
        lr.p1[s] = (SET_R ? 0.0 : lr.p1[s]) + bc * (l1 * (1.0 - dt * al) + l2 * (dt * be));
        lr.p2[s] = (SET_R ? 0.0 : lr.p2[s]) + bc * (l1 * (-dt * be) + l2 * (1.0 - dt * al));
        lwout += f * l1;
        Gb[(1 + 2 * K + 2 * L + l) * nB] += f * (-y1 * l1 - y2 * l2);
        Gb[(1 + 2 * K + 3 * L + l) * nB] += f * (-y2 * l1 + y1 * l2);
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
                               R *L0, R *L1, R *L2, R *L3, const int *__restrict__ pidx, double *lsh, int np) {
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
    const int w = pidx ? pidx[C] : -1;
    if (w >= 0) {                       // cell with an FP64 adjoint shadow
        double *o = lsh + w;
        o[0] += s0; o[np] += s1; o[2 * (long long)np] += s2; o[3 * (long long)np] += s3;
        L0[C] = (R)o[0]; L1[C] = (R)o[np]; L2[C] = (R)o[2 * (long long)np]; L3[C] = (R)o[3 * (long long)np];
        return;
    }
    L0[C] = (R)((double)L0[C] + s0);
    L1[C] = (R)((double)L1[C] + s1);
    L2[C] = (R)((double)L2[C] + s2);
    L3[C] = (R)((double)L3[C] + s3);
}

// Source gradients of one stage: d/d sval_s and d/d u_s[nn] of p[cell_s] += coef sval_s amp.
template <typename R>
__global__ void k_source_adj(const R *__restrict__ Lp, const int *__restrict__ cell, const double *__restrict__ sval,
                             const double *__restrict__ u_s, long long us_stride, int n_src, long long nn, double coef,
                             double *__restrict__ gs, double *__restrict__ gus, const int *__restrict__ pidx,
                             const double *__restrict__ lsh_p) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_src) return;
    const int C = cell[s];
    const int w = pidx ? pidx[C] : -1;
    const double lp = w >= 0 ? lsh_p[w] : (double)Lp[C];
    const long long idx = (long long)s * us_stride + nn;
    gs[s] += lp * coef * u_s[idx];
    if (gus) atomicAdd(gus + idx, lp * coef * sval[s]);
}

// IR misfit injection: Lp[rec_r] += dJ/dir[r][n] (sequential, receivers may share a cell).
template <typename R>
__global__ void k_receiver_adj(R *Lp, const int *__restrict__ cell, int n_rec, const double *__restrict__ dj,
                               long long stride, long long n, const int *__restrict__ pidx, double *lsh_p) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int r = 0; r < n_rec; ++r) {
        const int C = cell[r];
        const int w = pidx ? pidx[C] : -1;
        if (w >= 0) {
            lsh_p[w] += dj[r * stride + n];
            Lp[C] = (R)lsh_p[w];
        } else {
            Lp[C] = (R)((double)Lp[C] + dj[r * stride + n]);
        }
    }
}


// cPML adjoint (forward: functions3D.cuh:681-756, pml_apply). An updated cell in the
// band of axis a adds beta dt (kc_a D_a + psi_a) to p and to the a-velocity, where D_a is
// that field's axis-a flux difference over h and psi_a the step's auxiliary state, and
// the last stage advances psi_a <- b psi_a + a_a D_a. The adjoint of D_a therefore carries,
// beyond the main flux adjoint's seed beta dt / h lout, the extra seed
//   e = beta dt / h kc_a lout  (+ a_a / h lpsi on the last stage),
// lpsi being the adjoint of the advanced psi. k_pml_flux_adj adds the axis-a flux VJP of
// those extra seeds to the stage-input adjoint (the VJP is linear in its seeds, so this
// sum equals the VJP of the full seed); k_pml_psi_adj then turns lpsi into the adjoint of
// the step's psi: lpsi <- b lpsi + beta dt lout on the last stage, lpsi += beta dt lout on
// the others (all three stages read the same psi). The psi values themselves never enter.
struct PmlAdjAxis {
    const int *band;           // per position along the axis: -1, or the psi tile slot [0, 2d)
    const double *kc, *b, *a;
    double *lpsi_p, *lpsi_u;   // adjoints of psi_p and of the axis velocity's psi, tile layout
    int d;
};

// psi tile index of the cell (i, j, k) at band slot bnd of axis AX (fvm_device.hpp layout)
template <int AX>
__device__ inline long long pml_tile_index(int bnd, int i, int j, int k, int Nx, int Ny, int d) {
    if (AX == 0) return bnd + (long long)(2 * d) * (j + (long long)Ny * k);
    if (AX == 1) return i + (long long)Nx * (bnd + (long long)(2 * d) * k);
    return i + (long long)Nx * (j + (long long)Ny * bnd);
}

// Extra-seed flux VJP of axis AX on the cells whose windows reach the band: positions
// along AX from plist (npos of them), the two other axes over [3, N-3). One thread per
// cell; adds to the stage-input adjoint (and to its FP64 shadow on adjoint-patch cells).
template <typename S, typename R, int AX, bool LAST>
__global__ void __launch_bounds__(256)
k_pml_flux_adj(Grid g, FieldSet<S> V, Scl<R> sc, const R *__restrict__ L0, const R *__restrict__ L1,
               const R *__restrict__ L2, const R *__restrict__ L3, const uint8_t *__restrict__ upd, R *lin0, R *lin1,
               R *lin2, R *lin3, const int *__restrict__ pidx, const double *__restrict__ lsh_out, double *lsh_in,
               int np, PmlAdjAxis pa, const int *__restrict__ plist, int npos, AdjK<R> K, R ih) {
    const int ni = AX == 0 ? npos : g.Nx - 6, nj = AX == 1 ? npos : g.Ny - 6, nk = AX == 2 ? npos : g.Nz - 6;
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)ni * nj * nk) return;
    const int a0 = (int)(t % ni);
    t /= ni;
    const int a1 = (int)(t % nj), a2 = (int)(t / nj);
    const int i = AX == 0 ? plist[a0] : a0 + 3, j = AX == 1 ? plist[a1] : a1 + 3, k = AX == 2 ? plist[a2] : a2 + 3;
    const long long C = (long long)i + (long long)g.Nx * j + g.nxy * k;
    const long long s = AX == 0 ? 1LL : (AX == 1 ? (long long)g.Nx : g.nxy);
    const int pos = AX == 0 ? i : (AX == 1 ? j : k);
    const int nax = AX == 0 ? g.Nx : (AX == 1 ? g.Ny : g.Nz);
    constexpr int fu = 1 + AX;
    const R *Lf[4] = {L0, L1, L2, L3};
    const S *Vf[4] = {V.p, V.u, V.v, V.w};
    auto lraw = [&](int f, long long X) -> R {
        const int w = pidx ? pidx[X] : -1;
        return w >= 0 ? (R)lsh_out[(long long)f * np + w] : Lf[f][X];
    };
    // extra seed of field f (0 = p, 1 = the AX velocity) at the cell X = C + (pX - pos) s
    auto es = [&](int f, int pX) -> R {
        const int bnd = pa.band[pX];
        const long long X = C + (long long)(pX - pos) * s;
        if (bnd < 0 || !upd[X]) return R(0);
        R v = K.bdh * (R)pa.kc[pX] * lraw(f ? fu : 0, X);
        if (LAST) {
            const int ii = AX == 0 ? pX : i, jj = AX == 1 ? pX : j, kk = AX == 2 ? pX : k;
            const long long T = pml_tile_index<AX>(bnd, ii, jj, kk, g.Nx, g.Ny, pa.d);
            v += (R)pa.a[pX] * ih * (R)(f ? pa.lpsi_u[T] : pa.lpsi_p[T]);
        }
        return v;
    };
    auto vv = [&](int f, long long X) -> R { return Conv<S, R>::ld(Vf[f][X], sc.s[f]); };
    R ap = R(0), au = R(0);
#pragma unroll
    for (int q = -2; q <= 2; ++q) {
        const int cpos = pos - q;
        if (cpos < 2 || cpos > nax - 3) continue;
        const R ep0 = es(0, cpos), eu0 = es(1, cpos);
        const R ep1 = cpos + 1 <= nax - 1 ? es(0, cpos + 1) : R(0), eu1 = cpos + 1 <= nax - 1 ? es(1, cpos + 1) : R(0);
        const R epm = es(0, cpos - 1), eum = es(1, cpos - 1);
        if (ep0 == R(0) && eu0 == R(0) && ep1 == R(0) && eu1 == R(0) && epm == R(0) && eum == R(0)) continue;
        const long long Cq = C - q * s;
        R mpL, mpR, muL, muR, tpL, tpR, tuL, tuR;
        godunov_vjp(ep1 - ep0, eu1 - eu0, K.kp, K.kc, K.kr, mpL, tpR, muL, tuR);
        godunov_vjp(ep0 - epm, eu0 - eum, K.kp, K.kc, K.kr, tpL, mpR, tuL, muR);
        R gp[5], gu[5];
        weno_pair_vjp(vv(0, Cq - 2 * s), vv(0, Cq - s), vv(0, Cq), vv(0, Cq + s), vv(0, Cq + 2 * s), mpL, mpR, gp);
        weno_pair_vjp(vv(fu, Cq - 2 * s), vv(fu, Cq - s), vv(fu, Cq), vv(fu, Cq + s), vv(fu, Cq + 2 * s), muL, muR,
                      gu);
        ap += gp[q + 2];
        au += gu[q + 2];
    }
    if (ap == R(0) && au == R(0)) return;
    R *li[4] = {lin0, lin1, lin2, lin3};
    const int w = pidx ? pidx[C] : -1;
    if (w >= 0) {
        double *o0 = lsh_in + w, *o1 = lsh_in + (long long)fu * np + w;
        *o0 += (double)ap;
        *o1 += (double)au;
        li[0][C] = (R)*o0;
        li[fu][C] = (R)*o1;
        return;
    }
    li[0][C] += ap;
    li[fu][C] += au;
}

// Adjoint of the step's psi of axis AX, one thread per psi tile entry (see above).
template <typename R, int AX, bool LAST>
__global__ void __launch_bounds__(256)
k_pml_psi_adj(Grid g, const R *__restrict__ Lp, const R *__restrict__ Lu, const uint8_t *__restrict__ upd,
              const int *__restrict__ pidx, const double *__restrict__ lsh_out, int np, PmlAdjAxis pa, double bdt) {
    const int d2 = 2 * pa.d;
    const int n0 = AX == 0 ? d2 : g.Nx, n1 = AX == 1 ? d2 : g.Ny, n2 = AX == 2 ? d2 : g.Nz;
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)n0 * n1 * n2) return;
    const long long T = t;
    const int a0 = (int)(t % n0);
    t /= n0;
    const int a1 = (int)(t % n1), a2 = (int)(t / n1);
    const int bnd = AX == 0 ? a0 : (AX == 1 ? a1 : a2);
    const int nax = AX == 0 ? g.Nx : (AX == 1 ? g.Ny : g.Nz);
    // the position of this slot: low face NG + bnd, high face nax - NG - d + (bnd - d)
    const int pos = bnd < pa.d ? NG + bnd : nax - NG - pa.d + (bnd - pa.d);
    if (pos < 0 || pos >= nax || pa.band[pos] != bnd) return;
    const int i = AX == 0 ? pos : a0, j = AX == 1 ? pos : a1, k = AX == 2 ? pos : a2;
    if (i < NG || i >= g.Nx - NG || j < NG || j >= g.Ny - NG || k < NG || k >= g.Nz - NG) return;
    const long long X = (long long)i + (long long)g.Nx * j + g.nxy * k;
    double lp = 0.0, lu = 0.0;
    if (upd[X]) {
        const int w = pidx ? pidx[X] : -1;
        lp = w >= 0 ? lsh_out[w] : (double)Lp[X];
        lu = w >= 0 ? lsh_out[(long long)(1 + AX) * np + w] : (double)Lu[X];
    }
    const double b = LAST ? (double)(R)pa.b[pos] : 1.0;
    pa.lpsi_p[T] = b * pa.lpsi_p[T] + bdt * lp;
    pa.lpsi_u[T] = b * pa.lpsi_u[T] + bdt * lu;
}

}  // namespace fasesdev
