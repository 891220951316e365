// Tiled flux-stage kernel for gfx950: WENO5-Z + Godunov + SSP-RK3 stage + cPML with
// every face reconstructed once.
//
// Block = 64 lanes (x columns x0-1 .. x0+62) x NW waves (y rows j0-1 .. j0+NW-2),
// marching over TZ planes in z. Outputs are the inner 62 columns and NW-2 rows; the
// outer lane / wave of each side only supplies face states to its neighbour.
//   z faces: register pipeline along the march (face k+1/2 computed once, carried);
//   x faces: each lane reconstructs both states of its own cell, the right state of
//            the next cell comes from the neighbouring lane (shuffle);
//   y faces: both states per cell written to LDS, read by the rows above and below.
// p, u (x) and p, v (y) of the current plane sit in LDS with a 3-cell halo; p and w
// along z stay in registers.
//
// WENO5-Z arithmetic (functions3D.cuh:763-977, eps = 1e-10, q = 1) is restated on
// first differences of the 5-cell window: candidates c + linear combinations of the
// differences, smoothness indicators 13/12 s^2 + 1/4 t^2 from second differences s and
// the one-sided combinations t, weights d_r (1 + tau/(beta_r + eps)), normalised once.
#pragma once

#include "fvm_device.hpp"

namespace fasesdev {

template <typename R> __device__ inline R fast_rcp(R x);
template <> __device__ inline float fast_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <> __device__ inline double fast_rcp<double>(double x) {
    double r = __builtin_amdgcn_rcp(x);          // ~2^-26 seed, then two Newton steps
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}

// Both WENO5-Z states of the cell whose 5-cell window is (a,b,c,d,e), centre c:
//   A = V^-_{i+1/2} (linear weights 3/10, 3/5, 1/10 on the stencils ending at c+2, c+1, c)
//   B = V^+_{i-1/2} (mirrored weights 1/10, 3/5, 3/10)
template <typename R>
__device__ inline void weno_pair(R a, R b, R c, R d, R e, R &A, R &B) {
    const R d1 = b - a, d2 = c - b, d3 = d - c, d4 = e - d;
    const R s0 = d4 - d3, s1 = d3 - d2, s2 = d2 - d1;              // second differences
    const R t0 = fma(R(-3), d3, d4), t1 = d2 + d3, t2 = fma(R(3), d2, -d1);
    // 4 x smoothness indicators: 13/3 s^2 + t^2 (the weights only see ratios of
    // beta + eps and |beta_0 - beta_2|, so beta, tau and eps all carry the factor 4)
    const R k = R(13) / R(3);
    const R e0 = fma(k * s0, s0, fma(t0, t0, R(4e-10)));
    const R e1 = fma(k * s1, s1, fma(t1, t1, R(4e-10)));
    const R e2 = fma(k * s2, s2, fma(t2, t2, R(4e-10)));
    const R tau = fabs(e0 - e2);
    // unnormalised WENO-Z weights d_r (1 + tau/e_r) times e0 e1 e2:
    // g_r = (e_r + tau) * prod_{s != r} e_s
    const R g0 = (e0 + tau) * (e1 * e2), g1 = (e1 + tau) * (e0 * e2), g2 = (e2 + tau) * (e0 * e1);
    // candidate offsets from c, pre-multiplied by the linear weights
    //   left  (V^- at i+1/2): d = 3/10, 3/5, 1/10 on {1/3,5/6,-1/6}(c,d,e), {-1/6,5/6,1/3}(b,c,d),
    //                          {1/3,-7/6,11/6}(a,b,c)
    //   right (V^+ at i-1/2): d = 1/10, 3/5, 3/10 on {11/6,-7/6,1/3}(c,d,e), {1/3,5/6,-1/6}(b,c,d),
    //                          {-1/6,5/6,1/3}(a,b,c)
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
    A = fma(nl, fast_rcp(dl), c);
    B = fma(nr, fast_rcp(dr), c);
}

namespace tile {
constexpr int TXO = 62;         // output columns per block
constexpr int LX = 68;          // LDS columns: x0-3 .. x0+64
}  // namespace tile

template <typename R, int NW> struct TileLds {
    R p[NW + 4][tile::LX];      // rows j0-3 .. j0+NW
    R u[NW][tile::LX];
    R v[NW + 4][64];
    R ys[4][NW][64];            // y states: p^-, p^+, v^-, v^+ of every row
};

// Constants of the tiled kernel as unevaluated sums hi + lo (lo = 0 for FP64): in a
// lossless resonant room a constant rounded once (dt/h, 1/(2 rho), 1/3) is a systematic
// bias of the effective sound speed or amplitude that builds up over thousands of steps,
// so each carries its rounding residual.
template <typename R> struct Split {
    R hi, lo;
};
template <typename R>
__device__ inline R mul_s(const Split<R> &a, R x) { return fma(a.hi, x, a.lo * x); }
template <typename R>
__device__ inline R fma_s(const Split<R> &a, R x, R y) { return fma(a.hi, x, fma(a.lo, x, y)); }

// Godunov flux with kp = rho c^2 / 2, kc = c / 2, kr = 1/(2 rho):
// Fp = kp (uR + uL) - kc (pR - pL), Fu = kr (pR + pL) - kc (uR - uL).
template <typename R> struct TileK {
    Split<R> kp, kr, dth, third;   // rho c^2/2, 1/(2 rho), dt/h, 1/3
    R kc, ih;                       // c/2 (exact), 1/h (PML only)
};
template <typename R>
__device__ inline void godunov_f(R pL, R pR, R uL, R uR, const TileK<R> &K, R &Fp, R &Fu) {
    Fp = fma_s(K.kp, uR + uL, -K.kc * (pR - pL));
    Fu = fma_s(K.kr, pR + pL, -K.kc * (uR - uL));
}

template <typename R>
__device__ inline R shfl_next(R x) { return __shfl_down(x, 1, 64); }   // value of lane + 1
template <typename R>
__device__ inline R shfl_prev(R x) { return __shfl_up(x, 1, 64); }     // value of lane - 1

// Element at a 32-bit byte offset from a base pointer (the offset of a lane within a
// plane is computed once; the plane base changes per plane).
template <typename T> __device__ inline T &at(T *base, unsigned boff) {
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + boff);
}
template <typename T> __device__ inline const T &at(const T *base, unsigned boff) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + boff);
}

// Field pointers of one tiled stage launch: input set, base set U0, output set.
template <typename S> struct TileFields {
    const S *ip, *iu, *iv, *iw;
    const S *bp, *bu, *bv, *bw;
    S *op, *ou, *ov, *ow;
};

// One tile = 64 x NW threads; outputs 62 columns x (NW-2) rows x tz planes of tile
// (bx, by, bz) of an Nx x Ny x Nz grid.
template <typename S, typename R, int STAGE, int NW>
__device__ inline void flux_tile(TileLds<R, NW> &L, int Nx, int Ny, int Nz, const TileFields<S> &F,
                                 const uint8_t *__restrict__ upd, const PmlSet &pml, const TileK<R> &K, R dt,
                                 const Scl<R> &sc, int tz, int bx, int by, int bz) {
    using namespace tile;
    constexpr int LY = NW + 4, TYO = NW - 2;
    typedef Conv<S, R> CV;
    const S *__restrict__ ip = F.ip, *__restrict__ iu = F.iu, *__restrict__ iv = F.iv, *__restrict__ iw = F.iw;
    const S *__restrict__ bp = F.bp, *__restrict__ bu = F.bu, *__restrict__ bv = F.bv, *__restrict__ bw = F.bw;
    S *__restrict__ op = F.op, *__restrict__ ou = F.ou, *__restrict__ ov = F.ov, *__restrict__ ow = F.ow;
    // blockDim.x is one wavefront, so the row index is uniform per wave: keep it scalar
    const int lx = threadIdx.x, wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int tid = wy * 64 + lx;
    const int x0 = 3 + TXO * bx, j0 = 3 + TYO * by, k0 = 3 + tz * bz;
    const long long nxy = (long long)Nx * Ny;
    const int kend = min(k0 + tz, Nz - 3);
    const int i = x0 - 1 + lx, j = j0 - 1 + wy;
    const int ic = min(i, Nx - 1), jc = min(j, Ny - 1);               // clamped for loads
    const unsigned col = (unsigned)(ic + Nx * jc);                   // offset within a plane
    const unsigned cb = col * (unsigned)sizeof(S);                   // ... in bytes
    const bool inner_row = (wy >= 1 && wy <= NW - 2);
    const bool out_lane = (lx >= 1 && lx <= TXO) && i < Nx - 3 && j < Ny - 3 && inner_row;

    // halo items of this thread (p columns/rows, u columns, v rows): items tid, tid + 64 NW, ...
    constexpr int NITEM = 4 * LY + 4 * 64 + 4 * NW + 4 * 64;
    constexpr int NH = (NITEM + 64 * NW - 1) / (64 * NW);
    bool h_on[NH];
    const S *hsrc[NH];
    R hsc[NH];
    R *hdst[NH];
    unsigned h_off[NH];
#pragma unroll
    for (int q = 0; q < NH; ++q) {
        const int nP1 = 4 * LY, nP2 = 4 * 64, nU = 4 * NW, nV = 4 * 64;
        const int sideX[4] = {0, 1, LX - 2, LX - 1};
        const int sideY[4] = {0, 1, LY - 2, LY - 1};
        int t = tid + q * 64 * NW;
        int h_kind = -1, h_r = 0, h_c = 0;
        if (t < nP1) {
            h_kind = 0; h_c = sideX[t / LY]; h_r = t % LY;
        } else if ((t -= nP1) < nP2) {
            h_kind = 0; h_r = sideY[t / 64]; h_c = 2 + t % 64;
        } else if ((t -= nP2) < nU) {
            h_kind = 1; h_c = sideX[t / NW]; h_r = t % NW;
        } else if ((t -= nU) < nV) {
            h_kind = 2; h_r = sideY[t / 64]; h_c = t % 64;
        }
        h_on[q] = h_kind >= 0;
        h_off[q] = 0;
        if (h_kind >= 0) {
            int gx = (h_kind == 2) ? (x0 - 1 + h_c) : (x0 - 3 + h_c);
            int gy = (h_kind == 1) ? (j0 - 1 + h_r) : (j0 - 3 + h_r);
            gx = max(0, min(gx, Nx - 1));
            gy = max(0, min(gy, Ny - 1));
            h_off[q] = (unsigned)(gx + Nx * gy);
        }
        hsrc[q] = (h_kind == 0) ? ip : (h_kind == 1 ? iu : iv);
        hsc[q] = (h_kind == 0) ? sc.s[0] : (h_kind == 1 ? sc.s[1] : sc.s[2]);
        hdst[q] = (h_kind == 0) ? &L.p[h_r][h_c] : (h_kind == 1 ? &L.u[h_r][h_c] : &L.v[h_r][h_c]);
    }

    // z pipeline: window p/w at k-1..k+3 (index 0..4), carried A-state of cell k and
    // flux through k-1/2
    R pz[5] = {0, 0, 0, 0, 0}, wz[5] = {0, 0, 0, 0, 0};
    R pA = 0, wA = 0, FzP = 0, FzW = 0;
    if (inner_row) {
        R p7[7], w7[7];
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            const long long pq = nxy * (long long)(k0 - 3 + q);
            p7[q] = CV::ld(at(ip + pq, cb), sc.s[0]);
            w7[q] = CV::ld(at(iw + pq, cb), sc.s[3]);
        }
        R pAm, wAm, pB0, wB0, tmp0, tmp1;
        weno_pair(p7[0], p7[1], p7[2], p7[3], p7[4], pAm, tmp0);   // cell k0-1
        weno_pair(w7[0], w7[1], w7[2], w7[3], w7[4], wAm, tmp1);
        weno_pair(p7[1], p7[2], p7[3], p7[4], p7[5], pA, pB0);     // cell k0
        weno_pair(w7[1], w7[2], w7[3], w7[4], w7[5], wA, wB0);
        godunov_f(pAm, pB0, wAm, wB0, K, FzP, FzW);            // face k0-1/2
#pragma unroll
        for (int q = 0; q < 5; ++q) { pz[q] = p7[q + 2]; wz[q] = w7[q + 2]; }
    }
    // plane values in flight: own u, v (and p for the two outer rows), halo item
    R nu, nv, np = 0, nh[NH];
    {
        const long long pl = nxy * (long long)k0;
        nu = CV::ld(at(iu + pl, cb), sc.s[1]);
        nv = CV::ld(at(iv + pl, cb), sc.s[2]);
        if (!inner_row) np = CV::ld(at(ip + pl, cb), sc.s[0]);
#pragma unroll
        for (int q = 0; q < NH; ++q)                       // per-lane field: no uniform base
            nh[q] = h_on[q] ? CV::ld((hsrc[q] + pl)[h_off[q]], hsc[q]) : R(0);
    }

    for (int k = k0; k < kend; ++k) {
        const long long pl = nxy * (long long)k;
        // ---- fill the plane tiles from the values loaded during the previous plane
        const R pc = inner_row ? pz[1] : np;
        const R uc = nu, vc = nv;
        L.p[wy + 2][lx + 2] = pc;
        L.u[wy][lx + 2] = uc;
        L.v[wy + 2][lx] = vc;
#pragma unroll
        for (int q = 0; q < NH; ++q)
            if (h_on[q]) *hdst[q] = nh[q];
        __syncthreads();
        // ---- prefetch the next plane
        if (k + 1 < kend) {
            const long long pn = pl + nxy;
            nu = CV::ld(at(iu + pn, cb), sc.s[1]);
            nv = CV::ld(at(iv + pn, cb), sc.s[2]);
            if (!inner_row) np = CV::ld(at(ip + pn, cb), sc.s[0]);
#pragma unroll
            for (int q = 0; q < NH; ++q)
                if (h_on[q]) nh[q] = CV::ld((hsrc[q] + pn)[h_off[q]], hsc[q]);
        }
        // ---- y states of every row (incl. the two outer rows)
        {
            R A, B, Av, Bv;
            weno_pair(L.p[wy][lx + 2], L.p[wy + 1][lx + 2], pc, L.p[wy + 3][lx + 2], L.p[wy + 4][lx + 2], A, B);
            weno_pair(L.v[wy][lx], L.v[wy + 1][lx], vc, L.v[wy + 3][lx], L.v[wy + 4][lx], Av, Bv);
            L.ys[0][wy][lx] = A;
            L.ys[1][wy][lx] = B;
            L.ys[2][wy][lx] = Av;
            L.ys[3][wy][lx] = Bv;
        }
        // ---- x faces (shuffles are wave-wide: every lane of an inner row takes part)
        R FxP_p = 0, FxP_u = 0, FxM_p = 0, FxM_u = 0;
        if (inner_row) {
            R A, B, Au, Bu;
            weno_pair(L.p[wy + 2][lx], L.p[wy + 2][lx + 1], pc, L.p[wy + 2][lx + 3], L.p[wy + 2][lx + 4], A, B);
            weno_pair(L.u[wy][lx], L.u[wy][lx + 1], uc, L.u[wy][lx + 3], L.u[wy][lx + 4], Au, Bu);
            const R Bn = shfl_next(B), Bun = shfl_next(Bu);
            godunov_f(A, Bn, Au, Bun, K, FxP_p, FxP_u);          // face i+1/2
            FxM_p = shfl_prev(FxP_p);
            FxM_u = shfl_prev(FxP_u);
        }
        __syncthreads();

        R pAn = 0, wAn = 0, FzP_p = 0, FzP_w = 0, pnew = 0, wnew = 0;
        if (inner_row) {
            // ---- y faces j+1/2 and j-1/2
            R FyP_p, FyP_v, FyM_p, FyM_v;
            godunov_f(L.ys[0][wy][lx], L.ys[1][wy + 1][lx], L.ys[2][wy][lx], L.ys[3][wy + 1][lx], K, FyP_p, FyP_v);
            godunov_f(L.ys[0][wy - 1][lx], L.ys[1][wy][lx], L.ys[2][wy - 1][lx], L.ys[3][wy][lx], K, FyM_p, FyM_v);
            // ---- z face k+1/2 from the new cell k+1
            R pBn, wBn;
            weno_pair(pz[0], pz[1], pz[2], pz[3], pz[4], pAn, pBn);
            weno_pair(wz[0], wz[1], wz[2], wz[3], wz[4], wAn, wBn);
            godunov_f(pA, pBn, wA, wBn, K, FzP_p, FzP_w);
            const R wc = wz[1];
            // z window for the next plane: issue the lead loads before the update
            if (k + 1 < kend) {
                const long long p4 = nxy * (long long)(k + 4);
                pnew = CV::ld(at(ip + p4, cb), sc.s[0]);
                wnew = CV::ld(at(iw + p4, cb), sc.s[3]);
            }
            if (out_lane && at(upd + pl, col) == 1) {
                const R sp = -FxP_p + FxM_p - FyP_p + FyM_p - FzP_p + FzP;
                const R Du = -FxP_u + FxM_u, Dv = -FyP_v + FyM_v, Dw = -FzP_w + FzW;
                R uo, vo, wo, po, wgt;
                if (STAGE == 0) {
                    uo = fma_s(K.dth, Du, uc);
                    vo = fma_s(K.dth, Dv, vc);
                    wo = fma_s(K.dth, Dw, wc);
                    po = fma_s(K.dth, sp, pc);
                    wgt = R(1);
                } else {
                    const R p0 = CV::ld(at(bp + pl, cb), sc.s[0]), u0 = CV::ld(at(bu + pl, cb), sc.s[1]);
                    const R v0 = CV::ld(at(bv + pl, cb), sc.s[2]), w0 = CV::ld(at(bw + pl, cb), sc.s[3]);
                    if (STAGE == 1) {
                        uo = fma(R(0.75), u0, R(0.25) * fma_s(K.dth, Du, uc));
                        vo = fma(R(0.75), v0, R(0.25) * fma_s(K.dth, Dv, vc));
                        wo = fma(R(0.75), w0, R(0.25) * fma_s(K.dth, Dw, wc));
                        po = fma(R(0.75), p0, R(0.25) * fma_s(K.dth, sp, pc));
                        wgt = R(0.25);
                    } else {
                        uo = mul_s(K.third, fma(R(2), fma_s(K.dth, Du, uc), u0));
                        vo = mul_s(K.third, fma(R(2), fma_s(K.dth, Dv, vc), v0));
                        wo = mul_s(K.third, fma(R(2), fma_s(K.dth, Dw, wc), w0));
                        po = mul_s(K.third, fma(R(2), fma_s(K.dth, sp, pc), p0));
                        wgt = R(2) / R(3);
                    }
                }
                if (pml.on)
                    pml_apply<R, STAGE == 2>(pml, i, j, k, Nx, Ny, po, uo, vo, wo, (-FxP_p + FxM_p) * K.ih,
                                             (-FyP_p + FyM_p) * K.ih, (-FzP_p + FzP) * K.ih, Du * K.ih, Dv * K.ih,
                                             Dw * K.ih, wgt * dt);
                at(ou + pl, cb) = CV::st(uo, sc.is[1]);
                at(ov + pl, cb) = CV::st(vo, sc.is[2]);
                at(ow + pl, cb) = CV::st(wo, sc.is[3]);
                at(op + pl, cb) = CV::st(po, sc.is[0]);
            }
        }
        // advance the z pipeline (every row: the outer rows carry zeros, so the window
        // registers are the same on both sides of the row branch and nothing is copied)
        pA = pAn; wA = wAn; FzP = FzP_p; FzW = FzP_w;
        pz[0] = pz[1]; pz[1] = pz[2]; pz[2] = pz[3]; pz[3] = pz[4]; pz[4] = pnew;
        wz[0] = wz[1]; wz[1] = wz[2]; wz[2] = wz[3]; wz[3] = wz[4]; wz[4] = wnew;
    }
}

template <typename S, typename R, int STAGE, int NW>
__global__ void __launch_bounds__(64 * NW)
k_flux_tiled(int Nx, int Ny, int Nz, TileFields<S> F, const uint8_t *__restrict__ upd, PmlSet pml, TileK<R> K, R dt,
             Scl<R> sc, int tz) {
    __shared__ TileLds<R, NW> L;
    flux_tile<S, R, STAGE, NW>(L, Nx, Ny, Nz, F, upd, pml, K, dt, sc, tz, blockIdx.x, blockIdx.y, blockIdx.z);
}

}  // namespace fasesdev
