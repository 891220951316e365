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
    const R k13 = R(13) / R(12);
    const R b0 = fma(k13 * s0, s0, R(0.25) * (t0 * t0));
    const R b1 = fma(k13 * s1, s1, R(0.25) * (t1 * t1));
    const R b2 = fma(k13 * s2, s2, R(0.25) * (t2 * t2));
    const R eps = R(1e-10);
    const R tau = fabs(b0 - b2);
    const R q0 = tau * fast_rcp(b0 + eps), q1 = tau * fast_rcp(b1 + eps), q2 = tau * fast_rcp(b2 + eps);
    // candidate offsets from c
    const R c6 = R(1) / R(6), c3 = R(1) / R(3);
    const R a0 = fma(R(2) * c3, d3, -c6 * d4);        // {1/3,5/6,-1/6}(c,d,e) - c
    const R a1 = fma(c6, d2, c3 * d3);                // {-1/6,5/6,1/3}(b,c,d) - c
    const R a2 = fma(R(5) * c6, d2, -c3 * d1);        // {1/3,-7/6,11/6}(a,b,c) - c
    const R g0 = fma(R(-5) * c6, d3, c3 * d4);        // {11/6,-7/6,1/3}(c,d,e) - c
    const R g1 = fma(-c3, d2, -c6 * d3);              // {1/3,5/6,-1/6}(b,c,d) - c
    const R g2 = fma(R(-2) * c3, d2, c6 * d1);        // {-1/6,5/6,1/3}(a,b,c) - c
    const R al0 = fma(R(0.3), q0, R(0.3)), al1 = fma(R(0.6), q1, R(0.6)), al2 = fma(R(0.1), q2, R(0.1));
    const R ar0 = fma(R(0.1), q0, R(0.1)), ar1 = al1, ar2 = fma(R(0.3), q2, R(0.3));
    const R nl = fma(al0, a0, fma(al1, a1, al2 * a2));
    const R nr = fma(ar0, g0, fma(ar1, g1, ar2 * g2));
    A = fma(nl, fast_rcp(al0 + al1 + al2), c);
    B = fma(nr, fast_rcp(ar0 + ar1 + ar2), c);
}

namespace tile {
constexpr int NW = 16;          // waves (rows) per block
constexpr int TYO = NW - 2;     // output rows per block
constexpr int TXO = 62;         // output columns per block
constexpr int LX = 68;          // LDS columns: x0-3 .. x0+64
constexpr int LY = NW + 4;      // LDS rows:    j0-3 .. j0+NW
constexpr int TZ = 64;          // planes per block
}  // namespace tile

template <typename R> struct TileLds {
    R p[tile::LY][tile::LX];
    R u[tile::NW][tile::LX];
    R v[tile::LY][64];
    R ys[4][tile::NW][64];      // y states: p^-, p^+, v^-, v^+ of every row
};

// Godunov flux with contraction allowed (tiled kernel).
template <typename R>
__device__ inline void godunov_f(R pL, R pR, R uL, R uR, R rho, R c, R &Fp, R &Fu) {
    Fp = R(0.5) * (rho * c * c * (uR + uL) - c * (pR - pL));
    Fu = R(0.5) * ((pR + pL) / rho - c * (uR - uL));
}

template <typename R>
__device__ inline R shfl_next(R x) { return __shfl_down(x, 1, 64); }   // value of lane + 1
template <typename R>
__device__ inline R shfl_prev(R x) { return __shfl_up(x, 1, 64); }     // value of lane - 1

template <typename S, typename R, int STAGE>
__global__ void __launch_bounds__(64 * tile::NW)
k_flux_tiled(int Nx, int Ny, int Nz, const S *__restrict__ ip, const S *__restrict__ iu, const S *__restrict__ iv,
             const S *__restrict__ iw, const S *__restrict__ bp, const S *__restrict__ bu, const S *__restrict__ bv,
             const S *__restrict__ bw, S *__restrict__ op, S *__restrict__ ou, S *__restrict__ ov,
             S *__restrict__ ow, const uint8_t *__restrict__ upd, PmlSet pml, R rho, R c, R h, R dt, R s_in,
             R s_base, R is_out) {
    using namespace tile;
    __shared__ TileLds<R> L;
    typedef Conv<S, R> CV;
    const int lx = threadIdx.x, wy = threadIdx.y;
    const int tid = wy * 64 + lx;
    const int x0 = 3 + TXO * blockIdx.x, j0 = 3 + TYO * blockIdx.y, k0 = 3 + TZ * blockIdx.z;
    const long long nxy = (long long)Nx * Ny;
    const int kend = min(k0 + TZ, Nz - 3);
    const int i = x0 - 1 + lx, j = j0 - 1 + wy;
    const int ic = min(i, Nx - 1), jc = min(j, Ny - 1);               // clamped for loads
    const long long col = ic + (long long)Nx * jc;
    const bool inner_row = (wy >= 1 && wy <= NW - 2);
    const bool out_lane = (lx >= 1 && lx <= TXO) && i < Nx - 3 && j < Ny - 3 && inner_row;

    // halo item of this thread: p columns/rows, u columns, v rows (see LDS layout)
    int h_kind = -1, h_r = 0, h_c = 0;
    long long h_off = 0;
    {
        const int nP1 = 4 * LY, nP2 = 4 * 64, nU = 4 * NW, nV = 4 * 64;
        const int sideX[4] = {0, 1, LX - 2, LX - 1};
        const int sideY[4] = {0, 1, LY - 2, LY - 1};
        int t = tid;
        if (t < nP1) {
            h_kind = 0; h_c = sideX[t / LY]; h_r = t % LY;
        } else if ((t -= nP1) < nP2) {
            h_kind = 0; h_r = sideY[t / 64]; h_c = 2 + t % 64;
        } else if ((t -= nP2) < nU) {
            h_kind = 1; h_c = sideX[t / NW]; h_r = t % NW;
        } else if ((t -= nU) < nV) {
            h_kind = 2; h_r = sideY[t / 64]; h_c = t % 64;
        }
        if (h_kind >= 0) {
            // grid coordinates of the halo cell (columns x0-3+c for p/u, x0-1+c for v;
            // rows j0-3+r for p/v, j0-1+r for u), clamped into the grid
            int gx = (h_kind == 2) ? (x0 - 1 + h_c) : (x0 - 3 + h_c);
            int gy = (h_kind == 1) ? (j0 - 1 + h_r) : (j0 - 3 + h_r);
            gx = max(0, min(gx, Nx - 1));
            gy = max(0, min(gy, Ny - 1));
            h_off = gx + (long long)Nx * gy;
        }
    }

    // z pipeline: window p/w at k-1..k+3 (index 0..4), carried A-state of cell k and
    // flux through k-1/2
    R pz[5], wz[5];
    R pA = 0, wA = 0, FzP = 0, FzW = 0;
    if (inner_row) {
        R p7[7], w7[7];
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            const long long off = col + nxy * (long long)(k0 - 3 + q);
            p7[q] = CV::ld(ip[off], s_in);
            w7[q] = CV::ld(iw[off], s_in);
        }
        R pAm, wAm, pB0, wB0, tmp0, tmp1;
        weno_pair(p7[0], p7[1], p7[2], p7[3], p7[4], pAm, tmp0);   // cell k0-1
        weno_pair(w7[0], w7[1], w7[2], w7[3], w7[4], wAm, tmp1);
        weno_pair(p7[1], p7[2], p7[3], p7[4], p7[5], pA, pB0);     // cell k0
        weno_pair(w7[1], w7[2], w7[3], w7[4], w7[5], wA, wB0);
        godunov_f(pAm, pB0, wAm, wB0, rho, c, FzP, FzW);              // face k0-1/2
#pragma unroll
        for (int q = 0; q < 5; ++q) { pz[q] = p7[q + 2]; wz[q] = w7[q + 2]; }
    }

    for (int k = k0; k < kend; ++k) {
        const long long pl = nxy * (long long)k;
        // ---- fill the plane tiles
        const R pc = inner_row ? pz[1] : CV::ld(ip[col + pl], s_in);
        L.p[wy + 2][lx + 2] = pc;
        L.u[wy][lx + 2] = CV::ld(iu[col + pl], s_in);
        L.v[wy + 2][lx] = CV::ld(iv[col + pl], s_in);
        if (h_kind == 0) L.p[h_r][h_c] = CV::ld(ip[h_off + pl], s_in);
        else if (h_kind == 1) L.u[h_r][h_c] = CV::ld(iu[h_off + pl], s_in);
        else if (h_kind == 2) L.v[h_r][h_c] = CV::ld(iv[h_off + pl], s_in);
        __syncthreads();

        // ---- y states of every row (incl. the two outer rows)
        {
            R A, B, Av, Bv;
            weno_pair(L.p[wy][lx + 2], L.p[wy + 1][lx + 2], pc, L.p[wy + 3][lx + 2], L.p[wy + 4][lx + 2], A, B);
            weno_pair(L.v[wy][lx], L.v[wy + 1][lx], L.v[wy + 2][lx], L.v[wy + 3][lx], L.v[wy + 4][lx], Av, Bv);
            L.ys[0][wy][lx] = A;
            L.ys[1][wy][lx] = B;
            L.ys[2][wy][lx] = Av;
            L.ys[3][wy][lx] = Bv;
        }
        // ---- x faces (shuffles are wave-wide: every lane of an inner row takes part)
        R FxP_p = 0, FxP_u = 0, FxM_p = 0, FxM_u = 0, uc = 0;
        if (inner_row) {
            uc = L.u[wy][lx + 2];
            R A, B, Au, Bu;
            weno_pair(L.p[wy + 2][lx], L.p[wy + 2][lx + 1], pc, L.p[wy + 2][lx + 3], L.p[wy + 2][lx + 4], A, B);
            weno_pair(L.u[wy][lx], L.u[wy][lx + 1], uc, L.u[wy][lx + 3], L.u[wy][lx + 4], Au, Bu);
            const R Bn = shfl_next(B), Bun = shfl_next(Bu);
            godunov_f(A, Bn, Au, Bun, rho, c, FxP_p, FxP_u);          // face i+1/2
            FxM_p = shfl_prev(FxP_p);
            FxM_u = shfl_prev(FxP_u);
        }
        __syncthreads();

        if (inner_row) {
            // ---- y faces j+1/2 and j-1/2
            R FyP_p, FyP_v, FyM_p, FyM_v;
            godunov_f(L.ys[0][wy][lx], L.ys[1][wy + 1][lx], L.ys[2][wy][lx], L.ys[3][wy + 1][lx], rho, c, FyP_p, FyP_v);
            godunov_f(L.ys[0][wy - 1][lx], L.ys[1][wy][lx], L.ys[2][wy - 1][lx], L.ys[3][wy][lx], rho, c, FyM_p, FyM_v);
            const R vc = L.v[wy + 2][lx];
            // ---- z face k+1/2 from the new cell k+1
            R pAn, pBn, wAn, wBn;
            weno_pair(pz[0], pz[1], pz[2], pz[3], pz[4], pAn, pBn);
            weno_pair(wz[0], wz[1], wz[2], wz[3], wz[4], wAn, wBn);
            R FzP_p, FzP_w;
            godunov_f(pA, pBn, wA, wBn, rho, c, FzP_p, FzP_w);
            const R wc = wz[1];
            const long long C = col + pl;
            if (out_lane && upd[C] == 1) {
                const R sp = -FxP_p + FxM_p - FyP_p + FyM_p - FzP_p + FzP;
                R uo, vo, wo, po, wgt;
                if (STAGE == 0) {
                    uo = uc + dt * (-FxP_u + FxM_u) / h;
                    vo = vc + dt * (-FyP_v + FyM_v) / h;
                    wo = wc + dt * (-FzP_w + FzW) / h;
                    po = pc + dt * sp / h;
                    wgt = R(1);
                } else {
                    const R p0 = CV::ld(bp[C], s_base), u0 = CV::ld(bu[C], s_base);
                    const R v0 = CV::ld(bv[C], s_base), w0 = CV::ld(bw[C], s_base);
                    if (STAGE == 1) {
                        uo = R(0.75) * u0 + R(0.25) * (uc + dt * ((-FxP_u + FxM_u) / h));
                        vo = R(0.75) * v0 + R(0.25) * (vc + dt * ((-FyP_v + FyM_v) / h));
                        wo = R(0.75) * w0 + R(0.25) * (wc + dt * ((-FzP_w + FzW) / h));
                        po = R(0.75) * p0 + R(0.25) * (pc + dt * (sp / h));
                        wgt = R(0.25);
                    } else {
                        uo = (u0 + R(2) * (uc + dt * ((-FxP_u + FxM_u) / h))) / R(3);
                        vo = (v0 + R(2) * (vc + dt * ((-FyP_v + FyM_v) / h))) / R(3);
                        wo = (w0 + R(2) * (wc + dt * ((-FzP_w + FzW) / h))) / R(3);
                        po = (p0 + R(2) * (pc + dt * (sp / h))) / R(3);
                        wgt = R(2) / R(3);
                    }
                }
                if (pml.on)
                    pml_apply<R, STAGE == 2>(pml, i, j, k, Nx, Ny, po, uo, vo, wo, (-FxP_p + FxM_p) / h,
                                             (-FyP_p + FyM_p) / h, (-FzP_p + FzP) / h, (-FxP_u + FxM_u) / h,
                                             (-FyP_v + FyM_v) / h, (-FzP_w + FzW) / h, wgt * dt);
                ou[C] = CV::st(uo, is_out);
                ov[C] = CV::st(vo, is_out);
                ow[C] = CV::st(wo, is_out);
                op[C] = CV::st(po, is_out);
            }
            // advance the z pipeline
            pA = pAn; wA = wAn; FzP = FzP_p; FzW = FzP_w;
#pragma unroll
            for (int q = 0; q < 4; ++q) { pz[q] = pz[q + 1]; wz[q] = wz[q + 1]; }
            if (k + 1 < kend) {
                const long long off = col + nxy * (long long)(k + 4);
                pz[4] = CV::ld(ip[off], s_in);
                wz[4] = CV::ld(iw[off], s_in);
            }
        }
    }
}

}  // namespace fasesdev
