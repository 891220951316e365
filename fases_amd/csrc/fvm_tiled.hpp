// Tiled flux-stage kernel for gfx950: WENO5-Z + Godunov + SSP-RK3 stage + cPML with
// every face reconstructed once.
//
// Block = 64 lanes x NW waves (y rows j0 .. j0+NW-1), marching over tz planes in z, in
// one of two widths TX (picked per room by the autotuner; both give the same bits):
//   TX = 64: lanes are x columns x0 .. x0+63, every lane an output; the fluxes through
//            the tile's two x edges come from one wave that reconstructs the edge states
//            (A of x0-1 and x0+63, B of x0 and x0+64, p and u of every row) for the block;
//   TX = 62: lanes are x columns x0-1 .. x0+62, outputs the inner 62; the outer lanes
//            only supply x-face states to their neighbours (no edge work, but 2 of 64
//            lanes idle and more tiles on rooms whose width is a multiple of 64).
//   z faces: register pipeline along the march (face k+1/2 computed once, carried);
//   x faces: each lane reconstructs both states of its own cell, the right state of the
//            next cell comes from the neighbouring lane (shuffle), and so does the flux
//            through the cell's left face;
//   y faces: both states per cell written to LDS, read by the rows above and below; the
//            one state the block needs of each halo row (left of j0-1, right of j0+NW, p
//            and v) comes from four waves, 0, 1, NW-2, NW-1, i.e. one per SIMD.
// p, u (x) and p, v (y) of the current plane sit in LDS with a 3-cell halo; p and w
// along z stay in registers.
//
// WENO5-Z arithmetic (functions3D.cuh:763-977, eps = 1e-10, q = 1) is restated on
// first differences of the 5-cell window: candidates c + linear combinations of the
// differences, smoothness indicators 13/12 s^2 + 1/4 t^2 from second differences s and
// the one-sided combinations t, weights d_r (1 + tau/(beta_r + eps)), normalised once.
#pragma once

#include <climits>
#include <type_traits>

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

// Smoothness term of a 3-cell stencil with second difference s: 13/3 s^2 + 4 eps. Each
// stencil serves three cells (as beta_0, beta_1, beta_2), so a pipeline along an axis
// computes it once per cell and carries it to the next two.
template <typename R>
__device__ inline R weno_q(R s) {
    const R k = R(13) / R(3);
    return fma(k * s, s, R(4e-10));
}

// Unnormalised WENO-Z weights of a cell from the first differences d1..d4 of its 5-cell
// window (a,b,c,d,e): d1 = b - a, ..., d4 = e - d, and the smoothness terms
// q0 = weno_q(d4 - d3), q1 = weno_q(d3 - d2), q2 = weno_q(d2 - d1).
//   4 x smoothness indicators: e_r = 13/3 s^2 + t^2 + 4 eps (the weights only see ratios of
//   beta + eps and |beta_0 - beta_2|, so beta, tau and eps all carry the factor 4);
//   d_r (1 + tau/e_r) times e0 e1 e2: g_r = (e_r + tau) prod_{s != r} e_s.
// Reassociations of this arithmetic were measured in FP32 against FP64 on a 4000-step
// room (IR) and on the eps-level gradient of a 96x80x64 room: g_r = prod e_s tau +
// e0 e1 e2 (IR error 2.7e-5 instead of 8.0e-6) and denominators sharing 0.1 (g0 + g2) +
// 0.6 g1 (gradient error 3.8e-3 instead of 3.1e-4) each fail one of the two bars, so
// both states keep their own sums.
template <typename R> struct WenoG {
    R g0, g1, g2;
};
template <typename R>
__device__ inline WenoG<R> weno_g(R d1, R d2, R d3, R d4, R q0, R q1, R q2) {
    const R t0 = fma(R(-3), d3, d4), t1 = d2 + d3, t2 = fma(R(3), d2, -d1);
    const R e0 = fma(t0, t0, q0);
    const R e1 = fma(t1, t1, q1);
    const R e2 = fma(t2, t2, q2);
    const R tau = fabs(e0 - e2);
    return {(e0 + tau) * (e1 * e2), (e1 + tau) * (e0 * e2), (e2 + tau) * (e0 * e1)};
}
// Candidate offsets from the centre c, pre-multiplied by the linear weights:
//   left  (V^- at i+1/2): d = 3/10, 3/5, 1/10 on {1/3,5/6,-1/6}(c,d,e), {-1/6,5/6,1/3}(b,c,d),
//                          {1/3,-7/6,11/6}(a,b,c)
//   right (V^+ at i-1/2): d = 1/10, 3/5, 3/10 on {11/6,-7/6,1/3}(c,d,e), {1/3,5/6,-1/6}(b,c,d),
//                          {-1/6,5/6,1/3}(a,b,c)
template <typename R>
__device__ inline R weno_left(const WenoG<R> &g, R d1, R d2, R d3, R d4, R c) {
    const R a0 = fma(R(0.2), d3, R(-0.05) * d4);
    const R a1 = fma(R(0.1), d2, R(0.2) * d3);
    const R a2 = fma(R(1.0 / 12.0), d2, R(-1.0 / 30.0) * d1);
    const R nl = fma(g.g0, a0, fma(g.g1, a1, g.g2 * a2));
    return fma(nl, fast_rcp(fma(R(0.3), g.g0, fma(R(0.6), g.g1, R(0.1) * g.g2))), c);
}
// (weno_right is not written as the operation-for-operation mirror of weno_left: that
// order was measured to move the FP32 gradient of the 96x80x64 eps-level adjoint test
// from 3e-4 to 1.1e-2 of max |g| against FP64, on the one wall whose gradient is most
// sensitive to the forward's rounding.)
template <typename R>
__device__ inline R weno_right(const WenoG<R> &g, R d1, R d2, R d3, R d4, R c) {
    const R b0 = fma(R(-1.0 / 12.0), d3, R(1.0 / 30.0) * d4);
    const R b1 = fma(R(-0.2), d2, R(-0.1) * d3);
    const R b2 = fma(R(-0.2), d2, R(0.05) * d1);
    const R nr = fma(g.g0, b0, fma(g.g1, b1, g.g2 * b2));
    return fma(nr, fast_rcp(fma(R(0.1), g.g0, fma(R(0.6), g.g1, R(0.3) * g.g2))), c);
}

// Both WENO5-Z states of a cell from the differences and smoothness terms of its window
// and its centre c: A = V^-_{i+1/2}, B = V^+_{i-1/2}. Every caller goes through these
// functions, so a face state is the same value whether its differences were computed here
// or carried along an axis, and whether the cell's other state was wanted or not.
template <typename R>
__device__ inline void weno_core(R d1, R d2, R d3, R d4, R q0, R q1, R q2, R c, R &A, R &B) {
    const WenoG<R> g = weno_g(d1, d2, d3, d4, q0, q1, q2);
    A = weno_left(g, d1, d2, d3, d4, c);
    B = weno_right(g, d1, d2, d3, d4, c);
}

// Both WENO5-Z states of the cell whose 5-cell window is (a,b,c,d,e), centre c.
template <typename R>
__device__ inline void weno_pair(R a, R b, R c, R d, R e, R &A, R &B) {
    const R d1 = b - a, d2 = c - b, d3 = d - c, d4 = e - d;
    weno_core(d1, d2, d3, d4, weno_q(d4 - d3), weno_q(d3 - d2), weno_q(d2 - d1), c, A, B);
}
// Only the left state A = V^-_{i+1/2} (bit-identical to weno_pair's A) ...
template <typename R>
__device__ inline R weno_left_of(R a, R b, R c, R d, R e) {
    const R d1 = b - a, d2 = c - b, d3 = d - c, d4 = e - d;
    const WenoG<R> g = weno_g(d1, d2, d3, d4, weno_q(d4 - d3), weno_q(d3 - d2), weno_q(d2 - d1));
    return weno_left(g, d1, d2, d3, d4, c);
}
// ... and only the right state B = V^+_{i-1/2}.
template <typename R>
__device__ inline R weno_right_of(R a, R b, R c, R d, R e) {
    const R d1 = b - a, d2 = c - b, d3 = d - c, d4 = e - d;
    const WenoG<R> g = weno_g(d1, d2, d3, d4, weno_q(d4 - d3), weno_q(d3 - d2), weno_q(d2 - d1));
    return weno_right(g, d1, d2, d3, d4, c);
}

// One variable's window along the z march, carried from plane to plane: the centres
// u[k], u[k+1], u[k+2], the differences d1 = u[k]-u[k-1], d2, d3 = u[k+2]-u[k+1] and the
// smoothness terms of the stencils centred on k+1 (q1) and k (q2). With u[k+3] the pair
// of cell k+1 needs one new difference and one new smoothness term.
template <typename R> struct ZCarry {
    R c0, c1, c2, d1, d2, d3, q1, q2;
    __device__ inline void init(const R *u) {      // u[0..5] = planes k-3 .. k+2
        c0 = u[3]; c1 = u[4]; c2 = u[5];
        d1 = u[3] - u[2]; d2 = u[4] - u[3]; d3 = u[5] - u[4];
        q1 = weno_q(d3 - d2);
        q2 = weno_q(d2 - d1);
    }
    // states of cell k+1 from u[k+3]; then the window moves one plane on
    __device__ inline void step(R un, R &A, R &B) {
        const R d4 = un - c2;
        const R q0 = weno_q(d4 - d3);
        weno_core(d1, d2, d3, d4, q0, q1, q2, c1, A, B);
        c0 = c1; c1 = c2; c2 = un;
        d1 = d2; d2 = d3; d3 = d4;
        q2 = q1; q1 = q0;
    }
};

namespace tile {
constexpr int XHW = 2;          // wave that reconstructs the x-edge states (TX = 64)
// LDS columns of a tile of width TX: x0-3 .. x0+TX+2 (the own cell of lane l at column
// l + halo_x: 3 for TX = 64, 2 for TX = 62)
template <int TX> constexpr int lds_x() { return TX + 6; }
template <int TX> constexpr int halo_x() { return (TX + 6 - 64) / 2; }
}  // namespace tile

template <typename R, int NW, int TX> struct TileLds {
    R p[NW + 6][tile::lds_x<TX>()];      // rows j0-3 .. j0+NW+2
    R u[NW][tile::lds_x<TX>()];
    R v[NW + 6][64];
    R ys[4][NW + 2][64];        // y states p^-, p^+, v^-, v^+ of rows j0-1 .. j0+NW
    R xs[4][2][NW];             // x-edge states (p, u) of each row: A(x0-1), B(x0), A(x0+63), B(x0+64)
    R xf[2][2][NW];             // x-edge fluxes (p, u) of each row through x0-1/2 and x0+63+1/2
    R spare[64];                // target of the halo slots that carry no item
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

// Plane-relative buffer access: a resource descriptor whose base is one z-plane of a
// field (uniform, built in scalar registers) and the lane's 32-bit byte offset within
// the plane, so a load or store costs no per-lane 64-bit address arithmetic. The
// descriptor's range is the plane: the hardware returns 0 for an offset at or beyond
// it and drops such a store. (Data-format word of the gfx9 family; stride 0 = raw.)
__device__ inline __amdgpu_buffer_rsrc_t plane_rsrc(const void *plane, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(plane), (short)0, (int)bytes, 0x00020000);
}
template <typename T> struct PlaneIO;
template <> struct PlaneIO<float> {
    __device__ static inline float ld(__amdgpu_buffer_rsrc_t r, unsigned off) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
    }
    __device__ static inline void st(__amdgpu_buffer_rsrc_t r, unsigned off, float x) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, (int)off, 0, 0);
    }
};
template <> struct PlaneIO<double> {
    __device__ static inline double ld(__amdgpu_buffer_rsrc_t r, unsigned off) {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
    }
    __device__ static inline void st(__amdgpu_buffer_rsrc_t r, unsigned off, double x) {
        typedef unsigned u2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, x), r, (int)off, 0, 0);
    }
};
template <> struct PlaneIO<__half> {
    __device__ static inline __half ld(__amdgpu_buffer_rsrc_t r, unsigned off) {
        return __builtin_bit_cast(__half, __builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, 0));
    }
    __device__ static inline void st(__amdgpu_buffer_rsrc_t r, unsigned off, __half x) {
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, x), r, (int)off, 0, 0);
    }
};
__device__ inline uint8_t plane_ld_u8(const uint8_t *plane, unsigned bytes, unsigned off) {
    return __builtin_amdgcn_raw_buffer_load_b8(plane_rsrc(plane, bytes), (int)off, 0, 0);
}

// Field pointers of one tiled stage launch: input set, base set U0, output set.
template <typename S> struct TileFields {
    const S *ip, *iu, *iv, *iw;
    const S *bp, *bu, *bv, *bw;
    S *op, *ou, *ov, *ow;
};

// cPML factors of a tile's own column (x) and row (y), constant along the z march: band
// index (-1 outside), 1/kappa - 1, b and a of the psi update (pml_apply's tables).
template <typename R> struct PmlLane {
    int bx, by;
    R kx, bbx, ax, ky, bby, ay;
};
template <typename R>
__device__ inline PmlLane<R> pml_lane(const PmlSet &pm, int ic, int jc) {
    PmlLane<R> L{-1, -1, R(0), R(0), R(0), R(0), R(0), R(0)};
    L.bx = pm.x.band[ic];
    if (L.bx >= 0) { L.kx = (R)pm.x.kc[ic]; L.bbx = (R)pm.x.b[ic]; L.ax = (R)pm.x.a[ic]; }
    L.by = pm.y.band[jc];
    if (L.by >= 0) { L.ky = (R)pm.y.kc[jc]; L.bby = (R)pm.y.b[jc]; L.ay = (R)pm.y.a[jc]; }
    return L;
}
// psi of the lane's x band (px, ux) and the wave's y band (py, vy) at plane k, in compute
// precision (loaded a plane ahead by the tile).
template <typename R> struct PmlPsi {
    R px, ux, py, vy;
};
template <typename R>
__device__ inline PmlPsi<R> pml_psi_xy(const PmlSet &pml, const PmlLane<R> &PL, int i, int j, int k, int Nx, int Ny) {
    PmlPsi<R> q{R(0), R(0), R(0), R(0)};
    if (PL.bx >= 0) {
        const long long X = PL.bx + (long long)(2 * pml.d) * (j + (long long)Ny * k);
        q.px = (R)pml.psi_px[X];
        q.ux = (R)pml.psi_ux[X];
    }
    if (PL.by >= 0) {
        const long long Y = i + (long long)Nx * (PL.by + (long long)(2 * pml.d) * k);
        q.py = (R)pml.psi_py[Y];
        q.vy = (R)pml.psi_vy[Y];
    }
    return q;
}
// pml_apply with the x and y factors taken from PmlLane and the x, y psi from PmlPsi
// instead of the tables: the same operations in the same order, so the same bits.
template <typename R, bool ADVANCE>
__device__ inline void pml_tile(const PmlSet &pml, const PmlLane<R> &PL, const PmlPsi<R> &Q, bool z_band, int i, int j,
                                int k, int Nx, int Ny, R &po, R &uo, R &vo, R &wo, R Lpx, R Lpy, R Lpz, R Lux, R Lvy,
                                R Lwz, R wdt) {
#pragma clang fp contract(off)
    if (PL.bx >= 0) {
        const long long X = PL.bx + (long long)(2 * pml.d) * (j + (long long)Ny * k);
        const R px = Q.px, ux = Q.ux;
        po = po + wdt * (PL.kx * Lpx + px);
        uo = uo + wdt * (PL.kx * Lux + ux);
        if (ADVANCE) {
            pml.psi_px[X] = (double)(PL.bbx * px + PL.ax * Lpx);
            pml.psi_ux[X] = (double)(PL.bbx * ux + PL.ax * Lux);
        }
    }
    if (PL.by >= 0) {
        const long long Y = i + (long long)Nx * (PL.by + (long long)(2 * pml.d) * k);
        const R py = Q.py, vy = Q.vy;
        po = po + wdt * (PL.ky * Lpy + py);
        vo = vo + wdt * (PL.ky * Lvy + vy);
        if (ADVANCE) {
            pml.psi_py[Y] = (double)(PL.bby * py + PL.ay * Lpy);
            pml.psi_vy[Y] = (double)(PL.bby * vy + PL.ay * Lvy);
        }
    }
    const int bz = z_band ? pml.z.band[k] : -1;    // planes clear of the z bands skip the lookup
    if (bz >= 0) {
        const long long Z = i + (long long)Nx * (j + (long long)Ny * bz);
        const R kc = (R)pml.z.kc[k];
        const R pz = (R)pml.psi_pz[Z], wz = (R)pml.psi_wz[Z];
        po = po + wdt * (kc * Lpz + pz);
        wo = wo + wdt * (kc * Lwz + wz);
        if (ADVANCE) {
            const R b = (R)pml.z.b[k], a = (R)pml.z.a[k];
            pml.psi_pz[Z] = (double)(b * pz + a * Lpz);
            pml.psi_wz[Z] = (double)(b * wz + a * Lwz);
        }
    }
}

// One tile = 64 x NW threads; outputs TX columns x NW rows x tz planes of tile
// (bx, by, bz) of an Nx x Ny x Nz grid.
template <typename S, typename R, int STAGE, int NW, int TX, bool PML_FUSED>
__device__ inline void flux_tile(TileLds<R, NW, TX> &L, int Nx, int Ny, int Nz, const TileFields<S> &F,
                                 const uint8_t *__restrict__ upd, const PmlSet *__restrict__ pmlp, const TileK<R> &K, R dt,
                                 const Scl<R> &sc, int tz, int bx, int by, int bz) {
    using namespace tile;
    static_assert(TX == 64 || TX == 62, "tile width");
    constexpr int LX = lds_x<TX>(), HX = halo_x<TX>();
    typedef Conv<S, R> CV;
    const S *__restrict__ ip = F.ip, *__restrict__ iu = F.iu, *__restrict__ iv = F.iv, *__restrict__ iw = F.iw;
    const S *__restrict__ bp = F.bp, *__restrict__ bu = F.bu, *__restrict__ bv = F.bv, *__restrict__ bw = F.bw;
    S *__restrict__ op = F.op, *__restrict__ ou = F.ou, *__restrict__ ov = F.ov, *__restrict__ ow = F.ow;
    // blockDim.x is one wavefront, so the row index is uniform per wave: keep it scalar
    const int lx = threadIdx.x, wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int tid = wy * 64 + lx;
    const int x0 = 3 + TX * bx, j0 = 3 + NW * by, k0 = 3 + tz * bz;
    const long long nxy = (long long)Nx * Ny;
    const int kend = min(k0 + tz, Nz - 3);
    const int i = x0 - 3 + HX + lx, j = j0 + wy;      // lane lx sits at LDS column lx + HX
    const int ic = min(i, Nx - 1), jc = min(j, Ny - 1);               // clamped for loads
    const unsigned col = (unsigned)(ic + Nx * jc);                   // offset within a plane
    const unsigned cb = col * (unsigned)sizeof(S);                   // ... in bytes
    const unsigned pbytes = (unsigned)nxy * (unsigned)sizeof(S);     // bytes of one plane
    const bool out_lane = (TX == 64 || (lx >= 1 && lx <= 62)) && i < Nx - 3 && j < Ny - 3;
    // field value of this lane's column in the plane at element offset `pl`
    auto LD = [&](const S *f, long long pl, R scale) __device__ {
        return CV::ld(PlaneIO<S>::ld(plane_rsrc(f + pl, pbytes), cb), scale);
    };
    // ... the stored value itself (converted where it is used, so the wait on the load
    // lands there and not right after issue)
    auto LDR = [&](const S *f, long long pl) __device__ { return PlaneIO<S>::ld(plane_rsrc(f + pl, pbytes), cb); };
    auto ST = [&](S *f, long long pl, R x, R inv_scale) __device__ {
        PlaneIO<S>::st(plane_rsrc(f + pl, pbytes), cb, CV::st(x, inv_scale));
    };

    // halo items of this thread: items tid, tid + 64 NW, ... of
    //   p: rows 0-2 and NW+3..NW+5 (all LX columns), the HX columns of each side of the own
    //   rows;  u: the same side columns of the own rows;  v: rows 0-2 and NW+3..NW+5.
    // Each field's items start on a multiple of 64, so the field of a slot is uniform over
    // a wave: its loads are buffer loads on the plane descriptor (scalar base, per-lane
    // offset). A slot without an item loads at the plane's end (the hardware returns 0)
    // and stores into the spare row.
    constexpr int nP1 = 6 * LX, nP2 = 2 * HX * NW, nU = 2 * HX * NW, nV = 6 * 64;
    constexpr int kPEnd = (nP1 + nP2 + 63) / 64 * 64, kUEnd = kPEnd + (nU + 63) / 64 * 64, kVEnd = kUEnd + nV;
    constexpr int NH = (kVEnd + 64 * NW - 1) / (64 * NW);
    int h_kind[NH];             // 0 p, 1 u, 2 v, 3 none (wave-uniform)
    const S *hsrc[NH];          // its field (wave-uniform, scalar registers)
    R hsc[NH];                  // and scale
    R *hdst[NH];
    unsigned h_off[NH];         // byte offset within a plane
#pragma unroll
    for (int q = 0; q < NH; ++q) {
        const int sideX[6] = {0, 1, HX == 3 ? 2 : LX - 2, HX == 3 ? LX - 3 : LX - 1, LX - 2, LX - 1};
        const int sideY[6] = {0, 1, 2, NW + 3, NW + 4, NW + 5};
        const int t = tid + q * 64 * NW;
        const int tw = __builtin_amdgcn_readfirstlane(t);
        const int kind = tw < kPEnd ? 0 : (tw < kUEnd ? 1 : (tw < kVEnd ? 2 : 3));
        int h_r = -1, h_c = 0;
        if (kind == 0) {
            if (t < nP1) {
                h_r = sideY[t / LX]; h_c = t % LX;
            } else if (t - nP1 < nP2) {
                h_c = sideX[(t - nP1) / NW]; h_r = 3 + (t - nP1) % NW;
            }
        } else if (kind == 1) {
            if (t - kPEnd < nU) { h_c = sideX[(t - kPEnd) / NW]; h_r = (t - kPEnd) % NW; }
        } else if (kind == 2) {
            h_r = sideY[(t - kUEnd) / 64]; h_c = (t - kUEnd) % 64;
        }
        h_kind[q] = kind;
        {
            const unsigned long long a = (unsigned long long)(kind == 0 ? ip : (kind == 1 ? iu : iv));
            const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
            const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
            hsrc[q] = (const S *)(((unsigned long long)hi << 32) | lo);
            hsc[q] = kind == 0 ? sc.s[0] : (kind == 1 ? sc.s[1] : sc.s[2]);
        }
        h_off[q] = (unsigned)nxy * (unsigned)sizeof(S);             // out of range: reads 0
        hdst[q] = &L.spare[lx];
        if (h_r >= 0) {
            int gx = (kind == 2) ? (x0 - 3 + HX + h_c) : (x0 - 3 + h_c);
            int gy = (kind == 1) ? (j0 + h_r) : (j0 - 3 + h_r);
            gx = max(0, min(gx, Nx - 1));
            gy = max(0, min(gy, Ny - 1));
            h_off[q] = (unsigned)(gx + Nx * gy) * (unsigned)sizeof(S);
            hdst[q] = (kind == 0) ? &L.p[h_r][h_c] : (kind == 1 ? &L.u[h_r][h_c] : &L.v[h_r][h_c]);
        }
    }
    auto LDH = [&](int q, long long pl) __device__ { return PlaneIO<S>::ld(plane_rsrc(hsrc[q] + pl, pbytes), h_off[q]); };

    // z pipeline: carried windows of p and w (centres k..k+2, differences, smoothness
    // terms), the newest plane k+3 (pn, wn), the A-state of cell k and the flux through k-1/2
    ZCarry<R> zp, zw;
    R pn = R(0), wn = R(0);
    R pA, wA, FzP, FzW;
    {
        R p7[6], w7[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const long long pq = nxy * (long long)(k0 - 3 + q);
            p7[q] = LD(ip, pq, sc.s[0]);
            w7[q] = LD(iw, pq, sc.s[3]);
        }
        R pAm, wAm, pB0, wB0, tmp0, tmp1;
        weno_pair(p7[0], p7[1], p7[2], p7[3], p7[4], pAm, tmp0);   // cell k0-1
        weno_pair(w7[0], w7[1], w7[2], w7[3], w7[4], wAm, tmp1);
        weno_pair(p7[1], p7[2], p7[3], p7[4], p7[5], pA, pB0);     // cell k0
        weno_pair(w7[1], w7[2], w7[3], w7[4], w7[5], wA, wB0);
        godunov_f(pAm, pB0, wAm, wB0, K, FzP, FzW);            // face k0-1/2
        zp.init(p7);
        zw.init(w7);
    }
    // values in flight, in storage precision: own u, v, update flag, halo items, and p, w
    // of plane k+3 for the z window. Each plane converts what
    // the previous plane loaded and at once issues the loads of the next one, so a whole
    // plane of work covers their HBM latency.
    constexpr bool kRaw = !std::is_same<S, R>::value;      // storage needs a conversion
    S nu_s, nv_s, nh_s[NH], pend_p, pend_w;
    unsigned nflag;
    {
        const long long pl = nxy * (long long)k0;
        nu_s = LDR(iu, pl);
        nv_s = LDR(iv, pl);
        nflag = plane_ld_u8(upd + pl, (unsigned)nxy, col);
#pragma unroll
        for (int q = 0; q < NH; ++q)                       // per-lane field: no uniform base
            if (h_kind[q] != 3) nh_s[q] = LDH(q, pl);
        const long long p3 = nxy * (long long)(k0 + 3);
        pend_p = LDR(ip, p3);
        pend_w = LDR(iw, p3);
        if constexpr (!kRaw) {                 // this order converts at the plane's end
            pn = CV::ld(pend_p, sc.s[0]);
            wn = CV::ld(pend_w, sc.s[3]);
        }
    }
    // cPML: whether the tile's columns or rows reach a band (uniform per block); planes
    // k < pz_lo or k >= pz_hi reach one in z, so tiles clear of every band skip pml_apply.
    // The tables stay in memory (read only in that branch), which keeps the plane loop's
    // scalar registers for the field planes.
    int pz_lo = INT_MIN, pz_hi = INT_MAX;           // no plane
    int zc0 = INT_MIN, zc1 = INT_MAX;               // planes [zc0, zc1) lie in no z band
    PmlLane<R> PL{-1, -1, R(0), R(0), R(0), R(0), R(0), R(0)};
    if constexpr (PML_FUSED) {
        const PmlSet &pm = *pmlp;
        if (pm.on) {
            const bool pml_xy = x0 < pm.clear0[0] || x0 + TX > pm.clear1[0] || j0 < pm.clear0[1] ||
                                j0 + NW > pm.clear1[1];
            pz_lo = pml_xy ? INT_MAX : pm.clear0[2];    // every plane / the z bands
            pz_hi = pml_xy ? INT_MIN : pm.clear1[2];
            zc0 = pm.clear0[2];
            zc1 = pm.clear1[2];
            PL = pml_lane<R>(pm, ic, jc);
        }
    }
    // tiles reaching an x or y band prefetch the lane's psi of the next plane (their cells
    // otherwise wait on the psi loads inside the plane)
    const bool pml_xy_tile = PML_FUSED && pz_lo == INT_MAX;
    PmlPsi<R> psi_n{R(0), R(0), R(0), R(0)};
    if (pml_xy_tile && k0 < kend) psi_n = pml_psi_xy<R>(*pmlp, PL, ic, jc, k0, Nx, Ny);

    // One plane k.
    auto plane = [&](int k) __device__ __attribute__((always_inline)) {
        const long long pl = nxy * (long long)k;
        // ---- the values the previous plane loaded, into compute precision. With reduced
        // storage the z window's newest plane k+3 is converted here, and the next plane's
        // loads go out at once, a whole plane ahead of their conversion; with FP32 / FP64
        // storage (no conversion) the loads go out after the first barrier and the value is
        // taken at the end of the plane (measured faster each way).
        if constexpr (kRaw) {
            pn = CV::ld(pend_p, sc.s[0]);
            wn = CV::ld(pend_w, sc.s[3]);
        }
        const R pc = zp.c0, wc = zw.c0;
        const R uc = CV::ld(nu_s, sc.s[1]), vc = CV::ld(nv_s, sc.s[2]);

        const unsigned flag = nflag;
        const PmlPsi<R> psi_c = psi_n;
        R hv[NH];
#pragma unroll
        for (int q = 0; q < NH; ++q) hv[q] = h_kind[q] != 3 ? CV::ld(nh_s[q], hsc[q]) : R(0);
        // ---- the next plane's loads (p, w of plane k+4 for the z window)
        auto issue_next = [&]() __device__ __attribute__((always_inline)) {
            if (k + 1 >= kend) return;
            const long long pn = pl + nxy;
            nu_s = LDR(iu, pn);
            nv_s = LDR(iv, pn);
            nflag = plane_ld_u8(upd + pn, (unsigned)nxy, col);
#pragma unroll
            for (int q = 0; q < NH; ++q)
                if (h_kind[q] != 3) nh_s[q] = LDH(q, pn);
            const long long p4 = nxy * (long long)(k + 4);
            pend_p = LDR(ip, p4);
            pend_w = LDR(iw, p4);
            if (pml_xy_tile) psi_n = pml_psi_xy<R>(*pmlp, PL, ic, jc, k + 1, Nx, Ny);
        };
        if constexpr (kRaw) issue_next();
        // ---- fill the plane tiles
        L.p[wy + 3][lx + HX] = pc;
        L.u[wy][lx + HX] = uc;
        L.v[wy + 3][lx] = vc;
#pragma unroll
        for (int q = 0; q < NH; ++q)
            if (h_kind[q] != 3) *hdst[q] = hv[q];
        __syncthreads();
        if constexpr (!kRaw) issue_next();
        // RK base set of this plane: loaded here (TX = 62) or after the second barrier (TX =
        // 64, whose edge work would otherwise keep the four values live across more code)
        S b0 = S(), b1 = S(), b2 = S(), b3 = S();
        auto load_base = [&]() __device__ __attribute__((always_inline)) {
            if (STAGE != 0) {
                b0 = LDR(bp, pl);
                b1 = LDR(bu, pl);
                b2 = LDR(bv, pl);
                b3 = LDR(bw, pl);
            }
        };
        if constexpr (TX == 62) load_base();
        // ---- y states of the own row, and of the halo rows on four waves
        {
            R A, B, Av, Bv;
            weno_pair(L.p[wy + 1][lx + HX], L.p[wy + 2][lx + HX], pc, L.p[wy + 4][lx + HX], L.p[wy + 5][lx + HX], A, B);
            weno_pair(L.v[wy + 1][lx], L.v[wy + 2][lx], vc, L.v[wy + 4][lx], L.v[wy + 5][lx], Av, Bv);
            L.ys[0][wy + 1][lx] = A;
            L.ys[1][wy + 1][lx] = B;
            L.ys[2][wy + 1][lx] = Av;
            L.ys[3][wy + 1][lx] = Bv;
            // halo rows: the block needs the left state of row j0-1 (face j0-1/2) and the
            // right state of row j0+NW (face j0+NW-1/2), one reconstruction per SIMD
            if (wy == 0) {                           // p of row j0-1 (ys row 0, L.p rows 0..4)
                L.ys[0][0][lx] = weno_left_of(L.p[0][lx + HX], L.p[1][lx + HX], L.p[2][lx + HX], L.p[3][lx + HX],
                                              L.p[4][lx + HX]);
            } else if (wy == 1) {                    // v of row j0-1
                L.ys[2][0][lx] = weno_left_of(L.v[0][lx], L.v[1][lx], L.v[2][lx], L.v[3][lx], L.v[4][lx]);
            } else if (wy == NW - 2) {               // p of row j0+NW (ys row NW+1, L.p rows NW+1..NW+5)
                constexpr int r = NW + 1;
                L.ys[1][r][lx] = weno_right_of(L.p[r][lx + HX], L.p[r + 1][lx + HX], L.p[r + 2][lx + HX],
                                               L.p[r + 3][lx + HX], L.p[r + 4][lx + HX]);
            } else if (wy == NW - 1) {               // v of row j0+NW
                constexpr int r = NW + 1;
                L.ys[3][r][lx] = weno_right_of(L.v[r][lx], L.v[r + 1][lx], L.v[r + 2][lx], L.v[r + 3][lx],
                                               L.v[r + 4][lx]);
            } else if (TX == 64 && wy == XHW) {
                // x-edge faces of every row, x0-1/2 and x0+63+1/2: the four states
                // A(x0-1), B(x0), A(x0+63), B(x0+64) of p and u (weno_pair of the cell's
                // window, the same bits as the rows' own lanes), then the two fluxes, for
                // the rows' lanes 0 and 63
#pragma unroll
                for (int t = lx; t < 8 * NW; t += 64) {
                    const int r = t % NW, var = (t / NW) & 1, st = t / (2 * NW);
                    // window columns 0..4, 1..5, 64..68, 65..69; st odd: the right state
                    const R *w = (var ? &L.u[r][0] : &L.p[r + 3][0]) + ((st & 2) ? LX - 6 : 0) + (st & 1);
                    R A, B;
                    weno_pair(w[0], w[1], w[2], w[3], w[4], A, B);
                    L.xs[st][var][r] = (st & 1) ? B : A;
                }
                __builtin_amdgcn_wave_barrier();    // LDS operations of one wave complete in order
                if (lx < 2 * NW) {
                    const int r = lx % NW, f = lx / NW;
                    R Fp, Fu;
                    godunov_f(L.xs[2 * f][0][r], L.xs[2 * f + 1][0][r], L.xs[2 * f][1][r], L.xs[2 * f + 1][1][r], K, Fp, Fu);
                    L.xf[f][0][r] = Fp;
                    L.xf[f][1][r] = Fu;
                }
            }
        }
        // ---- x faces (shuffles are wave-wide: every lane takes part); with TX = 64, lane
        // 63's face i+1/2 and lane 0's face i-1/2 are the edge wave's fluxes
        R FxP_p, FxP_u, FxM_p, FxM_u;
        {
            R A, B, Au, Bu;
            weno_pair(L.p[wy + 3][lx + HX - 2], L.p[wy + 3][lx + HX - 1], pc, L.p[wy + 3][lx + HX + 1],
                      L.p[wy + 3][lx + HX + 2], A, B);
            weno_pair(L.u[wy][lx + HX - 2], L.u[wy][lx + HX - 1], uc, L.u[wy][lx + HX + 1], L.u[wy][lx + HX + 2], Au, Bu);
            const R Bn = shfl_next(B), Bun = shfl_next(Bu);
            godunov_f(A, Bn, Au, Bun, K, FxP_p, FxP_u);          // face i+1/2
            FxM_p = shfl_prev(FxP_p);
            FxM_u = shfl_prev(FxP_u);
        }
        __syncthreads();
        if constexpr (TX == 64) load_base();
        if constexpr (TX == 64) {
            FxM_p = lx == 0 ? L.xf[0][0][wy] : FxM_p;
            FxM_u = lx == 0 ? L.xf[0][1][wy] : FxM_u;
            FxP_p = lx == 63 ? L.xf[1][0][wy] : FxP_p;
            FxP_u = lx == 63 ? L.xf[1][1][wy] : FxP_u;
        }

        // ---- y faces j+1/2 and j-1/2
        R FyP_p, FyP_v, FyM_p, FyM_v;
        godunov_f(L.ys[0][wy + 1][lx], L.ys[1][wy + 2][lx], L.ys[2][wy + 1][lx], L.ys[3][wy + 2][lx], K, FyP_p,
                  FyP_v);
        godunov_f(L.ys[0][wy][lx], L.ys[1][wy + 1][lx], L.ys[2][wy][lx], L.ys[3][wy + 1][lx], K, FyM_p, FyM_v);
        // ---- z face k+1/2 from the new cell k+1 (the carried windows move on a plane)
        R pAn, wAn, pBn, wBn, FzP_p, FzP_w;
        zp.step(pn, pAn, pBn);
        zw.step(wn, wAn, wBn);
        godunov_f(pA, pBn, wA, wBn, K, FzP_p, FzP_w);
        if (out_lane && flag == 1) {
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
                const R p0 = CV::ld(b0, sc.s[0]), u0 = CV::ld(b1, sc.s[1]);
                const R v0 = CV::ld(b2, sc.s[2]), w0 = CV::ld(b3, sc.s[3]);
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
            if (PML_FUSED && (k < pz_lo || k >= pz_hi))
                pml_tile<R, STAGE == 2>(*pmlp, PL, psi_c, k < zc0 || k >= zc1, i, j, k, Nx, Ny, po, uo, vo, wo,
                                        (-FxP_p + FxM_p) * K.ih,
                                         (-FyP_p + FyM_p) * K.ih, (-FzP_p + FzP) * K.ih, Du * K.ih, Dv * K.ih,
                                         Dw * K.ih, wgt * dt);
            ST(ou, pl, uo, sc.is[1]);
            ST(ov, pl, vo, sc.is[2]);
            ST(ow, pl, wo, sc.is[3]);
            ST(op, pl, po, sc.is[0]);
        }
        // advance the z pipeline
        pA = pAn; wA = wAn; FzP = FzP_p; FzW = FzP_w;
        if constexpr (!kRaw) {
            pn = CV::ld(pend_p, sc.s[0]);
            wn = CV::ld(pend_w, sc.s[3]);
        }
    };
    // six planes per trip (the carried windows rotate with periods 3 and 2, so the
    // compiler renames instead of moving); kend is uniform, so every wave leaves at the
    // same plane and the barriers stay matched
    for (int k = k0; k < kend; k += 6) {
        plane(k);
        if (k + 1 >= kend) break;
        plane(k + 1);
        if (k + 2 >= kend) break;
        plane(k + 2);
        if (k + 3 >= kend) break;
        plane(k + 3);
        if (k + 4 >= kend) break;
        plane(k + 4);
        if (k + 5 >= kend) break;
        plane(k + 5);
    }
}

// XCD-aware tile order. Workgroups of a launch go to the 8 XCDs round-robin (id mod 8),
// and each XCD has its own L2: tile t of the x-fastest tile order is given to workgroup
// w such that XCD x works through the contiguous tile range [x q + min(x, r), ...) of
// q + (x < r) tiles (q = n / 8, r = n mod 8). Neighbouring tiles then share an L2 and
// its copy of their common halo rows and planes.
constexpr unsigned kXcds = 8;
__device__ inline unsigned xcd_tile(unsigned w, unsigned n) {
    const unsigned q = n / kXcds, r = n % kXcds, x = w % kXcds, loc = w / kXcds;
    return x * q + min(x, r) + loc;
}

// The 8-row tile is held to 80 VGPRs (two spilled) so that three blocks share a CU and
// cover each other's barriers: 717 us per launch (mean of the three stages) against 770
// with two blocks per CU at 84-96 VGPRs. Stages 1 and 2 fit because their RK base set is
// loaded in its own plane, after the first barrier, rather than a plane ahead (ahead, the
// four values stay live across the whole plane: 19-27 spills at 80 VGPRs, 905 us).
template <typename S, typename R, int STAGE, int NW, int TX>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(NW == 8 ? 6 : 1, 8)))
k_flux_tiled(int Nx, int Ny, int Nz, TileFields<S> F, const uint8_t *__restrict__ upd, const PmlSet *__restrict__ pml,
             TileK<R> K, R dt,
             Scl<R> sc, int tz, int nbx, int nby) {
    __shared__ TileLds<R, NW, TX> L;
    const unsigned t = xcd_tile(blockIdx.x, gridDim.x);
    const int bx = (int)(t % (unsigned)nbx), by = (int)((t / (unsigned)nbx) % (unsigned)nby);
    const int bz = (int)(t / ((unsigned)nbx * (unsigned)nby));
    flux_tile<S, R, STAGE, NW, TX, false>(L, Nx, Ny, Nz, F, upd, pml, K, dt, sc, tz, bx, by, bz);
}

// cPML of one stage on the packed list of band cells, after k_flux_tiled (which leaves
// the band cells at their plain RK update): each cell's flux differences along the axes
// whose band holds it are formed again with the tile's own reconstructions and fluxes
// (weno_pair / weno_left_of / weno_right_of, godunov_f on the stage input: the same values
// bit for bit), then pml_apply adds the stretching terms to the stored stage output and,
// on the last stage, advances psi (functions3D.cuh:681-756). Cells of an FP64 source patch
// (update flag 2) are left to the patch kernel. Out of the flux kernel, the band code no
// longer holds registers in every tile of the room.
template <typename S, typename R, int STAGE>
__global__ void __launch_bounds__(256)
k_pml_band(int Nx, int Ny, TileFields<S> F, const int *__restrict__ cells, int n,
           const PmlSet *__restrict__ pmlp, TileK<R> K, R dt, Scl<R> sc) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const long long C = cells[t];
    typedef Conv<S, R> CV;
    const PmlSet &pm = *pmlp;
    const long long nxy = (long long)Nx * Ny;
    const int i = (int)(C % Nx), j = (int)((C / Nx) % Ny), k = (int)(C / nxy);
    // (-F_{+1/2} + F_{-1/2}) / h of p and of the axis velocity f along stride st
    auto diffs = [&](long long st, const S *f, R sf, R &Lp, R &Lf) {
        R p[7], q[7];
#pragma unroll
        for (int m = 0; m < 7; ++m) {
            p[m] = CV::ld(F.ip[C + (m - 3) * st], sc.s[0]);
            q[m] = CV::ld(f[C + (m - 3) * st], sf);
        }
        R Ap, Bp, Aq, Bq;
        const R Amp = weno_left_of(p[0], p[1], p[2], p[3], p[4]);      // cell - 1: left state
        const R Amq = weno_left_of(q[0], q[1], q[2], q[3], q[4]);
        weno_pair(p[1], p[2], p[3], p[4], p[5], Ap, Bp);               // the cell: both
        weno_pair(q[1], q[2], q[3], q[4], q[5], Aq, Bq);
        const R Bnp = weno_right_of(p[2], p[3], p[4], p[5], p[6]);     // cell + 1: right state
        const R Bnq = weno_right_of(q[2], q[3], q[4], q[5], q[6]);
        R FPp, FPq, FMp, FMq;
        godunov_f(Ap, Bnp, Aq, Bnq, K, FPp, FPq);                      // face + 1/2
        godunov_f(Amp, Bp, Amq, Bq, K, FMp, FMq);                      // face - 1/2
        Lp = (-FPp + FMp) * K.ih;
        Lf = (-FPq + FMq) * K.ih;
    };
    R Lpx = R(0), Lux = R(0), Lpy = R(0), Lvy = R(0), Lpz = R(0), Lwz = R(0);
    const bool bx = pm.x.band[i] >= 0, by = pm.y.band[j] >= 0, bz = pm.z.band[k] >= 0;
    if (bx) diffs(1, F.iu, sc.s[1], Lpx, Lux);
    if (by) diffs(Nx, F.iv, sc.s[2], Lpy, Lvy);
    if (bz) diffs(nxy, F.iw, sc.s[3], Lpz, Lwz);
    // the velocities of the cell's band axes only (pml_apply leaves the others alone)
    R po = CV::ld(F.op[C], sc.s[0]);
    R uo = bx ? CV::ld(F.ou[C], sc.s[1]) : R(0);
    R vo = by ? CV::ld(F.ov[C], sc.s[2]) : R(0);
    R wo = bz ? CV::ld(F.ow[C], sc.s[3]) : R(0);
    const R wgt = STAGE == 0 ? R(1) : (STAGE == 1 ? R(0.25) : R(2) / R(3));
    pml_apply<R, STAGE == 2>(pm, i, j, k, Nx, Ny, po, uo, vo, wo, Lpx, Lpy, Lpz, Lux, Lvy, Lwz, wgt * dt);
    F.op[C] = CV::st(po, sc.is[0]);
    if (bx) F.ou[C] = CV::st(uo, sc.is[1]);
    if (by) F.ov[C] = CV::st(vo, sc.is[2]);
    if (bz) F.ow[C] = CV::st(wo, sc.is[3]);
}

}  // namespace fasesdev
