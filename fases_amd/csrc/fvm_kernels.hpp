// Device kernels and parameter blocks of the FVM step shared by the single-context
// time loop (fvm_context.hip) and the batched multi-job loop (fvm_batch.hip).
#pragma once

#include "fvm_device.hpp"
#include "fvm_tiled.hpp"

#include <hip/hip_runtime.h>
#include <type_traits>

namespace fasesdev {

constexpr int NG = 3;   // ghost cells per side (parsing3D.h:128)

// ------------------------------------------------------------------------------------
// Kernel parameter blocks
// ------------------------------------------------------------------------------------

template <typename S> struct FieldSet {
    S *p, *u, *v, *w;
};

struct Grid {
    int Nx, Ny, Nz;
    long long nxy;
};

// ------------------------------------------------------------------------------------
// Flux stage: WENO5-Z + Godunov + SSP-RK3 stage combination + cPML correction.
// STAGE 0: out = U0 + dt L(U0)                       (functions3D.cuh:1149-1152)
// STAGE 1: out = 3/4 U0 + 1/4 (in + dt L(in))        (:1230-1233)
// STAGE 2: out = (U0 + 2 (in + dt L(in))) / 3        (:1311-1314)
// One thread per cell; the cell is updated when i,j,k in [3, N-3) and S >= h.
// ------------------------------------------------------------------------------------
template <typename S>
__device__ inline S *fld(const FieldSet<S> &s, int f) {
    return f == 0 ? s.p : (f == 1 ? s.u : (f == 2 ? s.v : s.w));
}

// Field access of the grid kernel: stored fields in, stored fields out.
template <typename S, typename R> struct GridIO {
    FieldSet<S> in, base, out;
    Scl<R> sc;
    __device__ R ld_in(int f, long long C) const { return Conv<S, R>::ld(fld(in, f)[C], sc.s[f]); }
    __device__ R ld_base(int f, long long C) const { return Conv<S, R>::ld(fld(base, f)[C], sc.s[f]); }
    __device__ void st(int f, long long C, R v) const { fld(out, f)[C] = Conv<S, R>::st(v, sc.is[f]); }
};

// Field access of the source-patch kernel: cells with a patch index read and write their
// FP64 shadow (and refresh the stored copy); other cells read the stored fields.
template <typename S> struct PatchIO {
    FieldSet<S> in, base, out;
    const int *pidx;
    const double *sh_in, *sh_base;
    double *sh_out;
    int np;
    Scl<double> sc;
    // entry f of a scale array by selection (an index that varies per lane would put the
    // array in scratch memory)
    __device__ static double pick(const double (&a)[4], int f) {
        return f == 0 ? a[0] : (f == 1 ? a[1] : (f == 2 ? a[2] : a[3]));
    }
    __device__ double ld_in(int f, long long C) const {
        const int q = pidx[C];
        return q >= 0 ? sh_in[(long long)f * np + q] : Conv<S, double>::ld(fld(in, f)[C], pick(sc.s, f));
    }
    __device__ double ld_base(int f, long long C) const {
        const int q = pidx[C];
        return q >= 0 ? sh_base[(long long)f * np + q] : Conv<S, double>::ld(fld(base, f)[C], pick(sc.s, f));
    }
    __device__ void st(int f, long long C, double v) const {
        sh_out[(long long)f * np + pidx[C]] = v;
        fld(out, f)[C] = Conv<S, double>::st(v, pick(sc.is, f));
    }
};

// The same access for the stage update of one patch cell, whose patch index q is known:
// flux_cell_update touches only its own cell C, whose shadow entry is q (pidx[C] == q), so
// the index lookups of PatchIO go away; the same values, the same bits.
template <typename S> struct PatchOwnIO {
    const PatchIO<S> &io;
    int q;
    __device__ double ld_in(int f, long long) const { return io.sh_in[(long long)f * io.np + q]; }
    __device__ double ld_base(int f, long long) const { return io.sh_base[(long long)f * io.np + q]; }
    __device__ void st(int f, long long C, double v) const {
        io.sh_out[(long long)f * io.np + q] = v;
        fld(io.out, f)[C] = Conv<S, double>::st(v, PatchIO<S>::pick(io.sc.is, f));
    }
};

// Godunov flux of p and of the axis velocity fu through one face of cell C along stride
// st: the face -1/2 (side 0, window C-3st .. C+2st) or +1/2 (side 1, C-2st .. C+3st), the
// reference's per-face reconstruction (functions3D.cuh:763-1001).
template <typename IO, typename R>
__device__ inline void face_flux(long long C, const IO &io, int fu, long long st, int side, R rho, R c, R &Fp, R &Fu) {
#pragma clang fp contract(off)
    R pw[6], uw[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        pw[q] = io.ld_in(0, C + (q - 3 + side) * st);
        uw[q] = io.ld_in(fu, C + (q - 3 + side) * st);
    }
    R pl, pr, ul, ur;
    weno_face(pw[0], pw[1], pw[2], pw[3], pw[4], pw[5], pl, pr);
    weno_face(uw[0], uw[1], uw[2], uw[3], uw[4], uw[5], ul, ur);
    godunov(pl, pr, ul, ur, rho, c, Fp, Fu);
}

// The four face fluxes of cell (i, j, k) along one axis: p and the axis velocity fu
// (1 = u, 2 = v, 3 = w) through the faces +1/2 and -1/2.
template <typename IO, typename R>
__device__ inline void flux_axis_cell(const Grid &g, int i, int j, int k, const IO &io, int fu, long long st, R rho,
                                      R c, R &Fp_p, R &Fp_m, R &Fu_p, R &Fu_m) {
    const long long C = (long long)i + (long long)g.Nx * j + g.nxy * k;
    face_flux(C, io, fu, st, 1, rho, c, Fp_p, Fu_p);
    face_flux(C, io, fu, st, 0, rho, c, Fp_m, Fu_m);
}

// The same face flux with the tiled kernel's reconstruction (product-form WENO5-Z weights,
// one reciprocal per state) in FP64: the left state is cell C-1's (face -1/2) or C's (face
// +1/2) A, the right state the next cell's B. The FP64 source patch of reduced-precision
// storage uses this form; it agrees with the reference order to rounding, and a face's
// flux depends only on its six cells, whichever cell's thread forms it.
template <typename IO>
__device__ inline void face_flux_pf(long long C, const IO &io, int fu, long long st, int side, double rho, double c,
                                    double &Fp, double &Fu) {
    double pw[6], uw[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        pw[q] = io.ld_in(0, C + (q - 3 + side) * st);
        uw[q] = io.ld_in(fu, C + (q - 3 + side) * st);
    }
    const double pl = weno_left_of(pw[0], pw[1], pw[2], pw[3], pw[4]);
    const double pr = weno_right_of(pw[1], pw[2], pw[3], pw[4], pw[5]);
    const double ul = weno_left_of(uw[0], uw[1], uw[2], uw[3], uw[4]);
    const double ur = weno_right_of(uw[1], uw[2], uw[3], uw[4], uw[5]);
    godunov(pl, pr, ul, ur, rho, c, Fp, Fu);
}

// Stage update of cell (i, j, k) from its twelve face fluxes (x: F, y: G, z: H).
template <typename IO, typename R, int STAGE>
__device__ inline void flux_cell_update(const Grid &g, int i, int j, int k, const IO &io, const PmlSet &pml, R h,
                                        R dt, R FpP, R FpM, R FuP, R FuM, R GpP, R GpM, R GvP, R GvM, R HpP, R HpM,
                                        R HwP, R HwM, bool src = false, R src_add = R(0)) {
#pragma clang fp contract(off)
    const long long C = (long long)i + (long long)g.Nx * j + g.nxy * k;
    const R p_in = io.ld_in(0, C), u_in = io.ld_in(1, C), v_in = io.ld_in(2, C), w_in = io.ld_in(3, C);
    const R sp = -FpP + FpM - GpP + GpM - HpP + HpM;
    R uo, vo, wo, po, wgt;
    if (STAGE == 0) {
        uo = u_in + dt * (-FuP + FuM) / h;
        vo = v_in + dt * (-GvP + GvM) / h;
        wo = w_in + dt * (-HwP + HwM) / h;
        po = p_in + dt * sp / h;
        wgt = R(1);
    } else {
        const R p0 = io.ld_base(0, C), u0 = io.ld_base(1, C), v0 = io.ld_base(2, C), w0 = io.ld_base(3, C);
        if (STAGE == 1) {
            uo = R(0.75) * u0 + R(0.25) * (u_in + dt * ((-FuP + FuM) / h));
            vo = R(0.75) * v0 + R(0.25) * (v_in + dt * ((-GvP + GvM) / h));
            wo = R(0.75) * w0 + R(0.25) * (w_in + dt * ((-HwP + HwM) / h));
            po = R(0.75) * p0 + R(0.25) * (p_in + dt * (sp / h));
            wgt = R(0.25);
        } else {
            uo = (u0 + R(2) * (u_in + dt * ((-FuP + FuM) / h))) / R(3);
            vo = (v0 + R(2) * (v_in + dt * ((-GvP + GvM) / h))) / R(3);
            wo = (w0 + R(2) * (w_in + dt * ((-HwP + HwM) / h))) / R(3);
            po = (p0 + R(2) * (p_in + dt * (sp / h))) / R(3);
            wgt = R(2) / R(3);
        }
    }
    if (pml.on)
        pml_apply<R, STAGE == 2>(pml, i, j, k, g.Nx, g.Ny, po, uo, vo, wo, (-FpP + FpM) / h, (-GpP + GpM) / h,
                                 (-HpP + HpM) / h, (-FuP + FuM) / h, (-GvP + GvM) / h, (-HwP + HwM) / h, wgt * dt);
    if (src) po = po + src_add;         // a point source's injection (k_sources' arithmetic)
    io.st(1, C, uo);
    io.st(2, C, vo);
    io.st(3, C, wo);
    io.st(0, C, po);
}


template <typename IO, typename R, int STAGE>
__device__ inline void flux_cell(const Grid &g, int i, int j, int k, const IO &io, const PmlSet &pml, R rho, R c,
                                 R h, R dt) {
    R FpP, FpM, FuP, FuM, GpP, GpM, GvP, GvM, HpP, HpM, HwP, HwM;
    flux_axis_cell(g, i, j, k, io, 1, 1, rho, c, FpP, FpM, FuP, FuM);
    flux_axis_cell(g, i, j, k, io, 2, g.Nx, rho, c, GpP, GpM, GvP, GvM);
    flux_axis_cell(g, i, j, k, io, 3, g.nxy, rho, c, HpP, HpM, HwP, HwM);
    flux_cell_update<IO, R, STAGE>(g, i, j, k, io, pml, h, dt, FpP, FpM, FuP, FuM, GpP, GpM, GvP, GvM, HpP, HpM, HwP,
                                   HwM);
}

// Update-mask values: 0 = not updated (outside [3,N-3)^3 or S < h), 1 = updated by the
// grid kernel, 2 = updated by the patch kernel (reduced-precision storage only: cells
// near point sources, kept in FP64 shadows, see fases_set_sources).
template <typename S, typename R, int STAGE>
__global__ void __launch_bounds__(256)
k_flux_simple(Grid g, GridIO<S, R> io, const uint8_t *__restrict__ upd, PmlSet pml, R rho, R c, R h, R dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + NG;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + NG;
    const int k = blockIdx.z + NG;
    if (i >= g.Nx - NG || j >= g.Ny - NG || k >= g.Nz - NG) return;
    const long long C = (long long)i + (long long)g.Nx * j + g.nxy * k;
    if (upd[C] != 1) return;
    flux_cell<GridIO<S, R>, R, STAGE>(g, i, j, k, io, pml, rho, c, h, dt);
}

// The same stage update in FP64 arithmetic and FP64 shadow storage on the patch cells
// (face_flux_pf's reconstruction): all six face fluxes of one cell (the batched loop, one
// thread per cell) ...
template <typename S, int STAGE>
__device__ inline void patch_cell(const Grid &g, int i, int j, int k, const PatchIO<S> &io, const PmlSet &pml,
                                  double rho, double c, double h, double dt) {
    const long long C = (long long)i + (long long)g.Nx * j + g.nxy * k;
    double F[3][4];                             // per axis: Fp +1/2, Fp -1/2, Fu +1/2, Fu -1/2
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const long long st = a == 0 ? 1LL : (a == 1 ? (long long)g.Nx : g.nxy);
        face_flux_pf(C, io, a + 1, st, 1, rho, c, F[a][0], F[a][2]);
        face_flux_pf(C, io, a + 1, st, 0, rho, c, F[a][1], F[a][3]);
    }
    flux_cell_update<PatchIO<S>, double, STAGE>(g, i, j, k, io, pml, h, dt, F[0][0], F[0][1], F[0][2], F[0][3],
                                                F[1][0], F[1][1], F[1][2], F[1][3], F[2][0], F[2][1], F[2][2],
                                                F[2][3]);
}
// ... or in two launches with every face flux formed once: one thread per (cell, axis)
// forms the flux through the cell's face -1/2 into fm[(a np + q) 2 ..], plus one thread per
// listed edge item e (edges[e] = a np + q: cell q, axis a, whose +1 neighbour is no patch
// cell) the flux through its face +1/2 into fe[2 e ..]; then one thread per cell applies
// the stage update, taking the face +1/2 from its neighbour's fm (the same window, so the
// same bits) or from fe (eid[a np + q] = e).
template <typename S>
__global__ void __launch_bounds__(64)
k_flux_patch_faces(Grid g, PatchIO<S> io, const int *__restrict__ cells, const int *__restrict__ edges, int n_edge,
                   double rho, double c, double *__restrict__ fm, double *__restrict__ fe) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n3 = 3 * io.np;
    if (t >= n3 + n_edge) return;
    int a, q, side;
    double *o;
    const int u = t < n3 ? t : edges[t - n3];  // a np + q
    a = u / io.np; q = u - a * io.np;
    if (t < n3) {
        side = 0; o = fm + 2LL * t;
    } else {
        side = 1; o = fe + 2LL * (t - n3);
    }
    const long long st = a == 0 ? 1LL : (a == 1 ? (long long)g.Nx : g.nxy);
    face_flux_pf((long long)cells[q], io, a + 1, st, side, rho, c, o[0], o[1]);
}
// Point sources inside the patch, injected by the patch update instead of k_sources (the
// same arithmetic, applied to the same FP64 value, so the same bits): source index of each
// patch cell (psrc, -1: none; null: k_sources injects every source) and k_sources'
// parameters of the stage.
struct PatchSrc {
    const int *psrc;
    const double *sval, *u_s;
    long long us_stride, nn;
    double w, dt, c, rho, As, h;
};

template <typename S, int STAGE>
__global__ void __launch_bounds__(64)
k_flux_patch_update(Grid g, PatchIO<S> io, const int *__restrict__ cells, PmlSet pml, double h, double dt,
                    const double *__restrict__ fm, const double *__restrict__ fe, const int *__restrict__ eid,
                    PatchSrc ps) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= io.np) return;
    const int sidx = ps.psrc ? ps.psrc[t] : -1;
    double add = 0.0;
    if (sidx >= 0) {
        const double amp = ps.u_s[(long long)sidx * ps.us_stride + ps.nn];
        add = ps.w * ps.dt * ps.c * ps.c * ps.sval[sidx] * amp * ps.rho * ps.As / (ps.h * ps.h * ps.h);
    }
    const int C = cells[t];
    const int i = C % g.Nx, j = (C / g.Nx) % g.Ny, k = (int)(C / g.nxy);
    double F[3][4];                             // per axis: Fp +1/2, Fp -1/2, Fu +1/2, Fu -1/2
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const long long st = a == 0 ? 1LL : (a == 1 ? (long long)g.Nx : g.nxy);
        const double *m = fm + 2LL * ((long long)a * io.np + t);
        const int nb = io.pidx[C + st];
        const double *pp = nb >= 0 ? fm + 2LL * ((long long)a * io.np + nb) : fe + 2LL * eid[(long long)a * io.np + t];
        F[a][0] = pp[0]; F[a][1] = m[0]; F[a][2] = pp[1]; F[a][3] = m[1];
    }
    const PatchOwnIO<S> own{io, t};
    flux_cell_update<PatchOwnIO<S>, double, STAGE>(g, i, j, k, own, pml, h, dt, F[0][0], F[0][1], F[0][2], F[0][3],
                                                   F[1][0], F[1][1], F[1][2], F[1][3], F[2][0], F[2][1], F[2][2],
                                                   F[2][3], sidx >= 0, add);
}

// Copies the stored fields of the patch cells into an FP64 shadow set.
template <typename S>
__global__ void k_patch_gather(FieldSet<S> f, const int *__restrict__ cells, int np, double *__restrict__ sh,
                               Scl<double> sc) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    const int C = cells[t];
    for (int q = 0; q < 4; ++q) sh[(long long)q * np + t] = Conv<S, double>::ld(fld(f, q)[C], sc.s[q]);
}

// Largest |stored value| of each field of one set and a non-finite flag:
// out[0..3] = float bit patterns (non-negative floats order like unsigned ints), out[4] != 0
// when a value is Inf or NaN. out[] must be zero on entry.
// 16-byte groups of stored values: the reduction and the rescale move whole groups per
// load and store when all four fields start on a 16-byte boundary (hipMalloc'd fields
// do); the cells past the last whole group, or every cell otherwise, go one at a time.
template <typename S> struct alignas(16) Pack16 {
    S v[16 / sizeof(S)];
};
template <typename S> __device__ inline long long pack16_groups(const FieldSet<S> &f, long long n) {
    const unsigned long long a = (unsigned long long)f.p | (unsigned long long)f.u | (unsigned long long)f.v |
                                 (unsigned long long)f.w;
    return (a & 15ULL) ? 0LL : n / (long long)(16 / sizeof(S));
}

template <typename S>
__global__ void __launch_bounds__(256) k_absmax(FieldSet<S> f, long long n, unsigned *__restrict__ out) {
    constexpr int V = 16 / (int)sizeof(S);
    float m[4] = {0.f, 0.f, 0.f, 0.f};
    unsigned bad = 0;
    auto take = [&](int q, S v) {
        const float x = fabsf((float)v);
        if (!(x <= 3.0e38f)) bad = 1;
        else m[q] = fmaxf(m[q], x);
    };
    const long long t0 = blockIdx.x * (long long)blockDim.x + threadIdx.x, stride = (long long)gridDim.x * blockDim.x;
    const long long ng = pack16_groups(f, n);
    for (long long c = t0; c < ng; c += stride) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const Pack16<S> w = reinterpret_cast<const Pack16<S> *>(fld(f, q))[c];
#pragma unroll
            for (int e = 0; e < V; ++e) take(q, w.v[e]);
        }
    }
    for (long long c = ng * V + t0; c < n; c += stride) {
#pragma unroll
        for (int q = 0; q < 4; ++q) take(q, fld(f, q)[c]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
        for (int o = 32; o > 0; o >>= 1) m[q] = fmaxf(m[q], __shfl_xor(m[q], o, 64));
    // the block's waves through LDS, then one atomic per field and block (atomics from
    // every wave on the same five words serialise: 0.8 ms on a 53 M-cell set)
    __shared__ float wm[4][4];
    __shared__ unsigned wb[4];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) wm[wv][q] = m[q];
    }
    unsigned any_bad = __any(bad) ? 1u : 0u;
    if ((threadIdx.x & 63) == 0) wb[wv] = any_bad;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (int)(blockDim.x + 63) >> 6;
        unsigned b = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float x = wm[0][q];
            for (int w = 1; w < nw; ++w) x = fmaxf(x, wm[w][q]);
            atomicMax(out + q, __float_as_uint(x));
        }
        for (int w = 0; w < nw; ++w) b |= wb[w];
        if (b) atomicOr(out + 4, 1u);
    }
}

template <typename T> __global__ void k_fill(T *__restrict__ p, int n, T v) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// Multiplies the stored values of each field by a power of two (exact while in range).
template <typename S>
__global__ void __launch_bounds__(256) k_rescale(FieldSet<S> f, long long n, Scl<float> fac) {
    constexpr int V = 16 / (int)sizeof(S);
    const long long t0 = blockIdx.x * (long long)blockDim.x + threadIdx.x, stride = (long long)gridDim.x * blockDim.x;
    const long long ng = pack16_groups(f, n);
    for (long long c = t0; c < ng; c += stride) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            Pack16<S> *a = reinterpret_cast<Pack16<S> *>(fld(f, q));
            Pack16<S> w = a[c];
#pragma unroll
            for (int e = 0; e < V; ++e) w.v[e] = (S)((float)w.v[e] * fac.s[q]);
            a[c] = w;
        }
    }
    for (long long c = ng * V + t0; c < n; c += stride) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            S *a = fld(f, q);
            a[c] = (S)((float)a[c] * fac.s[q]);
        }
    }
}

// ------------------------------------------------------------------------------------
// Point sources (functions3D.cuh:1039-1047): p[cell] += w dt c^2 sval amp rho A_s / h^3
// with amp = u_s[nn] (PCS / Ricker) or sin(2 pi f nn dt/2) (sine). Sources sharing a cell
// are merged on the host, so the update is race-free.
// ------------------------------------------------------------------------------------
template <typename S, typename R>
__global__ void k_sources(S *__restrict__ p, const int *__restrict__ cell, const double *__restrict__ sval,
                          const double *__restrict__ freq, const double *__restrict__ u_s, long long us_stride,
                          int n_src, int opt, long long nn, double w, double dt, double c, double rho, double As, double h,
                          const int *__restrict__ pidx, double *__restrict__ sh_p, R scale, R inv_scale) {
#pragma clang fp contract(off)
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_src) return;
    typedef Conv<S, R> CV;
    (void)freq;
    const double amp = u_s[(long long)s * us_stride + nn];   // shared waveform (stride 0) or per-source table
    const double add = w * dt * c * c * sval[s] * amp * rho * As / (h * h * h);
    const int C = cell[s];
    const int q = pidx ? pidx[C] : -1;
    if (q >= 0) {                        // FP64 shadow of a reduced-precision field
        const double v = sh_p[q] + add;
        sh_p[q] = v;
        p[C] = Conv<S, double>::st(v, (double)inv_scale);
    } else {
        const R cur = CV::ld(p[C], scale);
        p[C] = CV::st((R)((double)cur + add), inv_scale);
    }
}

// ------------------------------------------------------------------------------------
// Immersed boundary + ADE on the packed boundary list (functions3D.cuh:462-565,
// ADE_solver :426-459). Probe values come from the stage INPUT fields; the blend
// q = (1-chi) q + chi q_b acts on the stage OUTPUT. All arithmetic is FP64.
// ADE stage (a,b,c): x_left = (a x_mid + b (x_right + dt xdot(x_right))) / c.
// ------------------------------------------------------------------------------------
// Probe geometry of boundary cell (i, j, k) from its signed distance s (the SDF's float
// value) and unit normal n (functions3D.cuh:466-479, filterABC.h:95-108): foot point
// W = C + s n, probe D = W - (s + 2h) n, the trilinear base corner of D clamped to the
// grid, the fractions of D inside that corner's cell, and the blend weight chi. The host
// builds the boundary list with this function and the boundary kernels recompute it per
// cell from (s, n): both sides round alike (no contraction, IEEE division and floor), so
// the kernels see the host's bits while the list holds 36 bytes per cell instead of 68.
struct ProbeGeo {
    int base;               // grid index of the trilinear base corner
    double sx, sy, sz;      // fractions of the probe point inside the base cell
    double chi;             // blend weight
};
__host__ __device__ inline ProbeGeo probe_geo(int i, int j, int k, int Nx, int Ny, int Nz, double s, double nx,
                                              double ny, double nz, double h, double tol) {
#pragma clang fp contract(off)
    const double Cx = (i + 0.5) * h, Cy = (j + 0.5) * h, Cz = (k + 0.5) * h;
    const double Wx = Cx + s * nx, Wy = Cy + s * ny, Wz = Cz + s * nz;
    const double l = s + 2.0 * h;
    const double Dx = Wx - l * nx, Dy = Wy - l * ny, Dz = Wz - l * nz;
    const double qx = Dx / h - 0.5, qy = Dy / h - 0.5, qz = Dz / h - 0.5;
    int bi = (int)floor(qx), bj = (int)floor(qy), bk = (int)floor(qz);
    bi = bi < 0 ? 0 : (bi > Nx - 2 ? Nx - 2 : bi);
    bj = bj < 0 ? 0 : (bj > Ny - 2 ? Ny - 2 : bj);
    bk = bk < 0 ? 0 : (bk > Nz - 2 ? Nz - 2 : bk);
    ProbeGeo r;
    r.base = bi + Nx * (bj + Ny * bk);
    r.sx = qx - (double)bi;
    r.sy = qy - (double)bj;
    r.sz = qz - (double)bk;
    r.chi = chi_band(s, h, tol);
    return r;
}

struct Bnd {
    int n;
    const int *cell;        // grid index of the boundary cell
    const int *mat;         // bits 0-7: material index (oid - 1); bit 8: probe on a cell centre
    const float *dist;      // signed distance S of the cell
    const double *nrm;      // [3][n] (entry-major): unit normal nx, ny, nz
    double h, tol;          // cell size; SDF distance where the blend weight reaches 0
};

// What the boundary kernels use of boundary cell b: its grid cell, material word, normal
// and probe geometry (consecutive b are consecutive addresses in every array).
struct BndCell {
    int C, mw;
    double nx, ny, nz;
    ProbeGeo q;
};
__device__ inline BndCell bnd_cell(const Grid &g, const Bnd &bd, int b) {
    BndCell r;
    r.C = bd.cell[b];
    r.mw = bd.mat[b];
    r.nx = bd.nrm[b];
    r.ny = bd.nrm[(long long)bd.n + b];
    r.nz = bd.nrm[2LL * bd.n + b];
    const unsigned C = (unsigned)r.C, nxy = (unsigned)g.nxy, nx = (unsigned)g.Nx;
    const unsigned k = C / nxy, rem = C - k * nxy, j = rem / nx, i = rem - j * nx;
    r.q = probe_geo((int)i, (int)j, (int)k, g.Nx, g.Ny, g.Nz, (double)bd.dist[b], r.nx, r.ny, r.nz, bd.h, bd.tol);
    return r;
}

// ADE state k of boundary cell b in an [K][nB] state buffer (pole-major, so the lanes of a
// wave, consecutive boundary cells, touch consecutive addresses).
__device__ inline long long ade_at(int k, int b, int nB) { return (long long)k * nB + b; }

struct Ade {
    int K, L;
    const double *Rinf, *A, *zeta, *B, *C, *alpha, *beta;
    const int *Keff, *Leff;   // poles that can contribute per material
};

template <typename S>
__device__ inline double tri_probe(const S *__restrict__ f, int base, int sx_, double sx, double sy, double sz,
                                   int nx, long long nxy, double scale) {
#pragma clang fp contract(off)
    typedef Conv<S, typename std::conditional<std::is_same<S, double>::value, double, float>::type> CV;
    (void)sx_;
    const double f000 = (double)CV::ld(f[base], scale);
    const double f100 = (double)CV::ld(f[base + 1], scale);
    const double f010 = (double)CV::ld(f[base + nx], scale);
    const double f110 = (double)CV::ld(f[base + nx + 1], scale);
    const double f001 = (double)CV::ld(f[base + nxy], scale);
    const double f101 = (double)CV::ld(f[base + nxy + 1], scale);
    const double f011 = (double)CV::ld(f[base + nxy + nx], scale);
    const double f111 = (double)CV::ld(f[base + nxy + nx + 1], scale);
    const double c00 = f000 * (1 - sx) + f100 * sx;
    const double c10 = f010 * (1 - sx) + f110 * sx;
    const double c01 = f001 * (1 - sx) + f101 * sx;
    const double c11 = f011 * (1 - sx) + f111 * sx;
    const double c0 = c00 * (1 - sy) + c10 * sy, c1 = c01 * (1 - sy) + c11 * sy;
    return c0 * (1 - sz) + c1 * sz;
}

// ADE state buffers of one stage: x_left = (a x_mid + b (x_right + dt xdot(x_right))) / c.
struct AdeBufs {
    double *phi_l;
    const double *phi_m, *phi_r;
    double *p1_l;
    const double *p1_m, *p1_r;
    double *p2_l;
    const double *p2_m, *p2_r;
};

template <typename S>
__device__ inline void boundary_cell(int b, const Grid &g, const FieldSet<S> &in, const FieldSet<S> &out,
                                     const Bnd &bd, const Ade &ade, double rho, double c, double dt, double as,
                                     double bs, double cs, const AdeBufs &ab, const Scl<double> &sc) {
#pragma clang fp contract(off)
    double *__restrict__ phi_l = ab.phi_l, *__restrict__ p1_l = ab.p1_l, *__restrict__ p2_l = ab.p2_l;
    const double *__restrict__ phi_m = ab.phi_m, *__restrict__ phi_r = ab.phi_r;
    const double *__restrict__ p1_m = ab.p1_m, *__restrict__ p1_r = ab.p1_r;
    const double *__restrict__ p2_m = ab.p2_m, *__restrict__ p2_r = ab.p2_r;
    typedef typename std::conditional<std::is_same<S, double>::value, double, float>::type R;
    typedef Conv<S, R> CV;
    const BndCell bc = bnd_cell(g, bd, b);
    const int C = bc.C, base = bc.q.base, mw = bc.mw;
    const int m = mw & 0xFF;
    const double nx = bc.nx, ny = bc.ny, nz = bc.nz, chi = bc.q.chi;
    double pD, uD, vD, wD;
    if (mw & 0x100) {
        // probe exactly on a cell centre (all trilinear fractions 0): the interpolant is
        // that cell's value, bit for bit
        pD = (double)CV::ld(in.p[base], (R)sc.s[0]);
        uD = (double)CV::ld(in.u[base], (R)sc.s[1]);
        vD = (double)CV::ld(in.v[base], (R)sc.s[2]);
        wD = (double)CV::ld(in.w[base], (R)sc.s[3]);
    } else {
        const double sx = bc.q.sx, sy = bc.q.sy, sz = bc.q.sz;
        pD = tri_probe(in.p, base, 0, sx, sy, sz, g.Nx, g.nxy, sc.s[0]);
        uD = tri_probe(in.u, base, 0, sx, sy, sz, g.Nx, g.nxy, sc.s[1]);
        vD = tri_probe(in.v, base, 0, sx, sy, sz, g.Nx, g.nxy, sc.s[2]);
        wD = tri_probe(in.w, base, 0, sx, sy, sz, g.Nx, g.nxy, sc.s[3]);
    }
    const double un = uD * nx + vD * ny + wD * nz;
    const double utx = uD - un * nx, uty = vD - un * ny, utz = wD - un * nz;
    const double Z0 = rho * c;
    const double wout = pD + Z0 * un;
    // ADE (vector-fitted admittance): real poles phi, complex pairs (psi1, psi2). A stage
    // with a = 0 (the first) does not read the mid buffer: its term is a zero either way.
    double sumA = 0.0, sumBC = 0.0;
    const int Ke = ade.Keff[m], Le = ade.Leff[m];
    const bool mid = as != 0.0;
    for (int k = 0; k < Ke; ++k) {
        const int mi = m * ade.K + k;
        const long long s = ade_at(k, b, bd.n);
        const double xr = phi_r[s];
        const double Lphi = wout - ade.zeta[mi] * xr;
        double xl = as * (mid ? phi_m[s] : 0.0) + bs * (xr + dt * Lphi);
        xl /= cs;
        phi_l[s] = xl;
        sumA += ade.A[mi] * xl;
    }
    for (int l = 0; l < Le; ++l) {
        const int mi = m * ade.L + l;
        const long long s = ade_at(l, b, bd.n);
        const double y1 = p1_r[s], y2 = p2_r[s];
        const double L1 = wout - (ade.alpha[mi] * y1 + ade.beta[mi] * y2);
        const double L2 = ade.beta[mi] * y1 - ade.alpha[mi] * y2;
        double z1 = as * (mid ? p1_m[s] : 0.0) + bs * (y1 + dt * L1);
        z1 /= cs;
        double z2 = as * (mid ? p2_m[s] : 0.0) + bs * (y2 + dt * L2);
        z2 /= cs;
        p1_l[s] = z1;
        p2_l[s] = z2;
        sumBC += ade.B[mi] * z1 + ade.C[mi] * z2;
    }
    const double win = ade.Rinf[m] * wout + sumA + sumBC;
    const double pb = 0.5 * (win + wout);
    const double unb = (wout - win) / (2.0 * Z0);
    const double ub = unb * nx + utx, vb = unb * ny + uty, wb = unb * nz + utz;
    const double u_o = (double)CV::ld(out.u[C], (R)sc.s[1]);
    const double v_o = (double)CV::ld(out.v[C], (R)sc.s[2]);
    const double w_o = (double)CV::ld(out.w[C], (R)sc.s[3]);
    const double p_o = (double)CV::ld(out.p[C], (R)sc.s[0]);
    out.u[C] = CV::st((R)((1.0 - chi) * u_o + chi * ub), (R)sc.is[1]);
    out.v[C] = CV::st((R)((1.0 - chi) * v_o + chi * vb), (R)sc.is[2]);
    out.w[C] = CV::st((R)((1.0 - chi) * w_o + chi * wb), (R)sc.is[3]);
    out.p[C] = CV::st((R)((1.0 - chi) * p_o + chi * pb), (R)sc.is[0]);
}

template <typename S>
__global__ void __launch_bounds__(256)
k_boundary(Grid g, FieldSet<S> in, FieldSet<S> out, Bnd bd, Ade ade, double rho, double c, double dt,
           double as, double bs, double cs, AdeBufs ab, Scl<double> sc) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= bd.n) return;
    boundary_cell<S>(b, g, in, out, bd, ade, rho, c, dt, as, bs, cs, ab, sc);
}

// ------------------------------------------------------------------------------------
// The same ADE recursion with the three SSP-RK3 stages of a step collapsed onto the
// step's start state x0 (reduced-precision storage; FP64 storage keeps boundary_cell,
// which is the reference's operation order). Per real pole, with a = 1 - dt zeta and
// w_s the stage's characteristic wout:
//   x1 = a x0 + dt w0
//   x2 = (3 x0 + a x1 + dt w1) / 4 = (3 + a^2)/4 x0 + a dt/4 w0 + dt/4 w1
//   x3 = (x0 + 2 (a x2 + dt w2)) / 3 = (1/3 + a (3 + a^2)/6) x0 + a^2 dt/6 w0 + a dt/6 w1 + 2 dt/3 w2
// and per complex pair Y = (psi1, psi2) the same with a -> M = [[1 - dt alpha, -dt beta],
// [dt beta, 1 - dt alpha]] acting on Y and dt w -> dt w e1. The stage's admittance sum
// (sum_k A_k x_{s+1,k} + sum_l B_l psi1 + C_l psi2) is then a dot product with x0 plus
// material constants times w0 .. ws, so the intermediate states x1, x2 never exist:
//   stage 0: reads x0, keeps T1 = (stage-1 x0 part) + G10 w0 and w0;
//   stage 1: reads T1, keeps w1;
//   stage 2: reads x0, w0, w1, writes x3.
// Per boundary cell and step that is 30 FP64 values moved instead of 64 (K = 4, L = 2).
// Every constant is formed on the host in FP64 (fases_set_materials).
// ------------------------------------------------------------------------------------
// Per-material coefficients, material m at coef + m * nc, nc = 6K + 14L + 4:
//   real pole k at 6k:        A, A a, A (3 + a^2)/4, 1/3 + a (3 + a^2)/6, a^2 dt/6, a dt/6
//   pair l at 6K + 14l:       B, C, [B C] M (2), [B C] (3I + M^2)/4 (2),
//                             P3 = I/3 + M (3I + M^2)/6 (4, row-major), dt/6 M^2 e1 (2), dt/6 M e1 (2)
//   tail at 6K + 14L:         Rinf, Rinf + G00, G10, Rinf + G11
enum AdeRkTail { kRkRinf = 0, kRkG00, kRkG10, kRkG11 };
struct AdeRk {
    int K, L, nc;
    const double *coef;
    const int *Keff, *Leff;
    double two_dt3;           // 2 dt / 3
};
// x0: the step's start state (read), x3: the step's end state (written at stage 2),
// aux: [3][nB] = T1, w0, w1 (entry-major).
struct AdeRkBufs {
    const double *x0_phi, *x0_p1, *x0_p2;
    double *x3_phi, *x3_p1, *x3_p2;
    double *aux;
};

template <typename S, int STAGE>
__device__ inline void boundary_cell_rk(int b, const Grid &g, const FieldSet<S> &in, const FieldSet<S> &out,
                                        const Bnd &bd, const AdeRk &ade, double Z0, double inv_2z0,
                                        const AdeRkBufs &ab, const Scl<double> &sc) {
    // No contraction beyond the explicit fma calls: whether the compiler fuses a product
    // into a following add depends on the basic blocks of the calling kernel, so the
    // single-context and batched kernels would otherwise round differently.
#pragma clang fp contract(off)
    typedef typename std::conditional<std::is_same<S, double>::value, double, float>::type R;
    typedef Conv<S, R> CV;
    const BndCell bc = bnd_cell(g, bd, b);
    const int C = bc.C, base = bc.q.base, mw = bc.mw;
    const int m = mw & 0xFF;
    const double nx = bc.nx, ny = bc.ny, nz = bc.nz, chi = bc.q.chi;
    double pD, uD, vD, wD;
    if (mw & 0x100) {
        pD = (double)CV::ld(in.p[base], (R)sc.s[0]);
        uD = (double)CV::ld(in.u[base], (R)sc.s[1]);
        vD = (double)CV::ld(in.v[base], (R)sc.s[2]);
        wD = (double)CV::ld(in.w[base], (R)sc.s[3]);
    } else {
        const double sx = bc.q.sx, sy = bc.q.sy, sz = bc.q.sz;
        pD = tri_probe(in.p, base, 0, sx, sy, sz, g.Nx, g.nxy, sc.s[0]);
        uD = tri_probe(in.u, base, 0, sx, sy, sz, g.Nx, g.nxy, sc.s[1]);
        vD = tri_probe(in.v, base, 0, sx, sy, sz, g.Nx, g.nxy, sc.s[2]);
        wD = tri_probe(in.w, base, 0, sx, sy, sz, g.Nx, g.nxy, sc.s[3]);
    }
    const double un = uD * nx + vD * ny + wD * nz;
    const double utx = uD - un * nx, uty = vD - un * ny, utz = wD - un * nz;
    const double wout = pD + Z0 * un;
    const double *__restrict__ cf = ade.coef + (long long)m * ade.nc;
    const double *__restrict__ tl = cf + 6 * ade.K + 14 * ade.L;
    const int Ke = ade.Keff[m], Le = ade.Leff[m];
    const int nB = bd.n;
    double win;
    if (STAGE == 0) {
        double s0 = 0.0, s1 = 0.0;
        for (int k = 0; k < Ke; ++k) {
            const double x = ab.x0_phi[ade_at(k, b, nB)];
            s0 = fma(cf[6 * k + 1], x, s0);
            s1 = fma(cf[6 * k + 2], x, s1);
        }
        for (int l = 0; l < Le; ++l) {
            const double *q = cf + 6 * ade.K + 14 * l;
            const long long s = ade_at(l, b, nB);
            const double y1 = ab.x0_p1[s], y2 = ab.x0_p2[s];
            s0 = fma(q[2], y1, fma(q[3], y2, s0));
            s1 = fma(q[4], y1, fma(q[5], y2, s1));
        }
        win = fma(tl[kRkG00], wout, s0);
        ab.aux[b] = fma(tl[kRkG10], wout, s1);
        ab.aux[(long long)nB + b] = wout;
    } else if (STAGE == 1) {
        win = fma(tl[kRkG11], wout, ab.aux[b]);
        ab.aux[2LL * nB + b] = wout;
    } else {
        const double w0 = ab.aux[(long long)nB + b], w1 = ab.aux[2LL * nB + b];
        const double w2 = ade.two_dt3 * wout;
        double sum = tl[kRkRinf] * wout;
        for (int k = 0; k < Ke; ++k) {
            const long long s = ade_at(k, b, nB);
            const double *q = cf + 6 * k;
            const double x3 = fma(q[3], ab.x0_phi[s], fma(q[4], w0, fma(q[5], w1, w2)));
            ab.x3_phi[s] = x3;
            sum = fma(q[0], x3, sum);
        }
        for (int l = 0; l < Le; ++l) {
            const double *q = cf + 6 * ade.K + 14 * l;
            const long long s = ade_at(l, b, nB);
            const double y1 = ab.x0_p1[s], y2 = ab.x0_p2[s];
            const double z1 = fma(q[6], y1, fma(q[7], y2, fma(q[10], w0, fma(q[12], w1, w2))));
            const double z2 = fma(q[8], y1, fma(q[9], y2, fma(q[11], w0, q[13] * w1)));
            ab.x3_p1[s] = z1;
            ab.x3_p2[s] = z2;
            sum = fma(q[0], z1, fma(q[1], z2, sum));
        }
        win = sum;
    }
    const double pb = 0.5 * (win + wout);
    const double unb = (wout - win) * inv_2z0;
    const double ub = unb * nx + utx, vb = unb * ny + uty, wb = unb * nz + utz;
    const double u_o = (double)CV::ld(out.u[C], (R)sc.s[1]);
    const double v_o = (double)CV::ld(out.v[C], (R)sc.s[2]);
    const double w_o = (double)CV::ld(out.w[C], (R)sc.s[3]);
    const double p_o = (double)CV::ld(out.p[C], (R)sc.s[0]);
    out.u[C] = CV::st((R)((1.0 - chi) * u_o + chi * ub), (R)sc.is[1]);
    out.v[C] = CV::st((R)((1.0 - chi) * v_o + chi * vb), (R)sc.is[2]);
    out.w[C] = CV::st((R)((1.0 - chi) * w_o + chi * wb), (R)sc.is[3]);
    out.p[C] = CV::st((R)((1.0 - chi) * p_o + chi * pb), (R)sc.is[0]);
}

template <typename S, int STAGE>
__global__ void __launch_bounds__(256)
k_boundary_rk(Grid g, FieldSet<S> in, FieldSet<S> out, Bnd bd, AdeRk ade, double Z0, double inv_2z0,
              AdeRkBufs ab, Scl<double> sc) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= bd.n) return;
    boundary_cell_rk<S, STAGE>(b, g, in, out, bd, ade, Z0, inv_2z0, ab, sc);
}

// Receiver gather after a step: ir[r*stride + n] = p(cell_r), vel[(r*stride + n)*3 + a].
template <typename S>
__global__ void k_receivers(FieldSet<S> f, const int *__restrict__ cell, int n_rec, double *__restrict__ ir,
                            double *__restrict__ vel, long long stride, long long n, Scl<double> sc) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rec) return;
    typedef typename std::conditional<std::is_same<S, double>::value, double, float>::type R;
    typedef Conv<S, R> CV;
    const int C = cell[r];
    ir[r * stride + n] = (double)CV::ld(f.p[C], (R)sc.s[0]);
    if (vel) {
        double *o = vel + 3 * (r * stride + n);
        o[0] = (double)CV::ld(f.u[C], (R)sc.s[1]);
        o[1] = (double)CV::ld(f.v[C], (R)sc.s[2]);
        o[2] = (double)CV::ld(f.w[C], (R)sc.s[3]);
    }
}

}  // namespace fasesdev
