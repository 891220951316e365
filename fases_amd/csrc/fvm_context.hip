// Solver context and time loop of the MI355X FVM core (C ABI in include/fases_core.h).
//
// One SSP-RK3 step = 3 x (flux stage kernel -> source kernel -> boundary kernel), as
// in 3D-solver/core/main3D.cu:1280-1320, with the reference's three field sets and
// its pointer rotation (U0 <-> U1 and phi_old <-> phi after every step). Fields are
// SoA p, vx, vy, vz over the padded grid, x fastest (functions3D.cuh:48-50).
// Differences from the reference by design: the boundary kernel walks a packed list
// of boundary cells instead of the whole grid, geometry constants of those cells are
// computed once at set-up, the ADE states are FP64 for every precision, receivers are
// gathered on the device into an FP64 buffer (no per-step host copies), and PML
// profiles are 1-D tables per axis.

#include "fvm_device.hpp"
#include "fvm_tiled.hpp"
#include "host.hpp"
#include "../../include/fases_core.h"

#include <hip/hip_runtime.h>
#include <cmath>
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

using namespace fasesdev;

namespace {

constexpr int NG = 3;   // ghost cells per side (parsing3D.h:128)
constexpr int kSourcePatch = 12;  // FP64 shadow radius around point sources (reduced storage)

#define HIPCHK(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);             \
            return FASES_E_HIP;                                                       \
        }                                                                             \
    } while (0)

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
    R s_in, s_base, is_out;
    __device__ R ld_in(int f, long long C) const { return Conv<S, R>::ld(fld(in, f)[C], s_in); }
    __device__ R ld_base(int f, long long C) const { return Conv<S, R>::ld(fld(base, f)[C], s_base); }
    __device__ void st(int f, long long C, R v) const { fld(out, f)[C] = Conv<S, R>::st(v, is_out); }
};

// Field access of the source-patch kernel: cells with a patch index read and write their
// FP64 shadow (and refresh the stored copy); other cells read the stored fields.
template <typename S> struct PatchIO {
    FieldSet<S> in, base, out;
    const int *pidx;
    const double *sh_in, *sh_base;
    double *sh_out;
    int np;
    double s_in, s_base, is_out;
    __device__ double ld_in(int f, long long C) const {
        const int q = pidx[C];
        return q >= 0 ? sh_in[(long long)f * np + q] : Conv<S, double>::ld(fld(in, f)[C], s_in);
    }
    __device__ double ld_base(int f, long long C) const {
        const int q = pidx[C];
        return q >= 0 ? sh_base[(long long)f * np + q] : Conv<S, double>::ld(fld(base, f)[C], s_base);
    }
    __device__ void st(int f, long long C, double v) const {
        sh_out[(long long)f * np + pidx[C]] = v;
        fld(out, f)[C] = Conv<S, double>::st(v, is_out);
    }
};

template <typename IO, typename R, int STAGE>
__device__ inline void flux_cell(const Grid &g, int i, int j, int k, const IO &io, const PmlSet &pml, R rho, R c,
                                 R h, R dt) {
#pragma clang fp contract(off)
    const long long C = (long long)i + (long long)g.Nx * j + g.nxy * k;

    // fluxes through the faces C+1/2 (P) and C-1/2 (M) along one axis
    auto flux_axis = [&](int fu, long long st, R &Fp_p, R &Fp_m, R &Fu_p, R &Fu_m) {
        R pw[7], uw[7];
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            pw[q] = io.ld_in(0, C + (q - 3) * st);
            uw[q] = io.ld_in(fu, C + (q - 3) * st);
        }
        R pl, pr, ul, ur;
        weno_face(pw[1], pw[2], pw[3], pw[4], pw[5], pw[6], pl, pr);   // face +1/2
        weno_face(uw[1], uw[2], uw[3], uw[4], uw[5], uw[6], ul, ur);
        godunov(pl, pr, ul, ur, rho, c, Fp_p, Fu_p);
        weno_face(pw[0], pw[1], pw[2], pw[3], pw[4], pw[5], pl, pr);   // face -1/2
        weno_face(uw[0], uw[1], uw[2], uw[3], uw[4], uw[5], ul, ur);
        godunov(pl, pr, ul, ur, rho, c, Fp_m, Fu_m);
    };

    R FpP, FpM, FuP, FuM, GpP, GpM, GvP, GvM, HpP, HpM, HwP, HwM;
    flux_axis(1, 1, FpP, FpM, FuP, FuM);
    flux_axis(2, g.Nx, GpP, GpM, GvP, GvM);
    flux_axis(3, g.nxy, HpP, HpM, HwP, HwM);

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
    io.st(1, C, uo);
    io.st(2, C, vo);
    io.st(3, C, wo);
    io.st(0, C, po);
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

// The same stage update in FP64 arithmetic and FP64 shadow storage on the patch cells.
template <typename S, int STAGE>
__global__ void __launch_bounds__(64)
k_flux_patch(Grid g, PatchIO<S> io, const int *__restrict__ cells, PmlSet pml, double rho, double c, double h,
             double dt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= io.np) return;
    const int C = cells[t];
    const int i = C % g.Nx, j = (C / g.Nx) % g.Ny, k = (int)(C / g.nxy);
    flux_cell<PatchIO<S>, double, STAGE>(g, i, j, k, io, pml, rho, c, h, dt);
}

// Copies the stored fields of the patch cells into an FP64 shadow set.
template <typename S>
__global__ void k_patch_gather(FieldSet<S> f, const int *__restrict__ cells, int np, double *__restrict__ sh,
                               double scale) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    const int C = cells[t];
    for (int q = 0; q < 4; ++q) sh[(long long)q * np + t] = Conv<S, double>::ld(fld(f, q)[C], scale);
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
template <typename G> struct Bnd {
    int n;
    const int *cell;        // grid index of the boundary cell
    const int *base;        // grid index of the trilinear base corner of the probe point
    const int *mat;         // bits 0-7: material index (oid - 1); bit 8: probe on a cell centre
    const G *geo;           // [8*n]: sx, sy, sz, nx, ny, nz, chi, pad (FP64 for FP64 fields)
};

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

template <typename S, typename G>
__global__ void __launch_bounds__(256)
k_boundary(Grid g, FieldSet<S> in, FieldSet<S> out, Bnd<G> bd, Ade ade, double rho, double c, double dt,
           double as, double bs, double cs, double *__restrict__ phi_l, const double *__restrict__ phi_m,
           const double *__restrict__ phi_r, double *__restrict__ p1_l, const double *__restrict__ p1_m,
           const double *__restrict__ p1_r, double *__restrict__ p2_l, const double *__restrict__ p2_m,
           const double *__restrict__ p2_r, double scale_in, double scale_out, double inv_scale_out) {
#pragma clang fp contract(off)
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= bd.n) return;
    typedef typename std::conditional<std::is_same<S, double>::value, double, float>::type R;
    typedef Conv<S, R> CV;
    const int C = bd.cell[b], base = bd.base[b], mw = bd.mat[b];
    const int m = mw & 0xFF;
    const G *q = bd.geo + 8 * (long long)b;
    const double sx = q[0], sy = q[1], sz = q[2], nx = q[3], ny = q[4], nz = q[5], chi = q[6];
    double pD, uD, vD, wD;
    if (mw & 0x100) {
        // probe exactly on a cell centre (all trilinear fractions 0): the interpolant is
        // that cell's value, bit for bit
        pD = (double)CV::ld(in.p[base], (R)scale_in);
        uD = (double)CV::ld(in.u[base], (R)scale_in);
        vD = (double)CV::ld(in.v[base], (R)scale_in);
        wD = (double)CV::ld(in.w[base], (R)scale_in);
    } else {
        pD = tri_probe(in.p, base, 0, sx, sy, sz, g.Nx, g.nxy, scale_in);
        uD = tri_probe(in.u, base, 0, sx, sy, sz, g.Nx, g.nxy, scale_in);
        vD = tri_probe(in.v, base, 0, sx, sy, sz, g.Nx, g.nxy, scale_in);
        wD = tri_probe(in.w, base, 0, sx, sy, sz, g.Nx, g.nxy, scale_in);
    }
    const double un = uD * nx + vD * ny + wD * nz;
    const double utx = uD - un * nx, uty = vD - un * ny, utz = wD - un * nz;
    const double Z0 = rho * c;
    const double wout = pD + Z0 * un;
    // ADE (vector-fitted admittance): real poles phi, complex pairs (psi1, psi2)
    double sumA = 0.0, sumBC = 0.0;
    const int Ke = ade.Keff[m], Le = ade.Leff[m];
    for (int k = 0; k < Ke; ++k) {
        const int mi = m * ade.K + k;
        const long long s = (long long)b * ade.K + k;
        const double xr = phi_r[s];
        const double Lphi = wout - ade.zeta[mi] * xr;
        double xl = as * phi_m[s] + bs * (xr + dt * Lphi);
        xl /= cs;
        phi_l[s] = xl;
        sumA += ade.A[mi] * xl;
    }
    for (int l = 0; l < Le; ++l) {
        const int mi = m * ade.L + l;
        const long long s = (long long)b * ade.L + l;
        const double y1 = p1_r[s], y2 = p2_r[s];
        const double L1 = wout - (ade.alpha[mi] * y1 + ade.beta[mi] * y2);
        const double L2 = ade.beta[mi] * y1 - ade.alpha[mi] * y2;
        double z1 = as * p1_m[s] + bs * (y1 + dt * L1);
        z1 /= cs;
        double z2 = as * p2_m[s] + bs * (y2 + dt * L2);
        z2 /= cs;
        p1_l[s] = z1;
        p2_l[s] = z2;
        sumBC += ade.B[mi] * z1 + ade.C[mi] * z2;
    }
    const double win = ade.Rinf[m] * wout + sumA + sumBC;
    const double pb = 0.5 * (win + wout);
    const double unb = (wout - win) / (2.0 * Z0);
    const double ub = unb * nx + utx, vb = unb * ny + uty, wb = unb * nz + utz;
    const double u_o = (double)CV::ld(out.u[C], scale_out);
    const double v_o = (double)CV::ld(out.v[C], scale_out);
    const double w_o = (double)CV::ld(out.w[C], scale_out);
    const double p_o = (double)CV::ld(out.p[C], scale_out);
    out.u[C] = CV::st((R)((1.0 - chi) * u_o + chi * ub), (R)inv_scale_out);
    out.v[C] = CV::st((R)((1.0 - chi) * v_o + chi * vb), (R)inv_scale_out);
    out.w[C] = CV::st((R)((1.0 - chi) * w_o + chi * wb), (R)inv_scale_out);
    out.p[C] = CV::st((R)((1.0 - chi) * p_o + chi * pb), (R)inv_scale_out);
}

// Receiver gather after a step: ir[r*stride + n] = p(cell_r), vel[(r*stride + n)*3 + a].
template <typename S>
__global__ void k_receivers(FieldSet<S> f, const int *__restrict__ cell, int n_rec, double *__restrict__ ir,
                            double *__restrict__ vel, long long stride, long long n, double scale) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rec) return;
    typedef typename std::conditional<std::is_same<S, double>::value, double, float>::type R;
    typedef Conv<S, R> CV;
    const int C = cell[r];
    ir[r * stride + n] = (double)CV::ld(f.p[C], (R)scale);
    if (vel) {
        double *o = vel + 3 * (r * stride + n);
        o[0] = (double)CV::ld(f.u[C], (R)scale);
        o[1] = (double)CV::ld(f.v[C], (R)scale);
        o[2] = (double)CV::ld(f.w[C], (R)scale);
    }
}

// ------------------------------------------------------------------------------------
// Host context
// ------------------------------------------------------------------------------------

struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct fases_ctx {
    fases_params prm;
    std::string err;
    hipStream_t stream = nullptr;
    bool have_geo = false, have_mat = false;
    int Nx = 0, Ny = 0, Nz = 0;
    long long N = 0, n_upd = 0;
    size_t esz = 8;                      // bytes per stored field value
    std::vector<DevBuf> bufs;            // every device allocation, freed at destroy
    size_t dev_bytes = 0;
    void *U[3][4] = {};                  // three field sets x (p, u, v, w)
    uint8_t *d_upd = nullptr;
    std::vector<uint8_t> h_upd;
    int n_patch = 0;                     // cells kept in FP64 shadows (near sources)
    int *d_patch = nullptr;              // their grid indices
    int *d_pidx = nullptr;               // grid -> patch index, -1 elsewhere (allocated on demand)
    double *shadow[3] = {};              // FP64 shadow per field set, [4][n_patch]
    std::vector<int> h_patch;
    // boundary list
    int nB = 0;
    int *d_bcell = nullptr, *d_bbase = nullptr, *d_bmat = nullptr;
    double *d_bgeo = nullptr;
    std::vector<int> h_bcell, h_bmat;
    // materials and ADE states
    int nMat = 0, K = 1, L = 1;
    double *d_Rinf = nullptr, *d_A = nullptr, *d_zeta = nullptr, *d_B = nullptr, *d_C = nullptr,
           *d_alpha = nullptr, *d_beta = nullptr;
    int *d_Keff = nullptr, *d_Leff = nullptr;
    double *phi[3] = {}, *psi1[3] = {}, *psi2[3] = {};   // [0]=phi, [1]=phi_old, [2]=phi_nu
    // PML
    PmlSet pml{};
    // sources / receivers
    int n_src = 0, src_opt = FASES_SRC_RICKER;
    long long n_samp = 0;
    int *d_scell = nullptr;
    double *d_sval = nullptr, *d_sfreq = nullptr, *d_us = nullptr;
    std::vector<double> h_sfreq;         // sine sources: frequencies (samples tabulated per run)
    int n_rec = 0;
    int *d_rcell = nullptr;
    double *d_ir = nullptr, *d_vel = nullptr;
    long long ir_cap = 0;
    long long step = 0;
    std::vector<uint8_t> h_oid;

    ~fases_ctx() {
        for (auto &b : bufs) if (b.ptr) (void)hipFree(b.ptr);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <typename T>
int dalloc(fases_ctx *ctx, T **out, size_t count) {
    *out = nullptr;
    if (count == 0) count = 1;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
        ctx->err = std::string("hipMalloc(") + std::to_string(count * sizeof(T)) + " B): " + hipGetErrorString(e);
        return FASES_E_NOMEM;
    }
    e = hipMemsetAsync(p, 0, count * sizeof(T), ctx->stream);
    if (e != hipSuccess) { (void)hipFree(p); ctx->err = hipGetErrorString(e); return FASES_E_HIP; }
    ctx->bufs.push_back({p, count * sizeof(T)});
    ctx->dev_bytes += count * sizeof(T);
    *out = static_cast<T *>(p);
    return FASES_OK;
}

void dfree_all(fases_ctx *ctx, void *keep_none = nullptr) {
    (void)keep_none;
    for (auto &b : ctx->bufs) if (b.ptr) (void)hipFree(b.ptr);
    ctx->bufs.clear();
    ctx->dev_bytes = 0;
}

template <typename T>
int upload(fases_ctx *ctx, T **dst, const T *src, size_t count) {
    int rc = dalloc(ctx, dst, count);
    if (rc) return rc;
    if (count) HIPCHK(hipMemcpyAsync(*dst, src, count * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return FASES_OK;
}

// cPML profile of one axis (functions3D.cuh:681-756) in FP64: depth x in [0,1] from the
// interior edge (3 + d, or N-4-d) to the outermost updated cell, sigma = sigma_max x^2,
// kappa = 1 + (kappa_max - 1) x^2, alpha = alpha_min + (alpha_max - alpha_min)(1 - x),
// alpha_min = alpha_max / 10.
void pml_axis_tables(const fases_params &P, int N, float key_lo, float key_hi, std::vector<int> &band,
                     std::vector<double> &kc, std::vector<double> &bb, std::vector<double> &aa) {
    const int d = P.pml_cells;
    const double m = 2.0;
    double smax = P.pml_sigma_max;
    if (!(smax > 0.0)) smax = 0.5 * P.c * (m + 1.0) * std::log(1.0 / 1e-7) / ((double)d * P.h);
    const double kmax = P.pml_kappa_max, amax = P.pml_alpha_max, amin = 0.1 * amax;
    band.assign((size_t)N, -1);
    kc.assign((size_t)N, 0.0);
    bb.assign((size_t)N, 1.0);
    aa.assign((size_t)N, 0.0);
    const double lo = key_lo, hi = key_hi;
    for (int i = 0; i < N; ++i) {
        const bool in_lo = key_lo > 0.f && i >= NG && i < NG + d;
        const bool in_hi = key_hi > 0.f && i >= N - NG - d && i <= N - 1 - NG;
        if (!(in_lo || in_hi)) continue;
        band[(size_t)i] = (i < NG + d) ? (i - NG) : (i - (N - NG - d) + d);
        const double xm = std::fmin(1.0, std::fmax(0.0, ((double)(NG + d) - (double)i) / (double)d));
        const double xp = std::fmin(1.0, std::fmax(0.0, ((double)i - ((double)(N - 1 - NG) - (double)d)) / (double)d));
        const double sig = lo * smax * std::pow(xm, m) + hi * smax * std::pow(xp, m);
        const double dep = lo * xm + hi * xp;
        const double kap = 1.0 + (kmax - 1.0) * std::pow(dep, m);
        const double alp = amin + (amax - amin) * (1.0 - dep);
        kc[(size_t)i] = 1.0 / kap - 1.0;
        const double b = std::exp(-(alp + sig / kap) * P.dt);
        bb[(size_t)i] = b;
        aa[(size_t)i] = (sig > 0.0) ? (sig / (kap * (sig + kap * alp))) * (b - 1.0) : 0.0;
    }
}

int set_error(fases_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    else fases::set_global_error(msg);
    return code;
}

template <typename S>
FieldSet<S> fset(fases_ctx *ctx, int s) {
    return FieldSet<S>{(S *)ctx->U[s][0], (S *)ctx->U[s][1], (S *)ctx->U[s][2], (S *)ctx->U[s][3]};
}

template <typename R>
Split<R> split_const(double v) {
    const R hi = (R)v;
    return Split<R>{hi, (R)(v - (double)hi)};
}

template <typename R>
TileK<R> tile_constants(const fases_params &P) {
    TileK<R> K;
    K.kp = split_const<R>(0.5 * P.rho * P.c * P.c);
    K.kr = split_const<R>(0.5 / P.rho);
    K.dth = split_const<R>(P.dt / P.h);
    K.third = split_const<R>(1.0 / 3.0);
    K.kc = (R)(0.5 * P.c);
    K.ih = (R)(1.0 / P.h);
    return K;
}

template <typename S, typename R, int NW>
int launch_tiled(fases_ctx *ctx, int stage, int in, int out, R rho, R c, R h, R dt) {
    (void)rho; (void)c; (void)h;
    const TileK<R> K = tile_constants<R>(ctx->prm);
    const FieldSet<S> fi = fset<S>(ctx, in), fb = fset<S>(ctx, 0), fo = fset<S>(ctx, out);
    const dim3 tb(64, NW);
    const dim3 tg((ctx->Nx - 6 + tile::TXO - 1) / tile::TXO, (ctx->Ny - 6 + (NW - 2) - 1) / (NW - 2),
                  (ctx->Nz - 6 + tile::TZ - 1) / tile::TZ);
    const R one = R(1);
    if (stage == 0)
        hipLaunchKernelGGL((k_flux_tiled<S, R, 0, NW>), tg, tb, 0, ctx->stream, ctx->Nx, ctx->Ny, ctx->Nz, fi.p, fi.u,
                           fi.v, fi.w, fb.p, fb.u, fb.v, fb.w, fo.p, fo.u, fo.v, fo.w, ctx->d_upd, ctx->pml, K, dt,
                           one, one, one);
    else if (stage == 1)
        hipLaunchKernelGGL((k_flux_tiled<S, R, 1, NW>), tg, tb, 0, ctx->stream, ctx->Nx, ctx->Ny, ctx->Nz, fi.p, fi.u,
                           fi.v, fi.w, fb.p, fb.u, fb.v, fb.w, fo.p, fo.u, fo.v, fo.w, ctx->d_upd, ctx->pml, K, dt,
                           one, one, one);
    else
        hipLaunchKernelGGL((k_flux_tiled<S, R, 2, NW>), tg, tb, 0, ctx->stream, ctx->Nx, ctx->Ny, ctx->Nz, fi.p, fi.u,
                           fi.v, fi.w, fb.p, fb.u, fb.v, fb.w, fo.p, fo.u, fo.v, fo.w, ctx->d_upd, ctx->pml, K, dt,
                           one, one, one);
    return FASES_OK;
}

template <typename S, typename R>
int launch_step(fases_ctx *ctx, long long n, std::vector<hipEvent_t> *kev) {
    int rc = FASES_OK;
    Grid g{ctx->Nx, ctx->Ny, ctx->Nz, (long long)ctx->Nx * ctx->Ny};
    const fases_params &P = ctx->prm;
    const R rho = (R)P.rho, c = (R)P.c, h = (R)P.h, dt = (R)P.dt;
    const R one = R(1);
    dim3 blk(64, 4, 1);
    dim3 grd((ctx->Nx - 2 * NG + 63) / 64, (ctx->Ny - 2 * NG + 3) / 4, ctx->Nz - 2 * NG);
    const double As = 6.0 * P.h * P.h;
    const int sb = 256;
    Ade ade{ctx->K, ctx->L, ctx->d_Rinf, ctx->d_A, ctx->d_zeta, ctx->d_B, ctx->d_C, ctx->d_alpha, ctx->d_beta,
            ctx->d_Keff, ctx->d_Leff};
    // Geometry stays FP64 for every field precision: a blend weight or normal rounded
    // once to FP32 is a bias that every step repeats (measured: FP32-room IR error x3.5).
    typedef double G;
    Bnd<G> bd{ctx->nB, ctx->d_bcell, ctx->d_bbase, ctx->d_bmat, ctx->d_bgeo};
    // (field set in, set out, RK weight, source sample, ADE a,b,c, ADE buffers left/mid/right)
    struct StageCfg { int in, out; double w; long long nn; double a, b, c; int l, mid, r; };
    const StageCfg cfg[3] = {
        {0, 1, 1.0, 2 * n, 0.0, 1.0, 1.0, 0, 1, 1},
        {1, 2, 0.25, 2 * n + 2, 3.0, 1.0, 4.0, 2, 1, 0},
        {2, 1, 2.0 / 3.0, 2 * n + 1, 1.0, 2.0, 3.0, 0, 1, 2},
    };
    for (int s = 0; s < 3; ++s) {
        const StageCfg &q = cfg[s];
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (kev) {
            HIPCHK(hipEventCreate(&e0));
            HIPCHK(hipEventCreate(&e1));
            kev->push_back(e0);
            kev->push_back(e1);
            HIPCHK(hipEventRecord(e0, ctx->stream));
        }
        const int kv = ctx->prm.kernel_variant;
        if (kv == 1 || (kv == 0 && ctx->prm.precision == FASES_FP64)) {
            const GridIO<S, R> gio{fset<S>(ctx, q.in), fset<S>(ctx, 0), fset<S>(ctx, q.out), one, one, one};
            if (s == 0)
                hipLaunchKernelGGL((k_flux_simple<S, R, 0>), grd, blk, 0, ctx->stream, g, gio, ctx->d_upd, ctx->pml,
                                   rho, c, h, dt);
            else if (s == 1)
                hipLaunchKernelGGL((k_flux_simple<S, R, 1>), grd, blk, 0, ctx->stream, g, gio, ctx->d_upd, ctx->pml,
                                   rho, c, h, dt);
            else
                hipLaunchKernelGGL((k_flux_simple<S, R, 2>), grd, blk, 0, ctx->stream, g, gio, ctx->d_upd, ctx->pml,
                                   rho, c, h, dt);
        } else {
            rc = launch_tiled<S, R, 16>(ctx, s, q.in, q.out, rho, c, h, dt);
            if (rc) return rc;
        }
        if (kev) HIPCHK(hipEventRecord(e1, ctx->stream));
        if (ctx->n_patch > 0) {
            const dim3 pg((ctx->n_patch + 63) / 64), pb(64);
            const PatchIO<S> pio{fset<S>(ctx, q.in), fset<S>(ctx, 0), fset<S>(ctx, q.out), ctx->d_pidx,
                                 ctx->shadow[q.in], ctx->shadow[0], ctx->shadow[q.out], ctx->n_patch, 1.0, 1.0, 1.0};
            if (s == 0)
                hipLaunchKernelGGL((k_flux_patch<S, 0>), pg, pb, 0, ctx->stream, g, pio, ctx->d_patch, ctx->pml,
                                   P.rho, P.c, P.h, P.dt);
            else if (s == 1)
                hipLaunchKernelGGL((k_flux_patch<S, 1>), pg, pb, 0, ctx->stream, g, pio, ctx->d_patch, ctx->pml,
                                   P.rho, P.c, P.h, P.dt);
            else
                hipLaunchKernelGGL((k_flux_patch<S, 2>), pg, pb, 0, ctx->stream, g, pio, ctx->d_patch, ctx->pml,
                                   P.rho, P.c, P.h, P.dt);
        }
        if (ctx->n_src > 0)
            hipLaunchKernelGGL((k_sources<S, R>), dim3((ctx->n_src + sb - 1) / sb), dim3(sb), 0, ctx->stream,
                               (S *)ctx->U[q.out][0], ctx->d_scell, ctx->d_sval, ctx->d_sfreq, ctx->d_us,
                               ctx->src_opt == FASES_SRC_SINE ? ctx->n_samp : 0LL, ctx->n_src, ctx->src_opt, q.nn, q.w, P.dt, P.c, P.rho, As, P.h,
                               ctx->n_patch > 0 ? ctx->d_pidx : nullptr, ctx->shadow[q.out], one, one);
        if (ctx->nB > 0)
            hipLaunchKernelGGL((k_boundary<S, G>), dim3((ctx->nB + sb - 1) / sb), dim3(sb), 0, ctx->stream, g,
                               fset<S>(ctx, q.in), fset<S>(ctx, q.out), bd, ade, P.rho, P.c, P.dt, q.a, q.b, q.c,
                               ctx->phi[q.l], ctx->phi[q.mid], ctx->phi[q.r], ctx->psi1[q.l], ctx->psi1[q.mid],
                               ctx->psi1[q.r], ctx->psi2[q.l], ctx->psi2[q.mid], ctx->psi2[q.r], 1.0, 1.0, 1.0);
    }
    // rotate: U0 <-> U1, phi_old <-> phi (main3D.cu:1312-1319)
    for (int f = 0; f < 4; ++f) std::swap(ctx->U[0][f], ctx->U[1][f]);
    std::swap(ctx->phi[0], ctx->phi[1]);
    std::swap(ctx->psi1[0], ctx->psi1[1]);
    std::swap(ctx->psi2[0], ctx->psi2[1]);
    std::swap(ctx->shadow[0], ctx->shadow[1]);
    return FASES_OK;
}

template <typename S>
int launch_receivers(fases_ctx *ctx, long long stride, long long n) {
    if (ctx->n_rec <= 0) return FASES_OK;
    hipLaunchKernelGGL((k_receivers<S>), dim3((ctx->n_rec + 63) / 64), dim3(64), 0, ctx->stream, fset<S>(ctx, 0),
                       ctx->d_rcell, ctx->n_rec, ctx->d_ir, ctx->d_vel, stride, n, 1.0);
    return FASES_OK;
}

template <typename S, typename R>
int run_impl(fases_ctx *ctx, long long nsteps, double *ir, double *vel, double *elapsed_ms, double *kernel_ms) {
    const long long need = (long long)ctx->n_rec * nsteps;
    if (ctx->n_src > 0 && ctx->src_opt != FASES_SRC_SINE && 2 * (ctx->step + nsteps) + 1 > ctx->n_samp)
        return set_error(ctx, FASES_E_ARG, "source waveform shorter than 2*(steps)+1 samples");
    if (ctx->n_src > 0 && ctx->src_opt == FASES_SRC_SINE && 2 * (ctx->step + nsteps) + 1 > ctx->n_samp) {
        // per-source sine samples sin(2 pi f nn dt/2), tabulated on the host (same libm
        // as a CPU evaluation of the reference formula)
        const long long ns = 2 * (ctx->step + nsteps) + 1;
        std::vector<double> tab((size_t)ctx->n_src * ns);
        for (int q = 0; q < ctx->n_src; ++q)
            for (long long t = 0; t < ns; ++t)
                tab[(size_t)q * ns + t] = std::sin(2.0 * M_PI * ctx->h_sfreq[(size_t)q] * (double)t * (0.5 * ctx->prm.dt));
        int rc = upload(ctx, &ctx->d_us, tab.data(), tab.size());
        if (rc) return rc;
        ctx->n_samp = ns;
    }
    if (ctx->n_rec > 0 && nsteps > ctx->ir_cap) {
        // grow the device receiver buffers
        int rc = dalloc(ctx, &ctx->d_ir, (size_t)need);
        if (rc) return rc;
        rc = dalloc(ctx, &ctx->d_vel, (size_t)need * 3);
        if (rc) return rc;
        ctx->ir_cap = nsteps;
    }
    if (ctx->n_src > 0 && ctx->src_opt != FASES_SRC_SINE && 2 * (ctx->step + nsteps) + 1 > ctx->n_samp)
        return set_error(ctx, FASES_E_ARG, "source waveform shorter than 2*(steps)+1 samples");
    hipEvent_t ev[2];
    for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    std::vector<hipEvent_t> kev;
    if (kernel_ms) kev.reserve((size_t)nsteps * 6);
    HIPCHK(hipEventRecord(ev[0], ctx->stream));
    for (long long q = 0; q < nsteps; ++q) {
        int rc = launch_step<S, R>(ctx, ctx->step, kernel_ms ? &kev : nullptr);
        if (rc) return rc;
        rc = launch_receivers<S>(ctx, nsteps, q);
        if (rc) return rc;
        ctx->step += 1;
    }
    HIPCHK(hipEventRecord(ev[1], ctx->stream));
    HIPCHK(hipEventSynchronize(ev[1]));
    HIPCHK(hipGetLastError());
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[1]));
    if (elapsed_ms) *elapsed_ms = ms;
    for (auto &e : ev) (void)hipEventDestroy(e);
    if (kernel_ms) {
        double tot = 0.0;
        for (size_t q = 0; q + 1 < kev.size(); q += 2) {
            float km = 0.f;
            HIPCHK(hipEventElapsedTime(&km, kev[q], kev[q + 1]));
            tot += km;
        }
        for (auto &e : kev) (void)hipEventDestroy(e);
        *kernel_ms = tot;
    }
    if (ctx->n_rec > 0) {
        if (ir) HIPCHK(hipMemcpy(ir, ctx->d_ir, (size_t)need * sizeof(double), hipMemcpyDeviceToHost));
        if (vel) HIPCHK(hipMemcpy(vel, ctx->d_vel, (size_t)need * 3 * sizeof(double), hipMemcpyDeviceToHost));
        if (ir) {
            for (long long i = 0; i < need; ++i)
                if (!std::isfinite(ir[i])) return set_error(ctx, FASES_E_NONFINITE, "receiver sample is not finite");
        }
    }
    return FASES_OK;
}

}  // namespace

extern "C" {

const char *fases_last_error(const fases_ctx *ctx) {
    if (ctx) return ctx->err.c_str();
    return fases::global_error();
}

int fases_create(const fases_params *p, fases_ctx **out) {
    if (!p || !out) return set_error(nullptr, FASES_E_ARG, "fases_create: null argument");
    *out = nullptr;
    if (!(p->h > 0) || !(p->dt > 0) || !(p->c > 0) || !(p->rho > 0))
        return set_error(nullptr, FASES_E_ARG, "fases_create: h, dt, c, rho must be > 0");
    if (p->precision != FASES_FP64 && p->precision != FASES_FP32 && p->precision != FASES_FP16)
        return set_error(nullptr, FASES_E_ARG, "fases_create: unknown precision");
    if (p->precision == FASES_FP16)
        return set_error(nullptr, FASES_E_ARG, "fases_create: FP16 storage is not built in this library yet");
    if (p->pml_cells < 1) return set_error(nullptr, FASES_E_ARG, "fases_create: pml_cells must be >= 1");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev < 1)
        return set_error(nullptr, FASES_E_HIP, std::string("no HIP device: ") + hipGetErrorString(e));
    if (p->device < 0 || p->device >= ndev) return set_error(nullptr, FASES_E_ARG, "fases_create: bad device index");
    e = hipSetDevice(p->device);
    if (e != hipSuccess) return set_error(nullptr, FASES_E_HIP, hipGetErrorString(e));
    std::unique_ptr<fases_ctx> ctx(new (std::nothrow) fases_ctx());
    if (!ctx) return set_error(nullptr, FASES_E_NOMEM, "fases_create: out of host memory");
    ctx->prm = *p;
    if (!(ctx->prm.tol > 0)) ctx->prm.tol = 3.0 * p->h;
    ctx->esz = (p->precision == FASES_FP64) ? 8 : (p->precision == FASES_FP32 ? 4 : 2);
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return set_error(nullptr, FASES_E_HIP, hipGetErrorString(e));
    *out = ctx.release();
    return FASES_OK;
}

int fases_destroy(fases_ctx *ctx) {
    if (!ctx) return FASES_OK;
    (void)hipSetDevice(ctx->prm.device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
    return FASES_OK;
}

int fases_set_geometry(fases_ctx *ctx, const float *S, const uint8_t *oid, int Nx, int Ny, int Nz) {
    if (!ctx || !S || !oid) return set_error(ctx, FASES_E_ARG, "fases_set_geometry: null argument");
    if (Nx < 2 * NG + 1 || Ny < 2 * NG + 1 || Nz < 2 * NG + 1)
        return set_error(ctx, FASES_E_ARG, "fases_set_geometry: grid smaller than 7 cells on an axis");
    const long long N = (long long)Nx * Ny * Nz;
    if (N >= (1LL << 31)) return set_error(ctx, FASES_E_ARG, "fases_set_geometry: grid exceeds 2^31 cells");
    HIPCHK(hipSetDevice(ctx->prm.device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    dfree_all(ctx);
    ctx->have_geo = ctx->have_mat = false;
    ctx->Nx = Nx; ctx->Ny = Ny; ctx->Nz = Nz; ctx->N = N;
    const fases_params &P = ctx->prm;
    const long long nxy = (long long)Nx * Ny;
    const float hf = (float)P.h;
    int rc;
    // fields: three sets of p, u, v, w
    for (int s = 0; s < 3; ++s)
        for (int f = 0; f < 4; ++f) {
            if (ctx->esz == 8) rc = dalloc(ctx, (double **)&ctx->U[s][f], (size_t)N);
            else if (ctx->esz == 4) rc = dalloc(ctx, (float **)&ctx->U[s][f], (size_t)N);
            else rc = dalloc(ctx, (__half **)&ctx->U[s][f], (size_t)N);
            if (rc) return rc;
        }
    // update mask: interior [3, N-3)^3 with S >= h (functions3D.cuh:1091-1097)
    std::vector<uint8_t> upd((size_t)N, 0);
    long long n_upd = 0;
    for (int k = NG; k < Nz - NG; ++k)
        for (int j = NG; j < Ny - NG; ++j)
            for (int i = NG; i < Nx - NG; ++i) {
                const long long C = i + (long long)Nx * j + nxy * k;
                if (!(S[C] < hf)) { upd[(size_t)C] = 1; ++n_upd; }
            }
    ctx->n_upd = n_upd;
    if ((rc = upload(ctx, &ctx->d_upd, upd.data(), (size_t)N))) return rc;
    ctx->h_upd = upd;
    ctx->n_patch = 0;
    ctx->d_patch = nullptr;
    ctx->d_pidx = nullptr;
    ctx->h_patch.clear();
    for (auto &sh : ctx->shadow) sh = nullptr;
    // boundary list in bmap order: cells in [3, N-3)^3 with oid in [1, 254]
    // (functions3D.cuh:466-479, filterABC.h:95-108) and their fixed geometry.
    std::vector<int> bcell, bbase, bmat;
    std::vector<double> bgeo;
    const double h = P.h, tol = P.tol;
    for (long long C = 0; C < N; ++C) {
        if (oid[C] == 0 || oid[C] == 0xFF) continue;
        const int i = (int)(C % Nx), j = (int)((C / Nx) % Ny), k = (int)(C / nxy);
        if (i < NG || i >= Nx - NG || j < NG || j >= Ny - NG || k < NG || k >= Nz - NG) continue;
        const double s = S[C];
        // normal n = -grad S/|grad S| (central differences), foot point W = C + s n,
        // probe D = W - (s + 2h) n, trilinear base corner of D (clamped to the grid).
        const double gx = ((double)S[C + 1] - (double)S[C - 1]) * (0.5 / h);
        const double gy = ((double)S[C + Nx] - (double)S[C - Nx]) * (0.5 / h);
        const double gz = ((double)S[C + nxy] - (double)S[C - nxy]) * (0.5 / h);
        double gn = std::sqrt(gx * gx + gy * gy + gz * gz);
        gn = std::fmax(gn, 1e-12);
        const double nx = -gx / gn, ny = -gy / gn, nz = -gz / gn;
        const double Cx = (i + 0.5) * h, Cy = (j + 0.5) * h, Cz = (k + 0.5) * h;
        const double Wx = Cx + s * nx, Wy = Cy + s * ny, Wz = Cz + s * nz;
        const double l = s + 2.0 * h;
        const double Dx = Wx - l * nx, Dy = Wy - l * ny, Dz = Wz - l * nz;
        const double qx = Dx / h - 0.5, qy = Dy / h - 0.5, qz = Dz / h - 0.5;
        int bi = (int)std::floor(qx), bj = (int)std::floor(qy), bk = (int)std::floor(qz);
        bi = std::max(0, std::min(bi, Nx - 2));
        bj = std::max(0, std::min(bj, Ny - 2));
        bk = std::max(0, std::min(bk, Nz - 2));
        bcell.push_back((int)C);
        bbase.push_back(bi + Nx * (bj + Ny * bk));
        const double fx = qx - (double)bi, fy = qy - (double)bj, fz = qz - (double)bk;
        const int centred = (fx == 0.0 && fy == 0.0 && fz == 0.0) ? 0x100 : 0;
        bmat.push_back(((int)oid[C] - 1) | centred);
        bgeo.push_back(fx);
        bgeo.push_back(fy);
        bgeo.push_back(fz);
        bgeo.push_back(nx);
        bgeo.push_back(ny);
        bgeo.push_back(nz);
        bgeo.push_back(chi_band(s, h, tol));
        bgeo.push_back(0.0);
    }
    ctx->nB = (int)bcell.size();
    if ((rc = upload(ctx, &ctx->d_bcell, bcell.data(), bcell.size()))) return rc;
    if ((rc = upload(ctx, &ctx->d_bbase, bbase.data(), bbase.size()))) return rc;
    if ((rc = upload(ctx, &ctx->d_bmat, bmat.data(), bmat.size()))) return rc;
    if ((rc = upload(ctx, &ctx->d_bgeo, bgeo.data(), bgeo.size()))) return rc;
    ctx->h_bcell = bcell;
    ctx->h_bmat = bmat;
    ctx->h_oid.assign(oid, oid + N);
    // PML tables and psi tiles (main3D.cu:804-828)
    const float *key = P.pml_keys;   // U D R L F B
    ctx->pml = PmlSet{};
    ctx->pml.d = P.pml_cells;
    ctx->pml.on = (key[0] != 0.f || key[1] != 0.f || key[2] != 0.f || key[3] != 0.f || key[4] != 0.f || key[5] != 0.f);
    {
        std::vector<int> band;
        std::vector<double> kc, bb, aa;
        auto axis = [&](PmlAxis &ax, int n, float lo, float hi) -> int {
            pml_axis_tables(P, n, lo, hi, band, kc, bb, aa);
            int r;
            int *db; double *dk, *dbb, *da;
            if ((r = upload(ctx, &db, band.data(), band.size()))) return r;
            if ((r = upload(ctx, &dk, kc.data(), kc.size()))) return r;
            if ((r = upload(ctx, &dbb, bb.data(), bb.size()))) return r;
            if ((r = upload(ctx, &da, aa.data(), aa.size()))) return r;
            ax.band = db; ax.kc = dk; ax.b = dbb; ax.a = da;
            return FASES_OK;
        };
        if ((rc = axis(ctx->pml.x, Nx, key[3], key[2]))) return rc;   // L = x-, R = x+
        if ((rc = axis(ctx->pml.y, Ny, key[5], key[4]))) return rc;   // B = y-, F = y+
        if ((rc = axis(ctx->pml.z, Nz, key[1], key[0]))) return rc;   // D = z-, U = z+
        const size_t d2 = 2 * (size_t)P.pml_cells;
        const size_t nx_t = ctx->pml.on ? d2 * Ny * Nz : 1, ny_t = ctx->pml.on ? (size_t)Nx * d2 * Nz : 1,
                     nz_t = ctx->pml.on ? (size_t)Nx * Ny * d2 : 1;
        if ((rc = dalloc(ctx, &ctx->pml.psi_px, nx_t))) return rc;
        if ((rc = dalloc(ctx, &ctx->pml.psi_ux, nx_t))) return rc;
        if ((rc = dalloc(ctx, &ctx->pml.psi_py, ny_t))) return rc;
        if ((rc = dalloc(ctx, &ctx->pml.psi_vy, ny_t))) return rc;
        if ((rc = dalloc(ctx, &ctx->pml.psi_pz, nz_t))) return rc;
        if ((rc = dalloc(ctx, &ctx->pml.psi_wz, nz_t))) return rc;
    }
    ctx->n_src = 0; ctx->n_rec = 0; ctx->ir_cap = 0; ctx->step = 0;
    ctx->d_ir = ctx->d_vel = nullptr;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->have_geo = true;
    return FASES_OK;
}

int fases_set_materials(fases_ctx *ctx, int nMat, int K, int L, const double *Rinf, const double *A,
                        const double *zeta, const double *B, const double *C, const double *alpha,
                        const double *beta) {
    if (!ctx) return FASES_E_ARG;
    if (!ctx->have_geo) return set_error(ctx, FASES_E_STATE, "fases_set_materials: call fases_set_geometry first");
    if (nMat < 1 || K < 1 || L < 1 || !Rinf || !A || !zeta || !B || !C || !alpha || !beta)
        return set_error(ctx, FASES_E_ARG, "fases_set_materials: bad argument");
    for (int mw : ctx->h_bmat)
        if ((mw & 0xFF) >= nMat) return set_error(ctx, FASES_E_ARG, "fases_set_materials: oid refers to a missing material");
    HIPCHK(hipSetDevice(ctx->prm.device));
    std::vector<int> Ke((size_t)nMat, 0), Le((size_t)nMat, 0);
    for (int m = 0; m < nMat; ++m) {
        for (int k = 0; k < K; ++k) if (A[(size_t)m * K + k] != 0.0) Ke[(size_t)m] = k + 1;
        for (int l = 0; l < L; ++l)
            if (B[(size_t)m * L + l] != 0.0 || C[(size_t)m * L + l] != 0.0) Le[(size_t)m] = l + 1;
    }
    int rc;
    ctx->nMat = nMat; ctx->K = K; ctx->L = L;
    if ((rc = upload(ctx, &ctx->d_Rinf, Rinf, (size_t)nMat))) return rc;
    if ((rc = upload(ctx, &ctx->d_A, A, (size_t)nMat * K))) return rc;
    if ((rc = upload(ctx, &ctx->d_zeta, zeta, (size_t)nMat * K))) return rc;
    if ((rc = upload(ctx, &ctx->d_B, B, (size_t)nMat * L))) return rc;
    if ((rc = upload(ctx, &ctx->d_C, C, (size_t)nMat * L))) return rc;
    if ((rc = upload(ctx, &ctx->d_alpha, alpha, (size_t)nMat * L))) return rc;
    if ((rc = upload(ctx, &ctx->d_beta, beta, (size_t)nMat * L))) return rc;
    if ((rc = upload(ctx, &ctx->d_Keff, Ke.data(), Ke.size()))) return rc;
    if ((rc = upload(ctx, &ctx->d_Leff, Le.data(), Le.size()))) return rc;
    for (int q = 0; q < 3; ++q) {
        if ((rc = dalloc(ctx, &ctx->phi[q], (size_t)ctx->nB * K))) return rc;
        if ((rc = dalloc(ctx, &ctx->psi1[q], (size_t)ctx->nB * L))) return rc;
        if ((rc = dalloc(ctx, &ctx->psi2[q], (size_t)ctx->nB * L))) return rc;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->have_mat = true;
    return FASES_OK;
}

int fases_set_sources(fases_ctx *ctx, int n, const int64_t *cell, const double *sval, int src_opt,
                      const double *u_s, int64_t n_samp, const double *freq) {
    if (!ctx) return FASES_E_ARG;
    if (!ctx->have_geo) return set_error(ctx, FASES_E_STATE, "fases_set_sources: call fases_set_geometry first");
    if (n < 0 || (n > 0 && (!cell || !sval))) return set_error(ctx, FASES_E_ARG, "fases_set_sources: bad argument");
    if (n > 0 && src_opt != FASES_SRC_SINE && (!u_s || n_samp < 1))
        return set_error(ctx, FASES_E_ARG, "fases_set_sources: waveform missing");
    if (n > 0 && src_opt == FASES_SRC_SINE && !freq)
        return set_error(ctx, FASES_E_ARG, "fases_set_sources: sine sources need freq");
    HIPCHK(hipSetDevice(ctx->prm.device));
    // merge sources that share a cell (the reference's non-atomic += would race, with the
    // same sum as the sequential result): same-cell sources add their drive amplitudes.
    std::vector<int> mc;
    std::vector<double> ms, mf;
    for (int s = 0; s < n; ++s) {
        if (cell[s] < 0 || cell[s] >= ctx->N) return set_error(ctx, FASES_E_RANGE, "source outside the grid");
        int found = -1;
        if (src_opt != FASES_SRC_SINE)
            for (size_t q = 0; q < mc.size(); ++q) if (mc[q] == (int)cell[s]) { found = (int)q; break; }
        if (found >= 0) ms[(size_t)found] += sval[s];
        else { mc.push_back((int)cell[s]); ms.push_back(sval[s]); mf.push_back(freq ? freq[s] : 0.0); }
    }
    int rc;
    // Reduced-precision storage: cells within kSourcePatch (Chebyshev) of a source keep
    // FP64 shadows of all three field sets and are updated in FP64 arithmetic. Inside a
    // step the source cell swings by the full per-stage injection (far above the local
    // pressure) and the RK combination cancels it; FP32 storage or arithmetic in that
    // neighbourhood seeds errors that propagate to every receiver (measured: IR rel L2
    // 1.7e-4 far from the source with no patch, 2e-6 with radius 12).
    if (ctx->prm.precision != FASES_FP64) {
        const int Nx = ctx->Nx, Ny = ctx->Ny, Nz = ctx->Nz;
        const long long nxy = (long long)Nx * Ny;
        for (auto &v : ctx->h_upd) if (v == 2) v = 1;
        std::vector<int> patch;
        std::vector<uint8_t> is_bnd((size_t)ctx->N, 0);
        for (int c0 : ctx->h_bcell) is_bnd[(size_t)c0] = 1;
        for (int c0 : mc) {
            const int si = c0 % Nx, sj = (int)((c0 / Nx) % Ny), sk = (int)(c0 / nxy);
            for (int k = std::max(0, sk - kSourcePatch); k <= std::min(Nz - 1, sk + kSourcePatch); ++k)
                for (int j = std::max(0, sj - kSourcePatch); j <= std::min(Ny - 1, sj + kSourcePatch); ++j)
                    for (int i = std::max(0, si - kSourcePatch); i <= std::min(Nx - 1, si + kSourcePatch); ++i) {
                        const long long C = i + (long long)Nx * j + nxy * k;
                        // boundary cells stay out: the boundary kernel blends their stored values
                        if (ctx->h_upd[(size_t)C] == 1 && !is_bnd[(size_t)C]) {
                            ctx->h_upd[(size_t)C] = 2;
                            patch.push_back((int)C);
                        }
                    }
        }
        HIPCHK(hipMemcpyAsync(ctx->d_upd, ctx->h_upd.data(), ctx->h_upd.size(), hipMemcpyHostToDevice, ctx->stream));
        ctx->n_patch = (int)patch.size();
        ctx->h_patch = patch;
        if (!patch.empty()) {
            if (!ctx->d_pidx && (rc = dalloc(ctx, &ctx->d_pidx, (size_t)ctx->N))) return rc;
            std::vector<int> pidx((size_t)ctx->N, -1);
            for (size_t q = 0; q < patch.size(); ++q) pidx[(size_t)patch[q]] = (int)q;
            HIPCHK(hipMemcpyAsync(ctx->d_pidx, pidx.data(), pidx.size() * sizeof(int), hipMemcpyHostToDevice,
                                  ctx->stream));
            if ((rc = upload(ctx, &ctx->d_patch, patch.data(), patch.size()))) return rc;
            for (int q = 0; q < 3; ++q) {
                if ((rc = dalloc(ctx, &ctx->shadow[q], (size_t)4 * patch.size()))) return rc;
                if (ctx->esz == 4)
                    hipLaunchKernelGGL((k_patch_gather<float>), dim3((ctx->n_patch + 63) / 64), dim3(64), 0,
                                       ctx->stream, fset<float>(ctx, q), ctx->d_patch, ctx->n_patch, ctx->shadow[q], 1.0);
            }
        }
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    ctx->n_src = (int)mc.size();
    ctx->src_opt = src_opt;
    ctx->n_samp = (src_opt == FASES_SRC_SINE) ? 0 : n_samp;
    ctx->h_sfreq = mf;
    if ((rc = upload(ctx, &ctx->d_scell, mc.data(), mc.size()))) return rc;
    if ((rc = upload(ctx, &ctx->d_sval, ms.data(), ms.size()))) return rc;
    if ((rc = upload(ctx, &ctx->d_sfreq, mf.data(), mf.size()))) return rc;
    if (src_opt != FASES_SRC_SINE && n > 0) {
        if ((rc = upload(ctx, &ctx->d_us, u_s, (size_t)n_samp))) return rc;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return FASES_OK;
}

int fases_set_receivers(fases_ctx *ctx, int n, const int64_t *cell) {
    if (!ctx) return FASES_E_ARG;
    if (!ctx->have_geo) return set_error(ctx, FASES_E_STATE, "fases_set_receivers: call fases_set_geometry first");
    if (n < 0 || (n > 0 && !cell)) return set_error(ctx, FASES_E_ARG, "fases_set_receivers: bad argument");
    std::vector<int> rc_((size_t)n);
    for (int r = 0; r < n; ++r) {
        if (cell[r] < 0 || cell[r] >= ctx->N) return set_error(ctx, FASES_E_RANGE, "Receiver(s) outside domain");
        rc_[(size_t)r] = (int)cell[r];
    }
    HIPCHK(hipSetDevice(ctx->prm.device));
    int rc;
    ctx->n_rec = n;
    if ((rc = upload(ctx, &ctx->d_rcell, rc_.data(), rc_.size()))) return rc;
    ctx->ir_cap = 0;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return FASES_OK;
}

static int clear_states(fases_ctx *ctx) {
    for (int q = 0; q < 3; ++q) {
        if (ctx->phi[q]) HIPCHK(hipMemsetAsync(ctx->phi[q], 0, (size_t)ctx->nB * ctx->K * sizeof(double), ctx->stream));
        if (ctx->psi1[q]) HIPCHK(hipMemsetAsync(ctx->psi1[q], 0, (size_t)ctx->nB * ctx->L * sizeof(double), ctx->stream));
        if (ctx->psi2[q]) HIPCHK(hipMemsetAsync(ctx->psi2[q], 0, (size_t)ctx->nB * ctx->L * sizeof(double), ctx->stream));
    }
    const size_t d2 = 2 * (size_t)ctx->prm.pml_cells;
    if (ctx->pml.on) {
        const size_t nx_t = d2 * ctx->Ny * ctx->Nz, ny_t = (size_t)ctx->Nx * d2 * ctx->Nz, nz_t = (size_t)ctx->Nx * ctx->Ny * d2;
        HIPCHK(hipMemsetAsync(ctx->pml.psi_px, 0, nx_t * 8, ctx->stream));
        HIPCHK(hipMemsetAsync(ctx->pml.psi_ux, 0, nx_t * 8, ctx->stream));
        HIPCHK(hipMemsetAsync(ctx->pml.psi_py, 0, ny_t * 8, ctx->stream));
        HIPCHK(hipMemsetAsync(ctx->pml.psi_vy, 0, ny_t * 8, ctx->stream));
        HIPCHK(hipMemsetAsync(ctx->pml.psi_pz, 0, nz_t * 8, ctx->stream));
        HIPCHK(hipMemsetAsync(ctx->pml.psi_wz, 0, nz_t * 8, ctx->stream));
    }
    ctx->step = 0;
    return FASES_OK;
}

int fases_set_fields(fases_ctx *ctx, const double *p, const double *u, const double *v, const double *w) {
    if (!ctx) return FASES_E_ARG;
    if (!ctx->have_geo) return set_error(ctx, FASES_E_STATE, "fases_set_fields: call fases_set_geometry first");
    HIPCHK(hipSetDevice(ctx->prm.device));
    const double *src[4] = {p, u, v, w};
    std::vector<uint8_t> tmp((size_t)ctx->N * ctx->esz);
    for (int f = 0; f < 4; ++f) {
        if (ctx->esz == 8) {
            double *t = (double *)tmp.data();
            for (long long c = 0; c < ctx->N; ++c) t[c] = src[f] ? src[f][c] : 0.0;
        } else {
            float *t = (float *)tmp.data();
            for (long long c = 0; c < ctx->N; ++c) t[c] = src[f] ? (float)src[f][c] : 0.f;
        }
        for (int s = 0; s < 3; ++s)
            HIPCHK(hipMemcpyAsync(ctx->U[s][f], tmp.data(), (size_t)ctx->N * ctx->esz, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    if (ctx->n_patch > 0) {
        const size_t np = (size_t)ctx->n_patch;
        std::vector<double> sh(4 * np);
        for (int f = 0; f < 4; ++f)
            for (size_t q = 0; q < np; ++q) sh[f * np + q] = src[f] ? src[f][ctx->h_patch[q]] : 0.0;
        for (int s = 0; s < 3; ++s)
            HIPCHK(hipMemcpyAsync(ctx->shadow[s], sh.data(), sh.size() * sizeof(double), hipMemcpyHostToDevice,
                                  ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    int rc = clear_states(ctx);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return FASES_OK;
}

int fases_reset(fases_ctx *ctx) {
    return fases_set_fields(ctx, nullptr, nullptr, nullptr, nullptr);
}

int fases_get_fields(const fases_ctx *cctx, double *p, double *u, double *v, double *w) {
    fases_ctx *ctx = const_cast<fases_ctx *>(cctx);
    if (!ctx) return FASES_E_ARG;
    if (!ctx->have_geo) return set_error(ctx, FASES_E_STATE, "fases_get_fields: no geometry");
    HIPCHK(hipSetDevice(ctx->prm.device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double *dst[4] = {p, u, v, w};
    std::vector<uint8_t> tmp((size_t)ctx->N * ctx->esz);
    for (int f = 0; f < 4; ++f) {
        if (!dst[f]) continue;
        HIPCHK(hipMemcpy(tmp.data(), ctx->U[0][f], (size_t)ctx->N * ctx->esz, hipMemcpyDeviceToHost));
        if (ctx->esz == 8) std::memcpy(dst[f], tmp.data(), (size_t)ctx->N * 8);
        else {
            const float *t = (const float *)tmp.data();
            for (long long c = 0; c < ctx->N; ++c) dst[f][c] = t[c];
        }
    }
    return FASES_OK;
}

int fases_run(fases_ctx *ctx, int64_t nsteps, double *ir, double *vel, double *elapsed_ms, double *kernel_ms) {
    if (!ctx) return FASES_E_ARG;
    if (!ctx->have_geo || !ctx->have_mat)
        return set_error(ctx, FASES_E_STATE, "fases_run: geometry and materials must be set first");
    if (nsteps < 0) return set_error(ctx, FASES_E_ARG, "fases_run: negative step count");
    if (nsteps == 0) return FASES_OK;
    HIPCHK(hipSetDevice(ctx->prm.device));
    if (ctx->prm.precision == FASES_FP64) return run_impl<double, double>(ctx, nsteps, ir, vel, elapsed_ms, kernel_ms);
    return run_impl<float, float>(ctx, nsteps, ir, vel, elapsed_ms, kernel_ms);
}

int fases_info(const fases_ctx *ctx, int64_t *n_cells, int64_t *n_bnd, int64_t *n_upd, int64_t *dev_bytes) {
    if (!ctx) return FASES_E_ARG;
    if (n_cells) *n_cells = ctx->N;
    if (n_bnd) *n_bnd = ctx->nB;
    if (n_upd) *n_upd = ctx->n_upd;
    if (dev_bytes) *dev_bytes = (int64_t)ctx->dev_bytes;
    return FASES_OK;
}

}  // extern "C"
