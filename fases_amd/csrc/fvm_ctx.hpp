// Solver context (the C ABI's opaque fases_ctx) and the host helpers shared by the
// single-context and batched time loops.
#pragma once

#include "fvm_kernels.hpp"
#include "host.hpp"
#include "../../include/fases_core.h"

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

using namespace fasesdev;

#define HIPCHK(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);             \
            return FASES_E_HIP;                                                       \
        }                                                                             \
    } while (0)

struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
};


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
    // face fluxes of the two-launch patch update: face -1/2 of every (axis, cell) [3][n_patch][2],
    // face +1/2 of the edge items [n_patch_edge][2]; edge items (a n_patch + q) and the item
    // of each (axis, cell) [3][n_patch] (-1: its +1 neighbour is a patch cell)
    double *d_patch_fm = nullptr, *d_patch_fe = nullptr;
    int *d_patch_edges = nullptr, *d_patch_eid = nullptr;
    int n_patch_edge = 0;
    int *d_patch_src = nullptr;          // source index of each patch cell (-1: none), when every
                                         // source lies in the patch (its update injects them)
    std::vector<int> h_patch;
    // boundary list
    int nB = 0;
    int *d_bcell = nullptr, *d_bmat = nullptr;
    float *d_bdist = nullptr;            // signed distance of each boundary cell
    double *d_bnrm = nullptr;            // unit normals, [3][nB]
    std::vector<int> h_bcell, h_bmat;
    std::vector<float> h_bdist;
    std::vector<double> h_bnrm;
    // materials and ADE states
    int nMat = 0, K = 1, L = 1;
    double *d_Rinf = nullptr, *d_A = nullptr, *d_zeta = nullptr, *d_B = nullptr, *d_C = nullptr,
           *d_alpha = nullptr, *d_beta = nullptr;
    int *d_Keff = nullptr, *d_Leff = nullptr;
    double *phi[3] = {}, *psi1[3] = {}, *psi2[3] = {};   // [0]=phi, [1]=phi_old, [2]=phi_nu; [K or L][nB]
    // stage-collapsed ADE of reduced-precision storage: per-material coefficients
    // (boundary_cell_rk's layout, ade_nc doubles each) and the [3][nB] stage values T1, w0, w1
    double *d_ade_rk = nullptr, *d_ade_aux = nullptr;
    int ade_nc = 0;
    // PML (device tables in pml, host copies per axis x, y, z)
    PmlSet pml{};
    PmlSet *d_pml = nullptr;      // device copy of pml (tiled path's cPML pass)
    int *d_pml_cells = nullptr;   // grid-kernel cells (update mask 1) inside a band, raster order
    int n_pml_cells = 0;
    std::vector<int> h_pml_cells; // the same before any source patch took cells out
    std::vector<int> h_pml_band[3];
    std::vector<double> h_pml_kc[3], h_pml_b[3], h_pml_a[3];
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
    // FP16 storage: field scale exponents (value = stored * 2^fexp[f]) and the reduction
    // buffer of the rescale pass (4 maxima + non-finite flag)
    int fexp[4] = {0, 0, 0, 0};
    float *d_slices = nullptr;           // animation planes of p, u, v, w (allocated on first use)
    size_t n_slices = 0;
    unsigned *d_amax = nullptr;
    int tz = 0;                          // z planes per block of the tiled kernel (0: not tuned)
    int nw = 16;                         // waves (rows) per block of the tiled kernel: 16 or 8
    int tx = 64;                         // columns per block of the tiled kernel: 64 or 62
    void *adj = nullptr;                 // adjoint checkpoints and buffers (fvm_adjoint.hip)
    void (*adj_release)(void *) = nullptr;

    void release_adjoint() {
        if (adj && adj_release) adj_release(adj);
        adj = nullptr;
        adj_release = nullptr;
    }
    ~fases_ctx() {
        release_adjoint();
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

// Frees one allocation of the context (nullptr: nothing) after the context's stream has
// drained, and drops it from the context's list and byte count.
int dfree(fases_ctx *ctx, void *p) {
    if (!p) return FASES_OK;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (size_t q = 0; q < ctx->bufs.size(); ++q)
        if (ctx->bufs[q].ptr == p) {
            ctx->dev_bytes -= ctx->bufs[q].bytes;
            ctx->bufs.erase(ctx->bufs.begin() + (long)q);
            HIPCHK(hipFree(p));
            return FASES_OK;
        }
    ctx->err = "dfree: pointer not owned by the context";
    return FASES_E_STATE;
}

// Replaces the buffer in *slot (freed, if any) by a new zeroed one of count elements, so
// repeated set-up calls keep the context's device footprint constant.
template <typename T>
int dalloc_replace(fases_ctx *ctx, T **slot, size_t count) {
    int rc = dfree(ctx, *slot);
    *slot = nullptr;
    if (rc) return rc;
    return dalloc(ctx, slot, count);
}

// Frees every allocation of the context and clears every pointer that named one (a later
// *_replace call must not meet a freed pointer).
void dfree_all(fases_ctx *ctx, void *keep_none = nullptr) {
    (void)keep_none;
    for (auto &b : ctx->bufs) if (b.ptr) (void)hipFree(b.ptr);
    ctx->bufs.clear();
    ctx->dev_bytes = 0;
    for (auto &set : ctx->U)
        for (auto &f : set) f = nullptr;
    ctx->d_upd = nullptr;
    ctx->d_patch = ctx->d_pidx = nullptr;
    for (auto &sh : ctx->shadow) sh = nullptr;
    ctx->d_patch_fm = ctx->d_patch_fe = nullptr;
    ctx->d_patch_edges = ctx->d_patch_eid = nullptr;
    ctx->d_patch_src = nullptr;
    ctx->d_bcell = ctx->d_bmat = nullptr;
    ctx->d_bdist = nullptr;
    ctx->d_bnrm = nullptr;
    ctx->d_Rinf = ctx->d_A = ctx->d_zeta = ctx->d_B = ctx->d_C = ctx->d_alpha = ctx->d_beta = nullptr;
    ctx->d_Keff = ctx->d_Leff = nullptr;
    for (int q = 0; q < 3; ++q) ctx->phi[q] = ctx->psi1[q] = ctx->psi2[q] = nullptr;
    ctx->d_ade_rk = ctx->d_ade_aux = nullptr;
    ctx->pml = PmlSet{};
    ctx->d_pml = nullptr;
    ctx->d_pml_cells = nullptr;
    ctx->d_scell = nullptr;
    ctx->d_sval = ctx->d_sfreq = ctx->d_us = nullptr;
    ctx->d_rcell = nullptr;
    ctx->d_ir = ctx->d_vel = nullptr;
    ctx->d_slices = nullptr;
    ctx->n_slices = 0;
    ctx->d_amax = nullptr;
}

template <typename T>
int upload(fases_ctx *ctx, T **dst, const T *src, size_t count) {
    int rc = dalloc(ctx, dst, count);
    if (rc) return rc;
    if (count) HIPCHK(hipMemcpyAsync(*dst, src, count * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return FASES_OK;
}

// upload() into the buffer in *slot, replacing it (see dalloc_replace).
template <typename T>
int upload_replace(fases_ctx *ctx, T **slot, const T *src, size_t count) {
    int rc = dfree(ctx, *slot);
    *slot = nullptr;
    if (rc) return rc;
    return upload(ctx, slot, src, count);
}

// Coefficients of the stage-collapsed ADE (boundary_cell_rk) for nMat materials, in FP64:
// per real pole a = 1 - dt zeta, per complex pair M = [[m, -n], [n, m]] with
// m = 1 - dt alpha, n = dt beta (the step x' = x + dt (w - zeta x), resp.
// psi1' = psi1 + dt (w - alpha psi1 - beta psi2), psi2' = psi2 + dt (beta psi1 - alpha psi2)).
inline std::vector<double> ade_rk_table(int nMat, int K, int L, const double *Rinf, const double *A,
                                        const double *zeta, const double *B, const double *C, const double *alpha,
                                        const double *beta, double dt) {
    const int nc = 6 * K + 14 * L + 4;
    std::vector<double> t((size_t)nMat * nc, 0.0);
    for (int m = 0; m < nMat; ++m) {
        double *cf = t.data() + (size_t)m * nc;
        double sumA = 0.0, sumAa = 0.0, sumB = 0.0, sumBC1 = 0.0;
        for (int k = 0; k < K; ++k) {
            const double Ak = A[(size_t)m * K + k], a = 1.0 - dt * zeta[(size_t)m * K + k];
            double *q = cf + 6 * k;
            q[0] = Ak;
            q[1] = Ak * a;
            q[2] = Ak * (3.0 + a * a) / 4.0;
            q[3] = 1.0 / 3.0 + a * (3.0 + a * a) / 6.0;
            q[4] = a * a * dt / 6.0;
            q[5] = a * dt / 6.0;
            sumA += Ak;
            sumAa += Ak * a;
        }
        for (int l = 0; l < L; ++l) {
            const double b_ = B[(size_t)m * L + l], c_ = C[(size_t)m * L + l];
            const double mm = 1.0 - dt * alpha[(size_t)m * L + l], nn = dt * beta[(size_t)m * L + l];
            // M, M^2, P2 = (3I + M^2)/4, P3 = I/3 + (2/3) M P2 (row-major 2x2)
            const double M[4] = {mm, -nn, nn, mm};
            const double M2[4] = {mm * mm - nn * nn, -2.0 * mm * nn, 2.0 * mm * nn, mm * mm - nn * nn};
            const double P2[4] = {(3.0 + M2[0]) / 4.0, M2[1] / 4.0, M2[2] / 4.0, (3.0 + M2[3]) / 4.0};
            double P3[4];
            for (int r = 0; r < 2; ++r)
                for (int s = 0; s < 2; ++s)
                    P3[2 * r + s] = (r == s ? 1.0 / 3.0 : 0.0) +
                                    (2.0 / 3.0) * (M[2 * r] * P2[s] + M[2 * r + 1] * P2[2 + s]);
            double *q = cf + 6 * K + 14 * l;
            q[0] = b_;
            q[1] = c_;
            q[2] = b_ * M[0] + c_ * M[2];      // [B C] M
            q[3] = b_ * M[1] + c_ * M[3];
            q[4] = b_ * P2[0] + c_ * P2[2];    // [B C] P2
            q[5] = b_ * P2[1] + c_ * P2[3];
            for (int e = 0; e < 4; ++e) q[6 + e] = P3[e];
            q[10] = dt / 6.0 * M2[0];          // dt/6 M^2 e1
            q[11] = dt / 6.0 * M2[2];
            q[12] = dt / 6.0 * M[0];           // dt/6 M e1
            q[13] = dt / 6.0 * M[2];
            sumB += b_;
            sumBC1 += b_ * mm + c_ * nn;       // [B C] M e1
        }
        double *tl = cf + 6 * K + 14 * L;
        tl[0] = Rinf[m];
        tl[1] = Rinf[m] + dt * (sumA + sumB);
        tl[2] = dt / 4.0 * (sumAa + sumBC1);
        tl[3] = Rinf[m] + dt / 4.0 * (sumA + sumB);
    }
    return t;
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
Scl<R> field_scales(const fases_ctx *ctx) {
    Scl<R> sc;
    for (int f = 0; f < 4; ++f) {
        sc.s[f] = (R)std::ldexp(1.0, ctx->fexp[f]);
        sc.is[f] = (R)std::ldexp(1.0, -ctx->fexp[f]);
    }
    return sc;
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

template <typename S, typename R, int NW, int TX>
int launch_tiled(fases_ctx *ctx, int stage, const FieldSet<S> &fi, const FieldSet<S> &fb, const FieldSet<S> &fo, R rho,
                 R c, R h, R dt, int tz) {
    (void)rho; (void)c; (void)h;
    const TileK<R> K = tile_constants<R>(ctx->prm);
    const dim3 tb(64, NW);
    const int nbx = (ctx->Nx - 6 + TX - 1) / TX, nby = (ctx->Ny - 6 + NW - 1) / NW;
    const int nbz = (ctx->Nz - 6 + tz - 1) / tz;
    const dim3 tg((unsigned)nbx * (unsigned)nby * (unsigned)nbz);     // 1-D: tiles ordered per XCD in the kernel
    const Scl<R> sc = field_scales<R>(ctx);
    const TileFields<S> F{fi.p, fi.u, fi.v, fi.w, fb.p, fb.u, fb.v, fb.w, fo.p, fo.u, fo.v, fo.w};
    if (stage == 0)
        hipLaunchKernelGGL((k_flux_tiled<S, R, 0, NW, TX>), tg, tb, 0, ctx->stream, ctx->Nx, ctx->Ny, ctx->Nz, F,
                           ctx->d_upd, ctx->d_pml, K, dt, sc, tz, nbx, nby);
    else if (stage == 1)
        hipLaunchKernelGGL((k_flux_tiled<S, R, 1, NW, TX>), tg, tb, 0, ctx->stream, ctx->Nx, ctx->Ny, ctx->Nz, F,
                           ctx->d_upd, ctx->d_pml, K, dt, sc, tz, nbx, nby);
    else
        hipLaunchKernelGGL((k_flux_tiled<S, R, 2, NW, TX>), tg, tb, 0, ctx->stream, ctx->Nx, ctx->Ny, ctx->Nz, F,
                           ctx->d_upd, ctx->d_pml, K, dt, sc, tz, nbx, nby);
    if (ctx->n_pml_cells > 0) {       // cPML of the band cells on the stage output
        const dim3 pg((unsigned)((ctx->n_pml_cells + 255) / 256));
        if (stage == 0)
            hipLaunchKernelGGL((k_pml_band<S, R, 0>), pg, dim3(256), 0, ctx->stream, ctx->Nx, ctx->Ny, F,
                               ctx->d_pml_cells, ctx->n_pml_cells, ctx->d_pml, K, dt, sc);
        else if (stage == 1)
            hipLaunchKernelGGL((k_pml_band<S, R, 1>), pg, dim3(256), 0, ctx->stream, ctx->Nx, ctx->Ny, F,
                               ctx->d_pml_cells, ctx->n_pml_cells, ctx->d_pml, K, dt, sc);
        else
            hipLaunchKernelGGL((k_pml_band<S, R, 2>), pg, dim3(256), 0, ctx->stream, ctx->Nx, ctx->Ny, F,
                               ctx->d_pml_cells, ctx->n_pml_cells, ctx->d_pml, K, dt, sc);
    }
    return FASES_OK;
}

bool uses_tiled(const fases_ctx *ctx) {
    const int kv = ctx->prm.kernel_variant;
    return !(kv == 1 || (kv == 0 && ctx->prm.precision == FASES_FP64));
}

// Tiled stage launch with a block height (16 waves: one block per CU and less halo; 8
// waves: three blocks per CU whose barriers do not wait on each other) and width (64
// output columns with edge work, or 62 without).
template <typename S, typename R>
int launch_tiled_nw(fases_ctx *ctx, int stage, const FieldSet<S> &fi, const FieldSet<S> &fb, const FieldSet<S> &fo,
                    R rho, R c, R h, R dt, int nw, int tz, int tx) {
    if (tx == 62)
        return nw == 8 ? launch_tiled<S, R, 8, 62>(ctx, stage, fi, fb, fo, rho, c, h, dt, tz)
                       : launch_tiled<S, R, 16, 62>(ctx, stage, fi, fb, fo, rho, c, h, dt, tz);
    return nw == 8 ? launch_tiled<S, R, 8, 64>(ctx, stage, fi, fb, fo, rho, c, h, dt, tz)
                   : launch_tiled<S, R, 16, 64>(ctx, stage, fi, fb, fo, rho, c, h, dt, tz);
}

// Tile autotuner: times a stage-0 launch (set 0 -> set 1) plus a stage-1 launch (sets
// 0, 1 -> set 2) for each candidate block width, height and z depth, best of four, and
// keeps the fastest. Sets 1 and 2 are rewritten by the next step before they are read and
// neither stage advances the PML states, so tuning changes no state; no candidate changes
// a result. A fixed tile_cols / tile_rows / tile_z restricts the search to that value.
// Depths: 16, 32, 64, 128 and, on a room whose 16-deep tiles do not fill the resident
// block slots of the chip (three 8-row blocks or one 16-row block per CU), the depth at
// which the tile count comes closest to the slot count without exceeding it.
template <typename S, typename R>
int autotune_tile(fases_ctx *ctx) {
    const int fix_tx = ctx->prm.tile_cols, fix_nw = ctx->prm.tile_rows, fix_tz = ctx->prm.tile_z;
    if (fix_tz > 0 && fix_nw > 0 && fix_tx > 0) { ctx->tz = fix_tz; ctx->nw = fix_nw; ctx->tx = fix_tx; return FASES_OK; }
    const fases_params &P = ctx->prm;
    const int nz = ctx->Nz - 6;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, P.device) != hipSuccess) ncu = 0;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    float best = 1e30f;
    int pick = fix_tz > 0 ? fix_tz : 64, pick_nw = fix_nw > 0 ? fix_nw : 16, pick_tx = fix_tx > 0 ? fix_tx : 64;
    for (int tx : {64, 62}) {
        if (fix_tx > 0 && tx != fix_tx) continue;
        for (int nw : {16, 8}) {
            if (fix_nw > 0 && nw != fix_nw) continue;
            std::vector<int> cand{16, 32, 64, 128};
            const long long cols = (long long)((ctx->Nx - 6 + tx - 1) / tx) * ((ctx->Ny - 6 + nw - 1) / nw);
            const long long slots = (long long)ncu * (nw == 8 ? 3 : 1);
            if (cols * ((nz + 15) / 16) < slots) {
                const long long zb = std::max<long long>(1, slots / cols);
                const int tf = std::max(4, (int)((nz + zb - 1) / zb));
                if (tf < 16) cand.insert(cand.begin(), tf);
            }
            for (int tz : cand) {
                if (fix_tz > 0 && tz != fix_tz) continue;
                if (fix_tz <= 0 && tz > 16 && (tz / 2) >= nz) break;   // deeper blocks than the grid add nothing
                float tmin = 1e30f;
                for (int rep = 0; rep < 4; ++rep) {
                    HIPCHK(hipEventRecord(e0, ctx->stream));
                    int rc = launch_tiled_nw<S, R>(ctx, 0, fset<S>(ctx, 0), fset<S>(ctx, 0), fset<S>(ctx, 1), (R)P.rho,
                                                   (R)P.c, (R)P.h, (R)P.dt, nw, tz, tx);
                    if (rc) return rc;
                    rc = launch_tiled_nw<S, R>(ctx, 1, fset<S>(ctx, 1), fset<S>(ctx, 0), fset<S>(ctx, 2), (R)P.rho,
                                               (R)P.c, (R)P.h, (R)P.dt, nw, tz, tx);
                    if (rc) return rc;
                    HIPCHK(hipEventRecord(e1, ctx->stream));
                    HIPCHK(hipEventSynchronize(e1));
                    float ms = 0.f;
                    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                    tmin = std::min(tmin, ms);
                }
                if (tmin < best) { best = tmin; pick = tz; pick_nw = nw; pick_tx = tx; }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(hipGetLastError());
    ctx->tz = pick;
    ctx->nw = pick_nw;
    ctx->tx = pick_tx;
    return FASES_OK;
}

// FP16 scale rule: the largest |value| of a field is stored near 2^kHalfTop, leaving a
// 2^(15 - kHalfTop) headroom to the FP16 maximum (65504) for growth until the next rescale
// and 2^(kHalfTop + 14) below it before values turn subnormal.
constexpr int kHalfTop = 6;

int exponent_for(double amax, int keep) {
    if (!(amax > 0.0)) return keep;
    int e;
    (void)std::frexp(amax, &e);     // amax in [2^(e-1), 2^e)
    return e - kHalfTop;
}

template <typename S>
int rescale_fields(fases_ctx *ctx) {
    if (!ctx->d_amax) {
        int rc = dalloc(ctx, &ctx->d_amax, 5);
        if (rc) return rc;
    }
    hipLaunchKernelGGL((k_fill<unsigned>), dim3(1), dim3(64), 0, ctx->stream, ctx->d_amax, 5, 0u);
    // 1024 blocks: 16 waves per CU stream the set, and 5 atomics per block meet in L2
    const int nb_max = (int)std::min<long long>(1024, (ctx->N + 255) / 256);
    hipLaunchKernelGGL((k_absmax<S>), dim3(nb_max), dim3(256), 0, ctx->stream, fset<S>(ctx, 0), ctx->N, ctx->d_amax);
    const int nb = (int)std::min<long long>(4096, (ctx->N + 255) / 256);
    unsigned hm[5];
    HIPCHK(hipMemcpyAsync(hm, ctx->d_amax, sizeof hm, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (hm[4]) return set_error(ctx, FASES_E_NONFINITE, "FP16 field storage overflowed (non-finite value)");
    Scl<float> fac;
    bool change = false;
    for (int f = 0; f < 4; ++f) {
        float m;
        std::memcpy(&m, &hm[f], sizeof m);
        const int e = exponent_for(std::ldexp((double)m, ctx->fexp[f]), ctx->fexp[f]);
        fac.s[f] = (float)std::ldexp(1.0, ctx->fexp[f] - e);
        fac.is[f] = 1.0f / fac.s[f];
        change |= (e != ctx->fexp[f]);
        ctx->fexp[f] = e;
    }
    if (change)
        for (int q = 0; q < 3; ++q)
            hipLaunchKernelGGL((k_rescale<S>), dim3(nb), dim3(256), 0, ctx->stream, fset<S>(ctx, q), ctx->N, fac);
    return FASES_OK;
}

// Steps (global index) at whose start an FP16 run rescales: 1, 2, 4, ... below the
// period (fields grow fastest from rest), then every period.
bool rescale_due(long long n, int period) {
    if (period <= 0 || n == 0) return false;
    if (n < period) return (n & (n - 1)) == 0;
    return n % period == 0;
}


// The boundary list of a context as the boundary kernels take it.
inline Bnd bnd_of(const fases_ctx *ctx) {
    return Bnd{ctx->nB, ctx->d_bcell, ctx->d_bmat, ctx->d_bdist, ctx->d_bnrm, ctx->prm.h, ctx->prm.tol};
}

// Stage-collapsed ADE parameters and buffers of a context (reduced-precision storage):
// x0 = phi[1] (the step's start state), x3 -> phi[0].
inline AdeRk ade_rk_of(const fases_ctx *ctx) {
    return AdeRk{ctx->K, ctx->L, ctx->ade_nc, ctx->d_ade_rk, ctx->d_Keff, ctx->d_Leff, 2.0 * ctx->prm.dt / 3.0};
}
inline AdeRkBufs ade_rk_bufs(const fases_ctx *ctx) {
    return AdeRkBufs{ctx->phi[1], ctx->psi1[1], ctx->psi2[1], ctx->phi[0], ctx->psi1[0], ctx->psi2[0], ctx->d_ade_aux};
}

// The three stages of step n (flux -> source patch -> sources -> boundary each) on
// explicit field sets: sets[0] = U^n (also the RK base), stage 0 writes sets[1], stage 1
// sets[2], stage 2 sets[3] (= sets[1] in the plain loop, whose caller then rotates);
// sh[] are the FP64 patch shadows of the same sets. ADE buffers rotate in the context.
template <typename S, typename R>
int launch_stages(fases_ctx *ctx, long long n, const FieldSet<S> sets[4], double *const sh[4],
                  std::vector<hipEvent_t> *kev) {
    int rc = FASES_OK;
    Grid g{ctx->Nx, ctx->Ny, ctx->Nz, (long long)ctx->Nx * ctx->Ny};
    const fases_params &P = ctx->prm;
    const R rho = (R)P.rho, c = (R)P.c, h = (R)P.h, dt = (R)P.dt;
    const Scl<R> sc = field_scales<R>(ctx);
    const Scl<double> scd = field_scales<double>(ctx);
    dim3 blk(64, 4, 1);
    dim3 grd((ctx->Nx - 2 * NG + 63) / 64, (ctx->Ny - 2 * NG + 3) / 4, ctx->Nz - 2 * NG);
    const double As = 6.0 * P.h * P.h;
    const int sb = 256;
    Ade ade{ctx->K, ctx->L, ctx->d_Rinf, ctx->d_A, ctx->d_zeta, ctx->d_B, ctx->d_C, ctx->d_alpha, ctx->d_beta,
            ctx->d_Keff, ctx->d_Leff};
    // Geometry stays FP64 for every field precision: a blend weight or normal rounded
    // once to FP32 is a bias that every step repeats (measured: FP32-room IR error x3.5).
    const Bnd bd = bnd_of(ctx);
    // (field set in, set out, RK weight, source sample, ADE a,b,c, ADE buffers left/mid/right)
    struct StageCfg { int in, out; double w; long long nn; double a, b, c; int l, mid, r; };
    const StageCfg cfg[3] = {
        {0, 1, 1.0, 2 * n, 0.0, 1.0, 1.0, 0, 1, 1},
        {1, 2, 0.25, 2 * n + 2, 3.0, 1.0, 4.0, 2, 1, 0},
        {2, 3, 2.0 / 3.0, 2 * n + 1, 1.0, 2.0, 3.0, 0, 1, 2},
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
        if (!uses_tiled(ctx)) {
            const GridIO<S, R> gio{sets[q.in], sets[0], sets[q.out], sc};
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
            rc = launch_tiled_nw<S, R>(ctx, s, sets[q.in], sets[0], sets[q.out], rho, c, h, dt, ctx->nw, ctx->tz, ctx->tx);
            if (rc) return rc;
        }
        if (kev) HIPCHK(hipEventRecord(e1, ctx->stream));
        if (ctx->n_patch > 0) {
            // FP64 update of the patch: every face flux once, then the stage update per cell
            const long long nf = 3LL * ctx->n_patch + ctx->n_patch_edge;
            const dim3 fg((unsigned)((nf + 63) / 64)), pg((ctx->n_patch + 63) / 64), pb(64);
            const PatchIO<S> pio{sets[q.in], sets[0], sets[q.out], ctx->d_pidx, sh[q.in], sh[0], sh[q.out],
                                 ctx->n_patch, scd};
            hipLaunchKernelGGL((k_flux_patch_faces<S>), fg, pb, 0, ctx->stream, g, pio, ctx->d_patch,
                               ctx->d_patch_edges, ctx->n_patch_edge, P.rho, P.c, ctx->d_patch_fm, ctx->d_patch_fe);
            const PatchSrc ps{ctx->n_src > 0 ? ctx->d_patch_src : nullptr, ctx->d_sval, ctx->d_us,
                              ctx->src_opt == FASES_SRC_SINE ? ctx->n_samp : 0LL, q.nn, q.w, P.dt, P.c, P.rho, As, P.h};
            if (s == 0)
                hipLaunchKernelGGL((k_flux_patch_update<S, 0>), pg, pb, 0, ctx->stream, g, pio, ctx->d_patch, ctx->pml,
                                   P.h, P.dt, ctx->d_patch_fm, ctx->d_patch_fe, ctx->d_patch_eid, ps);
            else if (s == 1)
                hipLaunchKernelGGL((k_flux_patch_update<S, 1>), pg, pb, 0, ctx->stream, g, pio, ctx->d_patch, ctx->pml,
                                   P.h, P.dt, ctx->d_patch_fm, ctx->d_patch_fe, ctx->d_patch_eid, ps);
            else
                hipLaunchKernelGGL((k_flux_patch_update<S, 2>), pg, pb, 0, ctx->stream, g, pio, ctx->d_patch, ctx->pml,
                                   P.h, P.dt, ctx->d_patch_fm, ctx->d_patch_fe, ctx->d_patch_eid, ps);
        }
        if (ctx->n_src > 0 && !(ctx->n_patch > 0 && ctx->d_patch_src))
            hipLaunchKernelGGL((k_sources<S, R>), dim3((ctx->n_src + sb - 1) / sb), dim3(sb), 0, ctx->stream,
                               sets[q.out].p, ctx->d_scell, ctx->d_sval, ctx->d_sfreq, ctx->d_us,
                               ctx->src_opt == FASES_SRC_SINE ? ctx->n_samp : 0LL, ctx->n_src, ctx->src_opt, q.nn, q.w, P.dt, P.c, P.rho, As, P.h,
                               ctx->n_patch > 0 ? ctx->d_pidx : nullptr, sh[q.out], sc.s[0], sc.is[0]);
        if (ctx->nB > 0 && std::is_same<S, double>::value)
            hipLaunchKernelGGL((k_boundary<S>), dim3((ctx->nB + sb - 1) / sb), dim3(sb), 0, ctx->stream, g,
                               sets[q.in], sets[q.out], bd, ade, P.rho, P.c, P.dt, q.a, q.b, q.c,
                               AdeBufs{ctx->phi[q.l], ctx->phi[q.mid], ctx->phi[q.r], ctx->psi1[q.l],
                                       ctx->psi1[q.mid], ctx->psi1[q.r], ctx->psi2[q.l], ctx->psi2[q.mid],
                                       ctx->psi2[q.r]},
                               scd);
        else if (ctx->nB > 0) {
            // reduced storage: the stage-collapsed ADE on the step's start state phi[1]
            // (x0), end state into phi[0] (rotated into phi[1] after the step)
            const AdeRk rk = ade_rk_of(ctx);
            const AdeRkBufs rb = ade_rk_bufs(ctx);
            const double Z0 = P.rho * P.c, iz = 1.0 / (2.0 * Z0);
            const dim3 bg((unsigned)((ctx->nB + sb - 1) / sb));
            if (s == 0)
                hipLaunchKernelGGL((k_boundary_rk<S, 0>), bg, dim3(sb), 0, ctx->stream, g, sets[q.in],
                                   sets[q.out], bd, rk, Z0, iz, rb, scd);
            else if (s == 1)
                hipLaunchKernelGGL((k_boundary_rk<S, 1>), bg, dim3(sb), 0, ctx->stream, g, sets[q.in],
                                   sets[q.out], bd, rk, Z0, iz, rb, scd);
            else
                hipLaunchKernelGGL((k_boundary_rk<S, 2>), bg, dim3(sb), 0, ctx->stream, g, sets[q.in],
                                   sets[q.out], bd, rk, Z0, iz, rb, scd);
        }
    }
    return FASES_OK;
}

// One plain step: the three stages on the context's sets, then the rotation.
template <typename S, typename R>
int launch_step(fases_ctx *ctx, long long n, std::vector<hipEvent_t> *kev) {
    const FieldSet<S> sets[4] = {fset<S>(ctx, 0), fset<S>(ctx, 1), fset<S>(ctx, 2), fset<S>(ctx, 1)};
    double *const sh[4] = {ctx->shadow[0], ctx->shadow[1], ctx->shadow[2], ctx->shadow[1]};
    int rc = launch_stages<S, R>(ctx, n, sets, sh, kev);
    if (rc) return rc;
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
                       ctx->d_rcell, ctx->n_rec, ctx->d_ir, ctx->d_vel, stride, n, field_scales<double>(ctx));
    return FASES_OK;
}

// FP16 rescale pass: measure max |value| per field on the current set, pick new
// exponents, multiply the stored values of all three sets by 2^(old - new). Fails with
// FASES_E_NONFINITE when a stored value has overflowed.
// Per-run set-up shared by the single-context and batched loops: sine sample tables
// long enough for the run, a waveform-length check, receiver buffers of nsteps rows.
// Chebyshev radius of the FP64 region around point sources under reduced-precision
// storage (forward shadows and adjoint): fases_params.source_patch, 12 when 0.
constexpr int kDefaultSourcePatch = 12;
inline int source_patch_radius(const fases_ctx *ctx) {
    return ctx->prm.source_patch > 0 ? ctx->prm.source_patch : kDefaultSourcePatch;
}

inline int prepare_run(fases_ctx *ctx, long long nsteps) {
    if (ctx->n_src > 0 && ctx->src_opt == FASES_SRC_SINE && 2 * (ctx->step + nsteps) + 1 > ctx->n_samp) {
        // per-source sine samples sin(2 pi f nn dt/2), tabulated on the host (same libm
        // as a CPU evaluation of the reference formula)
        const long long ns = 2 * (ctx->step + nsteps) + 1;
        std::vector<double> tab((size_t)ctx->n_src * ns);
        for (int q = 0; q < ctx->n_src; ++q)
            for (long long t = 0; t < ns; ++t)
                tab[(size_t)q * ns + t] = std::sin(2.0 * M_PI * ctx->h_sfreq[(size_t)q] * (double)t * (0.5 * ctx->prm.dt));
        int rc = upload_replace(ctx, &ctx->d_us, tab.data(), tab.size());
        if (rc) return rc;
        ctx->n_samp = ns;
    }
    if (ctx->n_rec > 0 && nsteps > ctx->ir_cap) {
        // grow the device receiver buffers
        int rc = dalloc_replace(ctx, &ctx->d_ir, (size_t)ctx->n_rec * nsteps);
        if (rc) return rc;
        rc = dalloc_replace(ctx, &ctx->d_vel, (size_t)ctx->n_rec * nsteps * 3);
        if (rc) return rc;
        ctx->ir_cap = nsteps;
    }
    if (ctx->n_src > 0 && ctx->src_opt != FASES_SRC_SINE && 2 * (ctx->step + nsteps) + 1 > ctx->n_samp)
        return set_error(ctx, FASES_E_ARG, "source waveform shorter than 2*(steps)+1 samples");
    return FASES_OK;
}

}  // namespace
