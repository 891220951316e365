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

#include "fvm_ctx.hpp"

#include <hip/hip_runtime.h>
#include <cmath>
#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

namespace {


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

// IEEE binary16 encoding of a double on the host, one rounding (to nearest even).
uint16_t half_bits(double x) {
    const uint16_t sign = std::signbit(x) ? 0x8000u : 0u;
    const double a = std::fabs(x);
    if (std::isnan(a)) return (uint16_t)(sign | 0x7E00u);
    if (a >= 65520.0) return (uint16_t)(sign | 0x7C00u);      // rounds past 65504
    int e;
    (void)std::frexp(a, &e);                                   // a in [2^(e-1), 2^e)
    const int q = std::max(e - 11, -24);                      // quantum: 11 significant bits, 2^-24 floor
    const double v = std::ldexp(std::nearbyint(std::ldexp(a, -q)), q);
    if (v == 0.0) return sign;
    if (v < 6.103515625e-05) return (uint16_t)(sign | (uint16_t)std::ldexp(v, 24));   // subnormal
    int ev;
    const double mv = std::frexp(v, &ev);                     // v = mv 2^ev, mv in [0.5, 1)
    return (uint16_t)(sign | (uint16_t)((ev + 14) << 10) | (uint16_t)((mv * 2.0 - 1.0) * 1024.0));
}

float half_value(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
    float v;
    if (e == 0) v = std::ldexp((float)m, -24);
    else if (e == 31) v = m ? NAN : INFINITY;
    else v = std::ldexp((float)(m | 0x400u), (int)e - 25);
    uint32_t b;
    std::memcpy(&b, &v, 4);
    b |= sign;
    std::memcpy(&v, &b, 4);
    return v;
}

constexpr long long kKernelEventEvery = 16;

template <typename S, typename R>
int run_impl(fases_ctx *ctx, long long nsteps, double *ir, double *vel, double *elapsed_ms, double *kernel_ms) {
    const long long need = (long long)ctx->n_rec * nsteps;
    {
        int rc = prepare_run(ctx, nsteps);
        if (rc) return rc;
    }
    if (uses_tiled(ctx) && ctx->tz == 0) {
        int rc = autotune_tile<S, R>(ctx);
        if (rc) return rc;
    }
    hipEvent_t ev[2];
    for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    // flux-stage events on every kKernelEventEvery-th step only: an event record is a queue
    // barrier with a cache release (~5 us on the 512x384x256 room), which on every step
    // would lengthen the run it measures by about 1.4 %
    std::vector<hipEvent_t> kev;
    long long sampled = 0;
    if (kernel_ms) kev.reserve((size_t)(nsteps / kKernelEventEvery + 1) * 6);
    HIPCHK(hipEventRecord(ev[0], ctx->stream));
    for (long long q = 0; q < nsteps; ++q) {
        if (std::is_same<S, __half>::value && rescale_due(ctx->step, ctx->prm.fp16_rescale_every)) {
            int rc = rescale_fields<S>(ctx);
            if (rc) return rc;
        }
        const bool sample = kernel_ms && q % kKernelEventEvery == 0;
        sampled += sample;
        int rc = launch_step<S, R>(ctx, ctx->step, sample ? &kev : nullptr);
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
        *kernel_ms = sampled > 0 ? tot * (double)nsteps / (double)sampled : 0.0;   // scaled to all steps
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

namespace {

// Three axis planes of one field into out = [YZ at x = i0 | XZ at y = j0 | XY at z = k0],
// each x-fastest over its two axes (the order main3D.cu:1326-1334 copies them in), as
// float32 (storage value times its power-of-two scale).
template <typename S>
__global__ void __launch_bounds__(256)
k_slices(const S *__restrict__ f, float scale, int Nx, int Ny, int Nz, int i0, int j0, int k0, float *__restrict__ out) {
    const long long nyz = (long long)Ny * Nz, nxz = (long long)Nx * Nz, nxy = (long long)Nx * Ny;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nyz + nxz + nxy) return;
    long long c;
    if (t < nyz) {
        c = i0 + (long long)Nx * t;                                  // t = j + Ny k
    } else if (t < nyz + nxz) {
        const long long q = t - nyz, i = q % Nx, k = q / Nx;
        c = i + Nx * (j0 + (long long)Ny * k);
    } else {
        c = (t - nyz - nxz) + nxy * k0;
    }
    out[t] = (float)f[c] * scale;
}
template <>
__global__ void __launch_bounds__(256)
k_slices<__half>(const __half *__restrict__ f, float scale, int Nx, int Ny, int Nz, int i0, int j0, int k0,
                 float *__restrict__ out) {
    const long long nyz = (long long)Ny * Nz, nxz = (long long)Nx * Nz, nxy = (long long)Nx * Ny;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nyz + nxz + nxy) return;
    long long c;
    if (t < nyz) {
        c = i0 + (long long)Nx * t;
    } else if (t < nyz + nxz) {
        const long long q = t - nyz, i = q % Nx, k = q / Nx;
        c = i + Nx * (j0 + (long long)Ny * k);
    } else {
        c = (t - nyz - nxz) + nxy * k0;
    }
    out[t] = __half2float(f[c]) * scale;
}

// Streaming copy of n 16-byte words, grid-stride with non-temporal loads and stores:
// the reference point for the HBM roofline (every byte read once and written once).
// Swept on the MI355X over 1-32 blocks of 256 per CU and 1-8 words in flight per
// thread: one word per thread and 3-4 blocks per CU reach ~5.65 TB/s on 2 GiB, deeper
// unrolling or more blocks 4.3-5.0 TB/s.
typedef unsigned copy_word __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_copy16(const copy_word *__restrict__ a, copy_word *__restrict__ b,
                                                long long n) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(&a[i]), &b[i]);
}

}  // namespace

extern "C" {

int fases_hbm_copy(int device, int64_t bytes, int reps, double *gbs) {
    if (!gbs || bytes < (1 << 20) || reps < 1) return set_error(nullptr, FASES_E_ARG, "fases_hbm_copy: bad argument");
    *gbs = 0.0;
    const long long n = bytes / 16;
    hipError_t e = hipSetDevice(device);
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipStream_t st = nullptr;
    float best = 1e30f;
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&a, (size_t)n * 16);
    if (e == hipSuccess) e = hipMalloc(&b, (size_t)n * 16);
    if (e == hipSuccess) e = hipMemsetAsync(a, 1, (size_t)n * 16, st);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    int dev_cus = 256;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, device);
    const unsigned grid = (unsigned)std::max(1, dev_cus) * 4u;
    for (int r = 0; r <= reps && e == hipSuccess; ++r) {     // launch 0 warms up
        e = hipEventRecord(e0, st);
        hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, st, (const copy_word *)a, (copy_word *)b, n);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess && r > 0) best = std::min(best, ms);
    }
    if (e == hipSuccess) *gbs = 2.0 * (double)n * 16.0 / (best * 1e-3) / 1e9;
    const std::string msg = (e == hipSuccess) ? std::string() : std::string("fases_hbm_copy: ") + hipGetErrorString(e);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (st) (void)hipStreamDestroy(st);
    return e == hipSuccess ? FASES_OK : set_error(nullptr, FASES_E_HIP, msg);
}

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
    if (p->pml_cells < 1) return set_error(nullptr, FASES_E_ARG, "fases_create: pml_cells must be >= 1");
    if (p->tile_rows != 0 && p->tile_rows != 8 && p->tile_rows != 16)
        return set_error(nullptr, FASES_E_ARG, "fases_create: tile_rows must be 0, 8 or 16");
    if (p->tile_cols != 0 && p->tile_cols != 62 && p->tile_cols != 64)
        return set_error(nullptr, FASES_E_ARG, "fases_create: tile_cols must be 0, 62 or 64");
    if (p->source_patch < 0 || p->source_patch > 64)
        return set_error(nullptr, FASES_E_ARG, "fases_create: source_patch must be in [0, 64]");
    if (p->tile_z < 0) return set_error(nullptr, FASES_E_ARG, "fases_create: tile_z must be >= 0");
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
    ctx->release_adjoint();
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
    ctx->n_patch_edge = 0;
    ctx->h_patch.clear();
    for (auto &sh : ctx->shadow) sh = nullptr;
    // boundary list in bmap order: cells in [3, N-3)^3 with oid in [1, 254]
    // (functions3D.cuh:466-479, filterABC.h:95-108): cell, material word, signed distance
    // and normal; the probe geometry follows from these (probe_geo, on the device per
    // stage and here once for the centred-probe flag).
    std::vector<int> bcell, bmat;
    std::vector<float> bdist;
    std::vector<double> bnrm;
    const double h = P.h, tol = P.tol;
    for (long long C = 0; C < N; ++C) {
        if (oid[C] == 0 || oid[C] == 0xFF) continue;
        const int i = (int)(C % Nx), j = (int)((C / Nx) % Ny), k = (int)(C / nxy);
        if (i < NG || i >= Nx - NG || j < NG || j >= Ny - NG || k < NG || k >= Nz - NG) continue;
        const double s = S[C];
        // normal n = -grad S/|grad S| (central differences)
        const double gx = ((double)S[C + 1] - (double)S[C - 1]) * (0.5 / h);
        const double gy = ((double)S[C + Nx] - (double)S[C - Nx]) * (0.5 / h);
        const double gz = ((double)S[C + nxy] - (double)S[C - nxy]) * (0.5 / h);
        double gn = std::sqrt(gx * gx + gy * gy + gz * gz);
        gn = std::fmax(gn, 1e-12);
        const double nx = -gx / gn, ny = -gy / gn, nz = -gz / gn;
        const ProbeGeo q = probe_geo(i, j, k, Nx, Ny, Nz, s, nx, ny, nz, h, tol);
        bcell.push_back((int)C);
        const int centred = (q.sx == 0.0 && q.sy == 0.0 && q.sz == 0.0) ? 0x100 : 0;
        bmat.push_back(((int)oid[C] - 1) | centred);
        bdist.push_back(S[C]);
        for (double v : {nx, ny, nz}) bnrm.push_back(v);
    }
    ctx->nB = (int)bcell.size();
    {   // cell-major normals -> entry-major [3][nB]
        std::vector<double> t(bnrm.size());
        for (size_t b = 0; b < bcell.size(); ++b)
            for (int a = 0; a < 3; ++a) t[(size_t)a * bcell.size() + b] = bnrm[b * 3 + a];
        bnrm.swap(t);
    }
    if ((rc = upload(ctx, &ctx->d_bcell, bcell.data(), bcell.size()))) return rc;
    if ((rc = upload(ctx, &ctx->d_bmat, bmat.data(), bmat.size()))) return rc;
    if ((rc = upload(ctx, &ctx->d_bdist, bdist.data(), bdist.size()))) return rc;
    if ((rc = upload(ctx, &ctx->d_bnrm, bnrm.data(), bnrm.size()))) return rc;
    ctx->h_bcell = bcell;
    ctx->h_bmat = bmat;
    ctx->h_bdist = bdist;
    ctx->h_bnrm = bnrm;
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
            const int a3 = (&ax == &ctx->pml.x) ? 0 : (&ax == &ctx->pml.y ? 1 : 2);
            ctx->h_pml_band[a3] = band;
            ctx->h_pml_kc[a3] = kc;
            ctx->h_pml_b[a3] = bb;
            ctx->h_pml_a[a3] = aa;
            int r;
            int *db; double *dk, *dbb, *da;
            if ((r = upload(ctx, &db, band.data(), band.size()))) return r;
            if ((r = upload(ctx, &dk, kc.data(), kc.size()))) return r;
            if ((r = upload(ctx, &dbb, bb.data(), bb.size()))) return r;
            if ((r = upload(ctx, &da, aa.data(), aa.size()))) return r;
            ax.band = db; ax.kc = dk; ax.b = dbb; ax.a = da;
            HIPCHK(hipStreamSynchronize(ctx->stream));   // the next axis refills the host tables
            return FASES_OK;
        };
        if ((rc = axis(ctx->pml.x, Nx, key[3], key[2]))) return rc;   // L = x-, R = x+
        if ((rc = axis(ctx->pml.y, Ny, key[5], key[4]))) return rc;   // B = y-, F = y+
        if ((rc = axis(ctx->pml.z, Nz, key[1], key[0]))) return rc;   // D = z-, U = z+
        const int nn[3] = {Nx, Ny, Nz};
        const float klo[3] = {key[3], key[5], key[1]}, khi[3] = {key[2], key[4], key[0]};
        for (int a = 0; a < 3; ++a) {
            ctx->pml.clear0[a] = (klo[a] > 0.f) ? NG + P.pml_cells : NG;
            ctx->pml.clear1[a] = (khi[a] > 0.f) ? nn[a] - NG - P.pml_cells : nn[a] - NG;
        }
        const size_t d2 = 2 * (size_t)P.pml_cells;
        const size_t nx_t = ctx->pml.on ? d2 * Ny * Nz : 1, ny_t = ctx->pml.on ? (size_t)Nx * d2 * Nz : 1,
                     nz_t = ctx->pml.on ? (size_t)Nx * Ny * d2 : 1;
        if ((rc = dalloc(ctx, &ctx->pml.psi_px, nx_t))) return rc;
        if ((rc = dalloc(ctx, &ctx->pml.psi_ux, nx_t))) return rc;
        if ((rc = dalloc(ctx, &ctx->pml.psi_py, ny_t))) return rc;
        if ((rc = dalloc(ctx, &ctx->pml.psi_vy, ny_t))) return rc;
        if ((rc = dalloc(ctx, &ctx->pml.psi_pz, nz_t))) return rc;
        if ((rc = dalloc(ctx, &ctx->pml.psi_wz, nz_t))) return rc;
        // device copy for the cPML kernels
        if ((rc = upload(ctx, &ctx->d_pml, &ctx->pml, 1))) return rc;
        // packed list of the updated cells inside any band (raster order), for the tiled
        // path's separate cPML pass
        std::vector<int> bandc;
        if (ctx->pml.on) {
            const auto &bxa = ctx->h_pml_band[0], &bya = ctx->h_pml_band[1], &bza = ctx->h_pml_band[2];
            for (int k = 0; k < Nz; ++k)
                for (int j = 0; j < Ny; ++j)
                    for (int i = 0; i < Nx; ++i) {
                        const long long C = i + (long long)Nx * (j + (long long)Ny * k);
                        if (ctx->h_upd[(size_t)C] == 1 && (bxa[(size_t)i] >= 0 || bya[(size_t)j] >= 0 || bza[(size_t)k] >= 0))
                            bandc.push_back((int)C);
                    }
        }
        ctx->n_pml_cells = (int)bandc.size();
        ctx->d_pml_cells = nullptr;
        if (!bandc.empty() && (rc = upload(ctx, &ctx->d_pml_cells, bandc.data(), bandc.size()))) return rc;
        HIPCHK(hipStreamSynchronize(ctx->stream));   // bandc is moved below
        ctx->h_pml_cells = std::move(bandc);
    }
    ctx->n_src = 0; ctx->n_rec = 0; ctx->ir_cap = 0; ctx->step = 0;
    ctx->tz = 0;
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
    if ((rc = upload_replace(ctx, &ctx->d_Rinf, Rinf, (size_t)nMat))) return rc;
    if ((rc = upload_replace(ctx, &ctx->d_A, A, (size_t)nMat * K))) return rc;
    if ((rc = upload_replace(ctx, &ctx->d_zeta, zeta, (size_t)nMat * K))) return rc;
    if ((rc = upload_replace(ctx, &ctx->d_B, B, (size_t)nMat * L))) return rc;
    if ((rc = upload_replace(ctx, &ctx->d_C, C, (size_t)nMat * L))) return rc;
    if ((rc = upload_replace(ctx, &ctx->d_alpha, alpha, (size_t)nMat * L))) return rc;
    if ((rc = upload_replace(ctx, &ctx->d_beta, beta, (size_t)nMat * L))) return rc;
    if ((rc = upload_replace(ctx, &ctx->d_Keff, Ke.data(), Ke.size()))) return rc;
    if ((rc = upload_replace(ctx, &ctx->d_Leff, Le.data(), Le.size()))) return rc;
    for (int q = 0; q < 3; ++q) {
        if ((rc = dalloc_replace(ctx, &ctx->phi[q], (size_t)ctx->nB * K))) return rc;
        if ((rc = dalloc_replace(ctx, &ctx->psi1[q], (size_t)ctx->nB * L))) return rc;
        if ((rc = dalloc_replace(ctx, &ctx->psi2[q], (size_t)ctx->nB * L))) return rc;
    }
    if (ctx->prm.precision != FASES_FP64) {       // stage-collapsed ADE (boundary_cell_rk)
        const std::vector<double> rk = ade_rk_table(nMat, K, L, Rinf, A, zeta, B, C, alpha, beta, ctx->prm.dt);
        ctx->ade_nc = 6 * K + 14 * L + 4;
        if ((rc = upload_replace(ctx, &ctx->d_ade_rk, rk.data(), rk.size()))) return rc;
        if ((rc = dalloc_replace(ctx, &ctx->d_ade_aux, (size_t)ctx->nB * 3))) return rc;
        HIPCHK(hipStreamSynchronize(ctx->stream));   // rk is a local
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
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // the adjoint's source-region maps are built from the source cells: rebuild them at
    // the next fases_adjoint_forward
    ctx->release_adjoint();
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
    // Reduced-precision storage: cells within the source radius (Chebyshev, default 12,
    // fases_params.source_patch) of a source keep FP64 shadows of all three field sets and
    // are updated in FP64 arithmetic. Inside a
    // step the source cell swings by the full per-stage injection (far above the local
    // pressure) and the RK combination cancels it; FP32 storage or arithmetic in that
    // neighbourhood seeds errors that propagate to every receiver (measured: IR rel L2
    // 1.7e-4 far from the source with no patch, 2e-6 with radius 12).
    if (ctx->prm.precision != FASES_FP64) {
        const int rad = source_patch_radius(ctx);
        const int Nx = ctx->Nx, Ny = ctx->Ny, Nz = ctx->Nz;
        const long long nxy = (long long)Nx * Ny;
        for (auto &v : ctx->h_upd) if (v == 2) v = 1;
        std::vector<int> patch;
        std::vector<uint8_t> is_bnd((size_t)ctx->N, 0);
        for (int c0 : ctx->h_bcell) is_bnd[(size_t)c0] = 1;
        for (int c0 : mc) {
            const int si = c0 % Nx, sj = (int)((c0 / Nx) % Ny), sk = (int)(c0 / nxy);
            for (int k = std::max(0, sk - rad); k <= std::min(Nz - 1, sk + rad); ++k)
                for (int j = std::max(0, sj - rad); j <= std::min(Ny - 1, sj + rad); ++j)
                    for (int i = std::max(0, si - rad); i <= std::min(Nx - 1, si + rad); ++i) {
                        const long long C = i + (long long)Nx * j + nxy * k;
                        // boundary cells stay out: the boundary kernel blends their stored values
                        if (ctx->h_upd[(size_t)C] == 1 && !is_bnd[(size_t)C]) {
                            ctx->h_upd[(size_t)C] = 2;
                            patch.push_back((int)C);
                        }
                    }
        }
        HIPCHK(hipMemcpyAsync(ctx->d_upd, ctx->h_upd.data(), ctx->h_upd.size(), hipMemcpyHostToDevice, ctx->stream));
        if (!ctx->h_pml_cells.empty()) {
            // the cPML pass covers the band's grid-kernel cells only: the patch update
            // applies the cPML of its own cells in FP64
            std::vector<int> bandc;
            for (int c0 : ctx->h_pml_cells)
                if (ctx->h_upd[(size_t)c0] == 1) bandc.push_back(c0);
            if ((rc = upload_replace(ctx, &ctx->d_pml_cells, bandc.data(), bandc.size()))) return rc;
            ctx->n_pml_cells = (int)bandc.size();
            HIPCHK(hipStreamSynchronize(ctx->stream));   // bandc is a local
        }
        ctx->n_patch = (int)patch.size();
        ctx->h_patch = patch;
        if (patch.empty()) {                       // no source any more: drop the old patch
            if ((rc = dfree(ctx, ctx->d_patch))) return rc;
            ctx->d_patch = nullptr;
            for (void *pp : {(void *)ctx->d_patch_fm, (void *)ctx->d_patch_fe, (void *)ctx->d_patch_edges,
                             (void *)ctx->d_patch_eid, (void *)ctx->d_patch_src})
                if ((rc = dfree(ctx, pp))) return rc;
            ctx->d_patch_fm = ctx->d_patch_fe = nullptr;
            ctx->d_patch_edges = ctx->d_patch_eid = ctx->d_patch_src = nullptr;
            ctx->n_patch_edge = 0;
            for (auto &sh : ctx->shadow) {
                if ((rc = dfree(ctx, sh))) return rc;
                sh = nullptr;
            }
        } else {
            if (!ctx->d_pidx && (rc = dalloc(ctx, &ctx->d_pidx, (size_t)ctx->N))) return rc;
            std::vector<int> pidx((size_t)ctx->N, -1);
            for (size_t q = 0; q < patch.size(); ++q) pidx[(size_t)patch[q]] = (int)q;
            HIPCHK(hipMemcpyAsync(ctx->d_pidx, pidx.data(), pidx.size() * sizeof(int), hipMemcpyHostToDevice,
                                  ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));   // pidx ends with this block
            if ((rc = upload_replace(ctx, &ctx->d_patch, patch.data(), patch.size()))) return rc;
            {   // face-flux buffers of the two-launch update: the edge items are the (axis, cell)
                // pairs whose +1 neighbour is no patch cell (their face +1/2 has no owner)
                const int np = (int)patch.size();
                std::vector<int> edges, eid((size_t)3 * np, -1);
                for (int a = 0; a < 3; ++a) {
                    const long long st = a == 0 ? 1LL : (a == 1 ? (long long)Nx : nxy);
                    for (int q = 0; q < np; ++q)
                        if (pidx[(size_t)(patch[(size_t)q] + st)] < 0) {
                            eid[(size_t)a * np + q] = (int)edges.size();
                            edges.push_back(a * np + q);
                        }
                }
                ctx->n_patch_edge = (int)edges.size();
                if ((rc = dalloc_replace(ctx, &ctx->d_patch_fm, (size_t)6 * np))) return rc;
                if ((rc = dalloc_replace(ctx, &ctx->d_patch_fe, std::max<size_t>(1, 2 * edges.size())))) return rc;
                if ((rc = upload_replace(ctx, &ctx->d_patch_edges, edges.data(), edges.size()))) return rc;
                if ((rc = upload_replace(ctx, &ctx->d_patch_eid, eid.data(), eid.size()))) return rc;
                // every source in the patch: the patch update injects them (no k_sources launch)
                std::vector<int> psrc((size_t)np, -1);
                bool all_in = !mc.empty();
                for (size_t q = 0; q < mc.size(); ++q) {
                    const int w = pidx[(size_t)mc[q]];
                    if (w < 0) all_in = false;
                    else psrc[(size_t)w] = (int)q;
                }
                if (all_in) {
                    if ((rc = upload_replace(ctx, &ctx->d_patch_src, psrc.data(), psrc.size()))) return rc;
                } else {
                    if ((rc = dfree(ctx, ctx->d_patch_src))) return rc;
                    ctx->d_patch_src = nullptr;
                }
                HIPCHK(hipStreamSynchronize(ctx->stream));   // edges, eid, psrc are locals
            }
            for (int q = 0; q < 3; ++q) {
                if ((rc = dalloc_replace(ctx, &ctx->shadow[q], (size_t)4 * patch.size()))) return rc;
                if (ctx->esz == 4)
                    hipLaunchKernelGGL((k_patch_gather<float>), dim3((ctx->n_patch + 63) / 64), dim3(64), 0,
                                       ctx->stream, fset<float>(ctx, q), ctx->d_patch, ctx->n_patch, ctx->shadow[q],
                                       field_scales<double>(ctx));
                else
                    hipLaunchKernelGGL((k_patch_gather<__half>), dim3((ctx->n_patch + 63) / 64), dim3(64), 0,
                                       ctx->stream, fset<__half>(ctx, q), ctx->d_patch, ctx->n_patch, ctx->shadow[q],
                                       field_scales<double>(ctx));
            }
        }
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    ctx->n_src = (int)mc.size();
    ctx->src_opt = src_opt;
    ctx->n_samp = (src_opt == FASES_SRC_SINE) ? 0 : n_samp;
    ctx->h_sfreq = mf;
    if ((rc = upload_replace(ctx, &ctx->d_scell, mc.data(), mc.size()))) return rc;
    if ((rc = upload_replace(ctx, &ctx->d_sval, ms.data(), ms.size()))) return rc;
    if ((rc = upload_replace(ctx, &ctx->d_sfreq, mf.data(), mf.size()))) return rc;
    if (src_opt != FASES_SRC_SINE && n > 0) {
        if ((rc = upload_replace(ctx, &ctx->d_us, u_s, (size_t)n_samp))) return rc;
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
    if ((rc = upload_replace(ctx, &ctx->d_rcell, rc_.data(), rc_.size()))) return rc;
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
        } else if (ctx->esz == 4) {
            float *t = (float *)tmp.data();
            for (long long c = 0; c < ctx->N; ++c) t[c] = src[f] ? (float)src[f][c] : 0.f;
        } else {
            double amax = 0.0;
            if (src[f])
                for (long long c = 0; c < ctx->N; ++c) amax = std::fmax(amax, std::fabs(src[f][c]));
            ctx->fexp[f] = exponent_for(amax, 0);
            const double is = std::ldexp(1.0, -ctx->fexp[f]);
            uint16_t *t = (uint16_t *)tmp.data();
            for (long long c = 0; c < ctx->N; ++c) t[c] = half_bits(src[f] ? src[f][c] * is : 0.0);
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

}  // extern "C"

namespace {

// Checks of both slice entry points; *per = floats of the three planes of one field.
int slices_args(fases_ctx *ctx, const char *who, int i0, int j0, int k0, const float *out, size_t *per) {
    if (!ctx || !out) return FASES_E_ARG;
    if (!ctx->have_geo) return set_error(ctx, FASES_E_STATE, std::string(who) + ": no geometry");
    const int Nx = ctx->Nx, Ny = ctx->Ny, Nz = ctx->Nz;
    if (i0 < 0 || i0 >= Nx || j0 < 0 || j0 >= Ny || k0 < 0 || k0 >= Nz)
        return set_error(ctx, FASES_E_RANGE, std::string(who) + ": slice plane outside the grid");
    *per = (size_t)Ny * Nz + (size_t)Nx * Nz + (size_t)Nx * Ny;
    return FASES_OK;
}

// Gathers the planes of p, u, v, w into the device buffer dst [4 * per] and copies them
// to the host array out (one copy), on the context's stream; returns once out is filled.
int slices_to_host(fases_ctx *ctx, int i0, int j0, int k0, size_t per, float *dst_dev, float *out) {
    const int Nx = ctx->Nx, Ny = ctx->Ny, Nz = ctx->Nz;
    const unsigned nb = (unsigned)((per + 255) / 256);
    for (int f = 0; f < 4; ++f) {
        float *dst = dst_dev + (size_t)f * per;
        if (ctx->esz == 8)
            hipLaunchKernelGGL(k_slices<double>, dim3(nb), dim3(256), 0, ctx->stream, (const double *)ctx->U[0][f],
                               1.0f, Nx, Ny, Nz, i0, j0, k0, dst);
        else if (ctx->esz == 4)
            hipLaunchKernelGGL(k_slices<float>, dim3(nb), dim3(256), 0, ctx->stream, (const float *)ctx->U[0][f],
                               1.0f, Nx, Ny, Nz, i0, j0, k0, dst);
        else
            hipLaunchKernelGGL(k_slices<__half>, dim3(nb), dim3(256), 0, ctx->stream, (const __half *)ctx->U[0][f],
                               (float)std::ldexp(1.0, ctx->fexp[f]), Nx, Ny, Nz, i0, j0, k0, dst);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipMemcpyAsync(out, dst_dev, 4 * per * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return FASES_OK;
}

}  // namespace

// Staging buffer of fases_get_slices_noalloc: a file-scope device array, one copy per GPU
// the library runs on, held until the process ends. In host code its name is not a device
// address: host code uses only the per-device pointer from hipGetSymbolAddress.
constexpr size_t kSliceScratchFloats = (size_t)1 << 21;      // 8 MiB: 2^21 floats
static __device__ __align__(256) float scratch[kSliceScratchFloats];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of `scratch` on the context's device, resolved once per device with that
// device current.
int slice_scratch(fases_ctx *ctx, float **ptr) {
    const int dev = ctx->prm.device;
    if (dev < 0 || dev >= kMaxScratchDevices)
        return set_error(ctx, FASES_E_ARG, "fases_get_slices_noalloc: device index above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        HIPCHK(hipSetDevice(dev));
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<float *>(p);
    }
    *ptr = g_scratch_ptr[dev];
    return FASES_OK;
}

}  // namespace

extern "C" {

int fases_get_slices(const fases_ctx *cctx, int i0, int j0, int k0, float *out) {
    fases_ctx *ctx = const_cast<fases_ctx *>(cctx);
    size_t per = 0;
    int rc = slices_args(ctx, "fases_get_slices", i0, j0, k0, out, &per);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->prm.device));
    if (ctx->n_slices < 4 * per) {
        rc = dalloc_replace(ctx, &ctx->d_slices, 4 * per);
        if (rc) return rc;
        ctx->n_slices = 4 * per;
    }
    return slices_to_host(ctx, i0, j0, k0, per, ctx->d_slices, out);
}

int fases_get_slices_noalloc(const fases_ctx *cctx, int i0, int j0, int k0, float *out) {
    fases_ctx *ctx = const_cast<fases_ctx *>(cctx);
    size_t per = 0;
    int rc = slices_args(ctx, "fases_get_slices_noalloc", i0, j0, k0, out, &per);
    if (rc) return rc;
    if (4 * per > kSliceScratchFloats)
        return set_error(ctx, FASES_E_ARG,
                         "fases_get_slices_noalloc: the planes need " + std::to_string(4 * per) +
                             " floats, above its static staging buffer of " + std::to_string(kSliceScratchFloats) +
                             " (use fases_get_slices)");
    HIPCHK(hipSetDevice(ctx->prm.device));
    float *buf = nullptr;
    if ((rc = slice_scratch(ctx, &buf))) return rc;
    return slices_to_host(ctx, i0, j0, k0, per, buf, out);
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
        else if (ctx->esz == 4) {
            const float *t = (const float *)tmp.data();
            for (long long c = 0; c < ctx->N; ++c) dst[f][c] = t[c];
        } else {
            const uint16_t *t = (const uint16_t *)tmp.data();
            const double sc = std::ldexp(1.0, ctx->fexp[f]);
            for (long long c = 0; c < ctx->N; ++c) dst[f][c] = (double)half_value(t[c]) * sc;
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
    if (ctx->prm.precision == FASES_FP16) return run_impl<__half, float>(ctx, nsteps, ir, vel, elapsed_ms, kernel_ms);
    return run_impl<float, float>(ctx, nsteps, ir, vel, elapsed_ms, kernel_ms);
}

int fases_pml_tables(const fases_ctx *ctx, int axis, int32_t *band, double *kc, double *b, double *a) {
    if (!ctx || axis < 0 || axis > 2) return FASES_E_ARG;
    if (!ctx->have_geo) return set_error(const_cast<fases_ctx *>(ctx), FASES_E_STATE, "fases_pml_tables: no geometry");
    const size_t n = ctx->h_pml_band[axis].size();
    for (size_t i = 0; i < n; ++i) {
        if (band) band[i] = ctx->h_pml_band[axis][i];
        if (kc) kc[i] = ctx->h_pml_kc[axis][i];
        if (b) b[i] = ctx->h_pml_b[axis][i];
        if (a) a[i] = ctx->h_pml_a[axis][i];
    }
    return FASES_OK;
}

int fases_tile_config(const fases_ctx *ctx, int *tile_x, int *tile_y, int *tile_z) {
    if (!ctx) return FASES_E_ARG;
    if (tile_x) *tile_x = ctx->tx;
    if (tile_y) *tile_y = ctx->nw;
    if (tile_z) *tile_z = ctx->tz;
    return FASES_OK;
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
