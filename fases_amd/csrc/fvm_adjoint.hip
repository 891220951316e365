// Reverse-time adjoint of the FVM run: gradients of an IR misfit J with respect to the
// ADE material parameters (R_inf, residues A, B, C, pole real parts zeta, alpha and the
// pair frequencies beta) and to the source drive (amplitude of each source and the
// shared waveform samples).
//
// Memory: the forward run stores a checkpoint (U^n, the ADE state, the FP64 patch
// shadows) every `every` steps. The backward sweep walks the segments last to first:
// it restores the segment's checkpoint, recomputes the segment forward into per-step
// slots (the inputs V0, V1, V2 of each step's three stages) and then runs the adjoint
// steps of the segment in reverse, each stage in reverse order: boundary + ADE adjoint,
// source gradients, flux adjoint, probe-adjoint gather.
//
// cPML: the checkpoints also hold the PML auxiliary states (the segment recompute needs
// them); the reverse sweep adds the PML part of the flux adjoint and carries the adjoint
// of the auxiliary states from step to step (k_pml_flux_adj, k_pml_psi_adj).
//
// FP16 storage: the forward rescales the fields at the steps the plain run does and keeps
// each step's scale exponents; the segment recompute applies the same exact power-of-two
// rescale to the step's start set, and every adjoint kernel of a step reads that step's
// stored stage inputs with that step's scales.
//
// Scope: FP16, FP32 or FP64 field storage.

#include "fvm_adjoint.hpp"
#include "fvm_ctx.hpp"

#include <cstring>
#include <numeric>

namespace {


// Tiled flux adjoint: waves per block (the inner rows - 4 rows are outputs) and the z
// planes per block. FP32: 16 rows, one block per CU at 95 VGPRs; 8-row blocks (two per
// CU, 104 VGPRs) were measured slower (reverse sweep 7.25 s vs 5.89 s on the 192x160x128
// room over 3000 steps): their halo rows cost more than the second block recovers.
// FP64: 8 rows (185 VGPRs, nothing spilled). The 16-row FP64 build has 128 VGPRs and
// spills ~180 of them to scratch, and that build gave gradients that changed from run to
// run (relative spread 1e-6, single entries off by 1e12 against the per-cell adjoint)
// although its source is the FP32 kernel's, which is bit-reproducible in any block order,
// as is the 8-row FP64 build.
template <typename R> constexpr int kAdjRowsOf = std::is_same<R, double>::value ? 8 : 16;
constexpr int kAdjTileZ = 32;

// z planes per block of the tiled flux adjoint. A block holds one CU (16 waves at 94
// VGPRs), so a room whose tiles at the default depth number fewer than the CUs leaves CUs
// idle for the whole launch; the depth is then cut so that the tile count comes as close
// to the CU count as it can without exceeding it (each block then marches fewer planes,
// plus its four pipeline-fill planes): on the 192x160x128 room 3 x 14 x 4 tiles of 32
// planes become 3 x 14 x 6 of 22.
int adjoint_tile_z(const fases_ctx *ctx, int nbx, int nby) {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->prm.device) != hipSuccess || ncu < 1)
        return kAdjTileZ;
    const int nz = ctx->Nz - 6;
    const long long cols = (long long)nbx * nby;
    if (cols * ((nz + kAdjTileZ - 1) / kAdjTileZ) >= ncu) return kAdjTileZ;
    const long long zb = std::max<long long>(1, ncu / cols);
    return std::max(8, (int)((nz + zb - 1) / zb));
}

// Device buffers of the adjoint, owned by the context (freed with it).
struct AdjBufs {
    std::vector<void *> allocs;
    int every = 0, n_ckpt = 0;
    long long nsteps = 0, step0 = 0;
    size_t esz = 0;
    int P = 0;                                   // gradient entries per material
    // sizes the buffers were made for (reused by a later forward of the same shape)
    long long key[8] = {};
    // checkpoints: fields (4 arrays), patch shadows [4 np], ADE state
    std::vector<void *> ck_f[4];
    std::vector<double *> ck_sh;
    std::vector<AdeState> ck_x;
    // FP16 storage: the field scale exponents in effect for each step's stages [nsteps][4]
    std::vector<int> fexp;
    // segment slots: V0[0..every], V1[0..every-1], V2[0..every-1] (+ shadows), ADE x0 per step
    std::vector<void *> v0[4], v1[4], v2[4];
    std::vector<double *> s0, s1, s2;
    std::vector<AdeState> xs;
    // adjoint fields (R precision): next (U^{n+1}), V2, V1, U0, and their FP64 shadows on
    // the source-patch cells ([4][np] each; reduced-precision storage only)
    void *lam[4][4] = {};
    double *lam_sh[4] = {};
    // adjoint patch: the forward patch cells within the FP64 source radius
    // (source_patch_radius: the FP64 adjoint covers the whole forward patch)
    int n_ap = 0;
    int *ap_cells = nullptr, *ap_idx = nullptr;
    // the patch dilated by two cells along each axis (every cell whose pair windows reach a
    // patch cell), grid -> dilated index, and its pair VJPs [3][nD][10]
    int nD = 0;
    int *d_cells = nullptr, *d_map = nullptr;
    double *pv = nullptr;
    AdeState lx[4] = {}, rx1 = {}, rx2 = {};
    double *lD = nullptr, *G = nullptr, *gs = nullptr, *gus = nullptr, *dj = nullptr;
    long long gus_n = 0;
    // cPML: psi checkpoints [6] per checkpoint, psi adjoints [6] (px, ux, py, vy, pz, wz),
    // per axis the positions whose flux windows reach the band
    std::vector<double *> ck_pml[6];
    double *lpsi[6] = {};
    int *plist[3] = {};
    int npos[3] = {};
    size_t pml_n[3] = {};
    // probe gather CSR
    int nT = 0;
    int *tgt = nullptr, *toff = nullptr, *eb = nullptr;
    double *ew = nullptr;
    ~AdjBufs() {
        for (void *p : allocs) (void)hipFree(p);
    }
};

void release_adj(void *p) { delete static_cast<AdjBufs *>(p); }

template <typename T> int adj_alloc(fases_ctx *ctx, AdjBufs &A, T **out, size_t count) {
    *out = nullptr;
    void *p = nullptr;
    const size_t bytes = std::max<size_t>(1, count) * sizeof(T);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        ctx->err = "adjoint: hipMalloc(" + std::to_string(bytes) + " B): " + hipGetErrorString(e);
        return FASES_E_NOMEM;
    }
    A.allocs.push_back(p);
    e = hipMemsetAsync(p, 0, bytes, ctx->stream);
    if (e != hipSuccess) { ctx->err = hipGetErrorString(e); return FASES_E_HIP; }
    *out = static_cast<T *>(p);
    return FASES_OK;
}

int alloc_ade(fases_ctx *ctx, AdjBufs &A, AdeState &x) {
    int rc;
    if ((rc = adj_alloc(ctx, A, &x.phi, (size_t)ctx->nB * ctx->K))) return rc;
    if ((rc = adj_alloc(ctx, A, &x.p1, (size_t)ctx->nB * ctx->L))) return rc;
    return adj_alloc(ctx, A, &x.p2, (size_t)ctx->nB * ctx->L);
}

int copy_ade(fases_ctx *ctx, const AdeState &dst, const double *phi, const double *p1, const double *p2) {
    const size_t nk = (size_t)ctx->nB * ctx->K * 8, nl = (size_t)ctx->nB * ctx->L * 8;
    HIPCHK(hipMemcpyAsync(dst.phi, phi, nk, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(dst.p1, p1, nl, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(dst.p2, p2, nl, hipMemcpyDeviceToDevice, ctx->stream));
    return FASES_OK;
}

int zero_ade(fases_ctx *ctx, const AdeState &x) {
    HIPCHK(hipMemsetAsync(x.phi, 0, (size_t)ctx->nB * ctx->K * 8, ctx->stream));
    HIPCHK(hipMemsetAsync(x.p1, 0, (size_t)ctx->nB * ctx->L * 8, ctx->stream));
    HIPCHK(hipMemsetAsync(x.p2, 0, (size_t)ctx->nB * ctx->L * 8, ctx->stream));
    return FASES_OK;
}

// Probe-adjoint CSR: for every (boundary cell, trilinear corner) the target grid cell
// and the corner weight wx wy wz, grouped by target.
int build_probe_csr(fases_ctx *ctx, AdjBufs &A) {
    const int nB = ctx->nB;
    const std::vector<int> &mat = ctx->h_bmat;
    struct E { int tgt, b; double w; };
    std::vector<E> es;
    es.reserve((size_t)nB * 8);
    const long long nx = ctx->Nx, nxy = (long long)ctx->Nx * ctx->Ny;
    for (int b = 0; b < nB; ++b) {
        // the probe geometry the boundary kernels compute for this cell
        const int C = ctx->h_bcell[(size_t)b];
        const int i = (int)(C % nx), j = (int)((C / nx) % ctx->Ny), k = (int)(C / nxy);
        const ProbeGeo q = probe_geo(i, j, k, ctx->Nx, ctx->Ny, ctx->Nz, (double)ctx->h_bdist[(size_t)b],
                                     ctx->h_bnrm[(size_t)b], ctx->h_bnrm[(size_t)nB + b],
                                     ctx->h_bnrm[2 * (size_t)nB + b], ctx->prm.h, ctx->prm.tol);
        if (mat[(size_t)b] & 0x100) { es.push_back({q.base, b, 1.0}); continue; }
        const double sx = q.sx, sy = q.sy, sz = q.sz;
        for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    const double w = (dx ? sx : 1.0 - sx) * (dy ? sy : 1.0 - sy) * (dz ? sz : 1.0 - sz);
                    es.push_back({(int)(q.base + dx + nx * dy + nxy * dz), b, w});
                }
    }
    std::stable_sort(es.begin(), es.end(), [](const E &a, const E &b) { return a.tgt < b.tgt; });
    std::vector<int> tgt, off, eb;
    std::vector<double> ew;
    for (size_t q = 0; q < es.size(); ++q) {
        if (q == 0 || es[q].tgt != es[q - 1].tgt) { tgt.push_back(es[q].tgt); off.push_back((int)q); }
        eb.push_back(es[q].b);
        ew.push_back(es[q].w);
    }
    off.push_back((int)es.size());
    A.nT = (int)tgt.size();
    int rc;
    if ((rc = adj_alloc(ctx, A, &A.tgt, tgt.size()))) return rc;
    if ((rc = adj_alloc(ctx, A, &A.toff, off.size()))) return rc;
    if ((rc = adj_alloc(ctx, A, &A.eb, eb.size()))) return rc;
    if ((rc = adj_alloc(ctx, A, &A.ew, ew.size()))) return rc;
    // on the context stream, behind the zero fill adj_alloc queued there (the stream is
    // non-blocking: a null-stream copy could land first and then be zeroed)
    HIPCHK(hipMemcpyAsync(A.tgt, tgt.data(), tgt.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(A.toff, off.data(), off.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(A.eb, eb.data(), eb.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(A.ew, ew.data(), ew.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));          // the host vectors end here
    return FASES_OK;
}

// The context's six psi arrays in the adjoint's order (px, ux, py, vy, pz, wz).
struct PmlPsi {
    double *a[6];
};
PmlPsi pml_psi(const fases_ctx *ctx) {
    return PmlPsi{{ctx->pml.psi_px, ctx->pml.psi_ux, ctx->pml.psi_py, ctx->pml.psi_vy, ctx->pml.psi_pz,
                   ctx->pml.psi_wz}};
}

// cPML buffers of the adjoint: psi checkpoints, psi adjoints and, per axis, the list of
// positions in [3, N-3) within 3 cells of a band cell (the reach of a WENO pair and its
// two faces), where the PML part of the flux adjoint is non-zero.
int setup_pml(fases_ctx *ctx, AdjBufs &A) {
    if (!ctx->pml.on) return FASES_OK;
    const size_t d2 = 2 * (size_t)ctx->pml.d;
    A.pml_n[0] = d2 * ctx->Ny * ctx->Nz;
    A.pml_n[1] = (size_t)ctx->Nx * d2 * ctx->Nz;
    A.pml_n[2] = (size_t)ctx->Nx * ctx->Ny * d2;
    int rc;
    for (int c = 0; c < A.n_ckpt; ++c)
        for (int a = 0; a < 6; ++a) {
            double *p;
            if ((rc = adj_alloc(ctx, A, &p, A.pml_n[a / 2]))) return rc;
            A.ck_pml[a].push_back(p);
        }
    for (int a = 0; a < 6; ++a)
        if ((rc = adj_alloc(ctx, A, &A.lpsi[a], A.pml_n[a / 2]))) return rc;
    const int nax[3] = {ctx->Nx, ctx->Ny, ctx->Nz};
    const int *band[3] = {ctx->pml.x.band, ctx->pml.y.band, ctx->pml.z.band};
    for (int ax = 0; ax < 3; ++ax) {
        std::vector<int> bh((size_t)nax[ax]), pl;
        HIPCHK(hipMemcpy(bh.data(), band[ax], bh.size() * 4, hipMemcpyDeviceToHost));
        for (int p = NG; p < nax[ax] - NG; ++p) {
            bool near = false;
            for (int q = std::max(0, p - 3); q <= std::min(nax[ax] - 1, p + 3) && !near; ++q) near = bh[(size_t)q] >= 0;
            if (near) pl.push_back(p);
        }
        A.npos[ax] = (int)pl.size();
        if (pl.empty()) continue;
        if ((rc = adj_alloc(ctx, A, &A.plist[ax], pl.size()))) return rc;
        HIPCHK(hipMemcpyAsync(A.plist[ax], pl.data(), pl.size() * 4, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));      // pl ends with this iteration
    }
    return FASES_OK;
}

int copy_pml(fases_ctx *ctx, AdjBufs &A, int c, bool save) {
    if (!ctx->pml.on) return FASES_OK;
    const PmlPsi psi = pml_psi(ctx);
    for (int a = 0; a < 6; ++a) {
        double *ck = A.ck_pml[a][(size_t)c];
        HIPCHK(hipMemcpyAsync(save ? ck : psi.a[a], save ? psi.a[a] : ck, A.pml_n[a / 2] * 8, hipMemcpyDeviceToDevice,
                              ctx->stream));
    }
    return FASES_OK;
}

template <typename S, typename R, bool LAST>
void launch_pml_adj(fases_ctx *ctx, AdjBufs &A, const Grid &g, const FieldSet<S> &vin, const Scl<R> &sc,
                    const FieldSet<R> &Lo, const FieldSet<R> &Li, const int *pidx, const double *lsh_out,
                    double *lsh_in, const AdjK<R> &K, double beta) {
    const PmlAxis *axs[3] = {&ctx->pml.x, &ctx->pml.y, &ctx->pml.z};
    const R ih = (R)(1.0 / ctx->prm.h);
    const int np = A.n_ap;
    for (int ax = 0; ax < 3; ++ax) {
        if (A.npos[ax] == 0) continue;
        const PmlAdjAxis pa{axs[ax]->band, axs[ax]->kc, axs[ax]->b, axs[ax]->a, A.lpsi[2 * ax], A.lpsi[2 * ax + 1],
                            ctx->pml.d};
        const long long ncell = ax == 0 ? (long long)A.npos[0] * (g.Ny - 6) * (g.Nz - 6)
                              : ax == 1 ? (long long)(g.Nx - 6) * A.npos[1] * (g.Nz - 6)
                                        : (long long)(g.Nx - 6) * (g.Ny - 6) * A.npos[2];
        const dim3 fg((unsigned)((ncell + 255) / 256)), pg((unsigned)((A.pml_n[ax] + 255) / 256));
        const R *Lu = ax == 0 ? Lo.u : (ax == 1 ? Lo.v : Lo.w);
#define FASES_PML_ADJ(AXV)                                                                                          \
        hipLaunchKernelGGL((k_pml_flux_adj<S, R, AXV, LAST>), fg, dim3(256), 0, ctx->stream, g, vin, sc, Lo.p, Lo.u,  \
                           Lo.v, Lo.w, ctx->d_upd, Li.p, Li.u, Li.v, Li.w, pidx, lsh_out, lsh_in, np, pa,             \
                           A.plist[AXV], A.npos[AXV], K, ih);                                                          \
        hipLaunchKernelGGL((k_pml_psi_adj<R, AXV, LAST>), pg, dim3(256), 0, ctx->stream, g, Lo.p, Lu, ctx->d_upd,     \
                           pidx, lsh_out, np, pa, beta * ctx->prm.dt)
        if (ax == 0) { FASES_PML_ADJ(0); }
        else if (ax == 1) { FASES_PML_ADJ(1); }
        else { FASES_PML_ADJ(2); }
#undef FASES_PML_ADJ
    }
}

template <typename S> FieldSet<S> fs_of(const std::vector<void *> (&v)[4], size_t q) {
    return FieldSet<S>{(S *)v[0][q], (S *)v[1][q], (S *)v[2][q], (S *)v[3][q]};
}

template <typename S, typename R>
int setup(fases_ctx *ctx, AdjBufs &A, long long nsteps, int every) {
    A.every = every;
    A.nsteps = nsteps;
    A.n_ckpt = (int)((nsteps + every - 1) / every);
    A.esz = sizeof(S);
    A.P = 1 + 2 * ctx->K + 4 * ctx->L;
    const size_t N = (size_t)ctx->N, np = (size_t)std::max(ctx->n_patch, 0);
    int rc;
    for (int c = 0; c < A.n_ckpt; ++c) {
        for (int f = 0; f < 4; ++f) {
            S *p;
            if ((rc = adj_alloc(ctx, A, &p, N))) return rc;
            A.ck_f[f].push_back(p);
        }
        double *sh = nullptr;
        if (np && (rc = adj_alloc(ctx, A, &sh, 4 * np))) return rc;
        A.ck_sh.push_back(sh);
        AdeState x{};
        if ((rc = alloc_ade(ctx, A, x))) return rc;
        A.ck_x.push_back(x);
    }
    const int ns = (int)std::min<long long>(every, nsteps);    // slots of the longest segment
    for (int q = 0; q <= ns; ++q) {
        for (int f = 0; f < 4; ++f) {
            S *p;
            if ((rc = adj_alloc(ctx, A, &p, N))) return rc;
            A.v0[f].push_back(p);
            if (q < ns) {
                if ((rc = adj_alloc(ctx, A, &p, N))) return rc;
                A.v1[f].push_back(p);
                if ((rc = adj_alloc(ctx, A, &p, N))) return rc;
                A.v2[f].push_back(p);
            }
        }
        double *sh = nullptr;
        if (np && (rc = adj_alloc(ctx, A, &sh, 4 * np))) return rc;
        A.s0.push_back(sh);
        if (q < ns) {
            if (np && (rc = adj_alloc(ctx, A, &sh, 4 * np))) return rc;
            A.s1.push_back(np ? sh : nullptr);
            if (np && (rc = adj_alloc(ctx, A, &sh, 4 * np))) return rc;
            A.s2.push_back(np ? sh : nullptr);
            AdeState x{};
            if ((rc = alloc_ade(ctx, A, x))) return rc;
            A.xs.push_back(x);
        }
    }
    for (int s = 0; s < 4; ++s)
        for (int f = 0; f < 4; ++f) {
            R *p;
            if ((rc = adj_alloc(ctx, A, &p, N))) return rc;
            A.lam[s][f] = p;
        }
    if (np) {
        // adjoint patch cells: forward patch cells near a source, in forward patch order
        std::vector<int> scell((size_t)std::max(ctx->n_src, 0));
        if (!scell.empty())
            HIPCHK(hipMemcpy(scell.data(), ctx->d_scell, scell.size() * 4, hipMemcpyDeviceToHost));
        const long long nx = ctx->Nx, nxy = (long long)ctx->Nx * ctx->Ny;
        const long long rad = source_patch_radius(ctx);
        std::vector<int> apc, api((size_t)ctx->N, -1);
        for (int c : ctx->h_patch) {
            const long long i = c % nx, j = (c / nx) % ctx->Ny, k = c / nxy;
            bool near = false;
            for (int s0 : scell) {
                const long long si = s0 % nx, sj = (s0 / nx) % ctx->Ny, sk = s0 / nxy;
                if (std::llabs(i - si) <= rad && std::llabs(j - sj) <= rad && std::llabs(k - sk) <= rad) {
                    near = true;
                    break;
                }
            }
            if (near) { api[(size_t)c] = (int)apc.size(); apc.push_back(c); }
        }
        A.n_ap = (int)apc.size();
        if ((rc = adj_alloc(ctx, A, &A.ap_cells, apc.size()))) return rc;
        if ((rc = adj_alloc(ctx, A, &A.ap_idx, api.size()))) return rc;
        // ordered after adj_alloc's zero fill on the (non-blocking) context stream
        HIPCHK(hipMemcpyAsync(A.ap_cells, apc.data(), apc.size() * 4, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(A.ap_idx, api.data(), api.size() * 4, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));      // apc / api end with this block
        for (auto &l : A.lam_sh)
            if ((rc = adj_alloc(ctx, A, &l, 4 * (size_t)std::max(A.n_ap, 1)))) return rc;
        {
            const int nax[3] = {ctx->Nx, ctx->Ny, ctx->Nz};
            const long long st[3] = {1, nx, nxy};
            std::vector<int> dc, dm((size_t)ctx->N, -1);
            for (int c : apc) {
                const int pos[3] = {(int)(c % nx), (int)((c / nx) % ctx->Ny), (int)(c / nxy)};
                for (int ax = 0; ax < 3; ++ax)
                    for (int q = -2; q <= 2; ++q) {
                        const int cpos = pos[ax] - q;
                        if (cpos < 2 || cpos > nax[ax] - 3) continue;
                        const long long X = c - q * st[ax];
                        if (dm[(size_t)X] < 0) { dm[(size_t)X] = (int)dc.size(); dc.push_back((int)X); }
                    }
            }
            A.nD = (int)dc.size();
            if ((rc = adj_alloc(ctx, A, &A.d_cells, dc.size()))) return rc;
            if ((rc = adj_alloc(ctx, A, &A.d_map, dm.size()))) return rc;
            if ((rc = adj_alloc(ctx, A, &A.pv, 30 * (size_t)std::max(A.nD, 1)))) return rc;
            HIPCHK(hipMemcpyAsync(A.d_cells, dc.data(), dc.size() * 4, hipMemcpyHostToDevice, ctx->stream));
            HIPCHK(hipMemcpyAsync(A.d_map, dm.data(), dm.size() * 4, hipMemcpyHostToDevice, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));  // dc / dm end with this block
        }
    }
    for (auto &x : A.lx)
        if ((rc = alloc_ade(ctx, A, x))) return rc;
    if ((rc = alloc_ade(ctx, A, A.rx1)) || (rc = alloc_ade(ctx, A, A.rx2))) return rc;
    if ((rc = adj_alloc(ctx, A, &A.lD, (size_t)ctx->nB * 4))) return rc;
    if ((rc = adj_alloc(ctx, A, &A.G, (size_t)ctx->nB * A.P))) return rc;
    if ((rc = adj_alloc(ctx, A, &A.gs, (size_t)std::max(ctx->n_src, 1)))) return rc;
    if ((rc = adj_alloc(ctx, A, &A.dj, (size_t)std::max(ctx->n_rec, 1) * nsteps))) return rc;
    if ((rc = setup_pml(ctx, A))) return rc;
    return build_probe_csr(ctx, A);
}

template <typename S, typename R>
int forward_impl(fases_ctx *ctx, long long nsteps, int every, double *ir) {
    const long long key[8] = {nsteps, every, ctx->N, ctx->nB, (long long)ctx->K * 1000 + ctx->L, ctx->n_patch,
                              (long long)ctx->n_src * 100000 + ctx->n_rec,
                              (long long)sizeof(S) + 100LL * (ctx->pml.on ? ctx->pml.d : 0)};
    AdjBufs *A = static_cast<AdjBufs *>(ctx->adj);
    const bool reuse = A && std::equal(key, key + 8, A->key);
    int rc;
    if (!reuse) {
        ctx->release_adjoint();
        A = new (std::nothrow) AdjBufs();
        if (!A) return set_error(ctx, FASES_E_NOMEM, "adjoint: out of host memory");
        ctx->adj = A;
        ctx->adj_release = release_adj;
        if ((rc = setup<S, R>(ctx, *A, nsteps, every))) { ctx->release_adjoint(); return rc; }
        std::copy(key, key + 8, A->key);
    }
    A->step0 = ctx->step;
    if ((rc = prepare_run(ctx, nsteps))) return rc;
    if (uses_tiled(ctx) && ctx->tz == 0 && (rc = autotune_tile<S, R>(ctx))) return rc;
    const size_t N = (size_t)ctx->N, np = (size_t)std::max(ctx->n_patch, 0);
    constexpr bool kHalf = std::is_same<S, __half>::value;
    A->fexp.assign(kHalf ? (size_t)nsteps * 4 : 0, 0);
    for (long long q = 0; q < nsteps; ++q) {
        if (kHalf) {                    // the plain run's rescale schedule (run_impl)
            if (rescale_due(ctx->step, ctx->prm.fp16_rescale_every) && (rc = rescale_fields<S>(ctx))) return rc;
            std::copy(ctx->fexp, ctx->fexp + 4, A->fexp.begin() + 4 * q);
        }
        if (q % every == 0) {
            const int c = (int)(q / every);
            for (int f = 0; f < 4; ++f)
                HIPCHK(hipMemcpyAsync(A->ck_f[f][(size_t)c], ctx->U[0][f], N * sizeof(S), hipMemcpyDeviceToDevice,
                                      ctx->stream));
            if (np)
                HIPCHK(hipMemcpyAsync(A->ck_sh[(size_t)c], ctx->shadow[0], 4 * np * 8, hipMemcpyDeviceToDevice,
                                      ctx->stream));
            if ((rc = copy_ade(ctx, A->ck_x[(size_t)c], ctx->phi[1], ctx->psi1[1], ctx->psi2[1]))) return rc;
            if ((rc = copy_pml(ctx, *A, c, true))) return rc;
        }
        if ((rc = launch_step<S, R>(ctx, ctx->step, nullptr))) return rc;
        if ((rc = launch_receivers<S>(ctx, nsteps, q))) return rc;
        ctx->step += 1;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipGetLastError());
    const long long need = (long long)ctx->n_rec * nsteps;
    if (ir && need > 0) {
        HIPCHK(hipMemcpy(ir, ctx->d_ir, (size_t)need * sizeof(double), hipMemcpyDeviceToHost));
        for (long long i = 0; i < need; ++i)
            if (!std::isfinite(ir[i])) return set_error(ctx, FASES_E_NONFINITE, "receiver sample is not finite");
    }
    return FASES_OK;
}

template <typename R> FieldSet<R> lam_set(AdjBufs &A, int s) {
    return FieldSet<R>{(R *)A.lam[s][0], (R *)A.lam[s][1], (R *)A.lam[s][2], (R *)A.lam[s][3]};
}

template <typename S, typename R>
int backward_impl(fases_ctx *ctx, const double *dJ, double *grad_mat, double *grad_sval, double *grad_us) {
    AdjBufs &A = *static_cast<AdjBufs *>(ctx->adj);
    const fases_params &P = ctx->prm;
    const long long nsteps = A.nsteps;
    const size_t N = (size_t)ctx->N, np = (size_t)std::max(ctx->n_patch, 0);
    const Grid g{ctx->Nx, ctx->Ny, ctx->Nz, (long long)ctx->Nx * ctx->Ny};
    Scl<R> sc = field_scales<R>(ctx);
    Scl<double> scd = field_scales<double>(ctx);
    // FP16 storage: step n's scale exponents (the forward's history) into the context
    constexpr bool kHalf = std::is_same<S, __half>::value;
    int fexp_end[4];
    std::copy(ctx->fexp, ctx->fexp + 4, fexp_end);
    auto step_scales = [&](long long n) {
        if (!kHalf) return;
        std::copy(A.fexp.begin() + 4 * (n - A.step0), A.fexp.begin() + 4 * (n - A.step0) + 4, ctx->fexp);
        sc = field_scales<R>(ctx);
        scd = field_scales<double>(ctx);
    };
    const Ade ade{ctx->K, ctx->L, ctx->d_Rinf, ctx->d_A, ctx->d_zeta, ctx->d_B, ctx->d_C, ctx->d_alpha, ctx->d_beta,
                  ctx->d_Keff, ctx->d_Leff};
    const Bnd bd = bnd_of(ctx);
    const double As = 6.0 * P.h * P.h;
    const double coef0 = P.dt * P.c * P.c * P.rho * As / (P.h * P.h * P.h);
    int rc;
    // inputs and zeroed accumulators
    if (ctx->n_rec > 0)
        HIPCHK(hipMemcpyAsync(A.dj, dJ, (size_t)ctx->n_rec * nsteps * 8, hipMemcpyHostToDevice, ctx->stream));
    for (int s = 0; s < 4; ++s) {
        for (int f = 0; f < 4; ++f) HIPCHK(hipMemsetAsync(A.lam[s][f], 0, N * sizeof(R), ctx->stream));
        if (A.n_ap) HIPCHK(hipMemsetAsync(A.lam_sh[s], 0, 4 * (size_t)A.n_ap * 8, ctx->stream));
    }
    const int nap = A.n_ap;
    const int *pidx = nap ? A.ap_idx : nullptr;       // cells with an FP64 adjoint
    for (auto &x : A.lx)
        if ((rc = zero_ade(ctx, x))) return rc;
    HIPCHK(hipMemsetAsync(A.G, 0, (size_t)ctx->nB * A.P * 8, ctx->stream));
    HIPCHK(hipMemsetAsync(A.gs, 0, (size_t)std::max(ctx->n_src, 1) * 8, ctx->stream));
    for (int a = 0; a < 6; ++a)
        if (A.lpsi[a]) HIPCHK(hipMemsetAsync(A.lpsi[a], 0, A.pml_n[a / 2] * 8, ctx->stream));
    const long long gus_n = (ctx->n_src > 0) ? ctx->n_samp * (ctx->src_opt == FASES_SRC_SINE ? ctx->n_src : 1) : 0;
    if (grad_us && gus_n > 0 && ctx->src_opt != FASES_SRC_SINE) {
        if (A.gus_n < gus_n) {
            if ((rc = adj_alloc(ctx, A, &A.gus, (size_t)gus_n))) return rc;
            A.gus_n = gus_n;
        }
        HIPCHK(hipMemsetAsync(A.gus, 0, (size_t)gus_n * 8, ctx->stream));
    }
    double *gus = (grad_us && ctx->src_opt != FASES_SRC_SINE && gus_n > 0) ? A.gus : nullptr;
    // logical adjoint sets: index into A.lam
    int iNext = 0, iV2 = 1, iV1 = 2, iU0 = 3;
    // ADE adjoints: lx[ix3] = lambda of X^{n+1}, ix0..ix2 zeroed per step
    int ix0 = 0, ix1 = 1, ix2 = 2, ix3 = 3;
    const dim3 fb(64, 4, 1), fgrid((ctx->Nx - 6 + 63) / 64, (ctx->Ny - 6 + 3) / 4, ctx->Nz - 6);
    const int sb = 256;
    const dim3 bgrid((ctx->nB + sb - 1) / sb), tgrid((A.nT + sb - 1) / sb);
    const R kp = (R)(0.5 * P.rho * P.c * P.c), kc = (R)(0.5 * P.c), kr = (R)(0.5 / P.rho);
    // flux adjoint: tiled kernel unless the per-cell one was asked for
    const bool tiled = P.adjoint_variant != 1;
    const int anx = (ctx->Nx - 6 + adjtile::TX - 1) / adjtile::TX;
    const int any = (ctx->Ny - 6 + kAdjRowsOf<R> - 5) / (kAdjRowsOf<R> - 4);
    const int atz = adjoint_tile_z(ctx, anx, any);
    dim3 agrid(anx, any, (ctx->Nz - 6 + atz - 1) / atz);
    // variant 2: the same tiles in the XCD-aware order, on a 1-D grid
    const int anbx = P.adjoint_variant == 2 ? (int)agrid.x : 0, anby = (int)agrid.y;
    if (anbx) agrid = dim3(agrid.x * agrid.y * agrid.z);
    for (int c = A.n_ckpt - 1; c >= 0; --c) {
        const long long n0 = A.step0 + (long long)c * A.every;
        const long long len = std::min<long long>(A.every, A.step0 + nsteps - n0);
        // restore the checkpoint into slot 0 and the context's ADE state
        for (int f = 0; f < 4; ++f)
            HIPCHK(hipMemcpyAsync(A.v0[f][0], A.ck_f[f][(size_t)c], N * sizeof(S), hipMemcpyDeviceToDevice,
                                  ctx->stream));
        if (np)
            HIPCHK(hipMemcpyAsync(A.s0[0], A.ck_sh[(size_t)c], 4 * np * 8, hipMemcpyDeviceToDevice, ctx->stream));
        {
            const AdeState &x = A.ck_x[(size_t)c];
            const size_t nk = (size_t)ctx->nB * ctx->K * 8, nl = (size_t)ctx->nB * ctx->L * 8;
            HIPCHK(hipMemcpyAsync(ctx->phi[1], x.phi, nk, hipMemcpyDeviceToDevice, ctx->stream));
            HIPCHK(hipMemcpyAsync(ctx->psi1[1], x.p1, nl, hipMemcpyDeviceToDevice, ctx->stream));
            HIPCHK(hipMemcpyAsync(ctx->psi2[1], x.p2, nl, hipMemcpyDeviceToDevice, ctx->stream));
        }
        if ((rc = copy_pml(ctx, A, c, false))) return rc;
        // recompute the segment into the slots
        for (long long q = 0; q < len; ++q) {
            if ((rc = copy_ade(ctx, A.xs[(size_t)q], ctx->phi[1], ctx->psi1[1], ctx->psi2[1]))) return rc;
            const FieldSet<S> sets[4] = {fs_of<S>(A.v0, (size_t)q), fs_of<S>(A.v1, (size_t)q),
                                         fs_of<S>(A.v2, (size_t)q), fs_of<S>(A.v0, (size_t)q + 1)};
            if (kHalf && q > 0) {
                // the forward's rescale at the start of this step: the same exact power-of-two
                // factors on the step's start set (the checkpoint of a segment's first step
                // was taken after it)
                const int *pe = &A.fexp[4 * (size_t)(n0 + q - 1 - A.step0)], *ce = &A.fexp[4 * (size_t)(n0 + q - A.step0)];
                if (!std::equal(pe, pe + 4, ce)) {
                    Scl<float> fac;
                    for (int f = 0; f < 4; ++f) {
                        fac.s[f] = (float)std::ldexp(1.0, pe[f] - ce[f]);
                        fac.is[f] = 1.0f / fac.s[f];
                    }
                    const int nb = (int)std::min<long long>(4096, (ctx->N + 255) / 256);
                    hipLaunchKernelGGL((k_rescale<S>), dim3(nb), dim3(256), 0, ctx->stream, sets[0], ctx->N, fac);
                }
            }
            step_scales(n0 + q);
            double *const sh[4] = {A.s0[(size_t)q], np ? A.s1[(size_t)q] : nullptr, np ? A.s2[(size_t)q] : nullptr,
                                   A.s0[(size_t)q + 1]};
            if ((rc = launch_stages<S, R>(ctx, n0 + q, sets, sh, nullptr))) return rc;
            std::swap(ctx->phi[0], ctx->phi[1]);
            std::swap(ctx->psi1[0], ctx->psi1[1]);
            std::swap(ctx->psi2[0], ctx->psi2[1]);
        }
        // adjoint steps of the segment, last first
        for (long long q = len - 1; q >= 0; --q) {
            const long long n = n0 + q;
            step_scales(n);
            const FieldSet<S> V0 = fs_of<S>(A.v0, (size_t)q), V1 = fs_of<S>(A.v1, (size_t)q),
                              V2 = fs_of<S>(A.v2, (size_t)q);
            const AdeState &X0 = A.xs[(size_t)q];
            if (ctx->n_rec > 0)
                hipLaunchKernelGGL((k_receiver_adj<R>), dim3(1), dim3(64), 0, ctx->stream, (R *)A.lam[iNext][0],
                                   ctx->d_rcell, ctx->n_rec, A.dj, nsteps, n - A.step0, pidx, A.lam_sh[iNext]);
            if (ctx->nB > 0)
                hipLaunchKernelGGL((k_ade_replay<S>), bgrid, dim3(sb), 0, ctx->stream, g, V0, V1, bd, ade, P.rho, P.c,
                                   P.dt, X0, A.rx1, A.rx2, scd);
            // stage table: (input set, output adjoint, input adjoint, ADE right state, lnew, lr,
            //               a, b, c, RK alpha / beta, source weight and sample)
            struct St {
                FieldSet<S> vin; const double *vsh; int lout, lin; const AdeState *xr; int lnew, lr;
                double a, b, c, alpha, beta, w; long long nn;
            };
            const double *sh0 = np ? A.s0[(size_t)q] : nullptr, *sh1 = np ? A.s1[(size_t)q] : nullptr,
                         *sh2 = np ? A.s2[(size_t)q] : nullptr;
            const St st[3] = {
                {V0, sh0, iV1, iU0, &X0, ix1, ix0, 0.0, 1.0, 1.0, 0.0, 1.0, 1.0, 2 * n},
                {V1, sh1, iV2, iV1, &A.rx1, ix2, ix1, 3.0, 1.0, 4.0, 0.75, 0.25, 0.25, 2 * n + 2},
                {V2, sh2, iNext, iV2, &A.rx2, ix3, ix2, 1.0, 2.0, 3.0, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, 2 * n + 1},
            };
            for (int s = 2; s >= 0; --s) {
                const St &t = st[s];
                const FieldSet<R> Lo = lam_set<R>(A, t.lout), Li = lam_set<R>(A, t.lin), Lb = lam_set<R>(A, iU0);
                // first writers of the step's ADE adjoints: stage 2 sets x0 (mid) and x2 (r),
                // stage 1 sets x1 (r), stage 0 adds to x0
                if (ctx->nB > 0) {
                    if (s == 2)
                        hipLaunchKernelGGL((k_boundary_adj<S, R, true, true>), bgrid, dim3(sb), 0, ctx->stream, g, t.vin,
                                           bd, ade, P.rho, P.c, P.dt, t.a, t.b, 1.0 / t.c, X0, *t.xr, A.lx[t.lnew],
                                           A.lx[ix0], A.lx[t.lr], Lo.p, Lo.u, Lo.v, Lo.w, A.lD, A.G, A.P, scd);
                    else if (s == 1)
                        hipLaunchKernelGGL((k_boundary_adj<S, R, false, true>), bgrid, dim3(sb), 0, ctx->stream, g,
                                           t.vin, bd, ade, P.rho, P.c, P.dt, t.a, t.b, 1.0 / t.c, X0, *t.xr, A.lx[t.lnew],
                                           A.lx[ix0], A.lx[t.lr], Lo.p, Lo.u, Lo.v, Lo.w, A.lD, A.G, A.P, scd);
                    else
                        hipLaunchKernelGGL((k_boundary_adj<S, R, false, false>), bgrid, dim3(sb), 0, ctx->stream, g,
                                           t.vin, bd, ade, P.rho, P.c, P.dt, t.a, t.b, 1.0 / t.c, X0, *t.xr, A.lx[t.lnew],
                                           A.lx[ix0], A.lx[t.lr], Lo.p, Lo.u, Lo.v, Lo.w, A.lD, A.G, A.P, scd);
                }
                if (ctx->n_src > 0)
                    hipLaunchKernelGGL((k_source_adj<R>), dim3((ctx->n_src + 63) / 64), dim3(64), 0, ctx->stream, Lo.p,
                                       ctx->d_scell, ctx->d_sval, ctx->d_us,
                                       ctx->src_opt == FASES_SRC_SINE ? ctx->n_samp : 0LL, ctx->n_src, t.nn,
                                       t.w * coef0, A.gs, gus, pidx, A.lam_sh[t.lout]);
                const AdjK<R> K{kp, kc, kr, (R)(t.beta * P.dt / P.h), (R)t.alpha, (R)t.beta};
                if (tiled) {
                    if (s == 2)
                        hipLaunchKernelGGL((k_flux_adj_tiled<S, R, false, false, true, kAdjRowsOf<R>>), agrid, dim3(64, kAdjRowsOf<R>), 0,
                                           ctx->stream, ctx->Nx, ctx->Ny, ctx->Nz, t.vin, sc, Lo.p, Lo.u, Lo.v, Lo.w,
                                           ctx->d_upd, Li.p, Li.u, Li.v, Li.w, Lb.p, Lb.u, Lb.v, Lb.w, K, atz, anbx, anby);
                    else if (s == 1)
                        hipLaunchKernelGGL((k_flux_adj_tiled<S, R, false, true, true, kAdjRowsOf<R>>), agrid, dim3(64, kAdjRowsOf<R>), 0,
                                           ctx->stream, ctx->Nx, ctx->Ny, ctx->Nz, t.vin, sc, Lo.p, Lo.u, Lo.v, Lo.w,
                                           ctx->d_upd, Li.p, Li.u, Li.v, Li.w, Lb.p, Lb.u, Lb.v, Lb.w, K, atz, anbx, anby);
                    else
                        hipLaunchKernelGGL((k_flux_adj_tiled<S, R, true, false, false, kAdjRowsOf<R>>), agrid, dim3(64, kAdjRowsOf<R>), 0,
                                           ctx->stream, ctx->Nx, ctx->Ny, ctx->Nz, t.vin, sc, Lo.p, Lo.u, Lo.v, Lo.w,
                                           ctx->d_upd, Li.p, Li.u, Li.v, Li.w, Lb.p, Lb.u, Lb.v, Lb.w, K, atz, anbx, anby);
                } else if (s == 2)
                    hipLaunchKernelGGL((k_flux_adj_cell<S, R, false, false, true>), fgrid, fb, 0, ctx->stream, g,
                                       t.vin, sc, Lo.p, Lo.u, Lo.v, Lo.w, ctx->d_upd, Li.p, Li.u, Li.v, Li.w, Lb.p,
                                       Lb.u, Lb.v, Lb.w, K);
                else if (s == 1)
                    hipLaunchKernelGGL((k_flux_adj_cell<S, R, false, true, true>), fgrid, fb, 0, ctx->stream, g,
                                       t.vin, sc, Lo.p, Lo.u, Lo.v, Lo.w, ctx->d_upd, Li.p, Li.u, Li.v, Li.w, Lb.p,
                                       Lb.u, Lb.v, Lb.w, K);
                else
                    hipLaunchKernelGGL((k_flux_adj_cell<S, R, true, false, false>), fgrid, fb, 0, ctx->stream, g,
                                       t.vin, sc, Lo.p, Lo.u, Lo.v, Lo.w, ctx->d_upd, Li.p, Li.u, Li.v, Li.w, Lb.p,
                                       Lb.u, Lb.v, Lb.w, K);
                if (nap) {
                    const AdjK<double> Kd{0.5 * P.rho * P.c * P.c, 0.5 * P.c, 0.5 / P.rho, t.beta * P.dt / P.h,
                                          t.alpha, t.beta};
                    const dim3 pg((unsigned)((nap + 63) / 64));
                    const dim3 qg((unsigned)((3LL * A.nD + 63) / 64));
                    const PatchAdjIn<S> pin{t.vin, t.vsh, ctx->d_pidx, (int)np, scd};
                    hipLaunchKernelGGL((k_flux_adj_patch_part<S, R>), qg, dim3(64), 0, ctx->stream, g, pin, Lo.p,
                                       Lo.u, Lo.v, Lo.w, A.lam_sh[t.lout], ctx->d_upd, A.d_cells, A.nD, A.ap_idx, nap,
                                       Kd, A.pv);
                    if (s == 2)
                        hipLaunchKernelGGL((k_flux_adj_patch<S, R, false, false, true>), pg, dim3(64), 0, ctx->stream,
                                           g, Lo.p, Lo.u, Lo.v, Lo.w, A.lam_sh[t.lout], A.ap_cells, A.ap_idx, nap,
                                           A.d_map, A.nD, A.pv, Li.p, Li.u, Li.v, Li.w, A.lam_sh[t.lin], Lb.p, Lb.u,
                                           Lb.v, Lb.w, A.lam_sh[iU0], Kd);
                    else if (s == 1)
                        hipLaunchKernelGGL((k_flux_adj_patch<S, R, false, true, true>), pg, dim3(64), 0, ctx->stream,
                                           g, Lo.p, Lo.u, Lo.v, Lo.w, A.lam_sh[t.lout], A.ap_cells, A.ap_idx, nap,
                                           A.d_map, A.nD, A.pv, Li.p, Li.u, Li.v, Li.w, A.lam_sh[t.lin], Lb.p, Lb.u,
                                           Lb.v, Lb.w, A.lam_sh[iU0], Kd);
                    else
                        hipLaunchKernelGGL((k_flux_adj_patch<S, R, true, false, false>), pg, dim3(64), 0, ctx->stream,
                                           g, Lo.p, Lo.u, Lo.v, Lo.w, A.lam_sh[t.lout], A.ap_cells, A.ap_idx, nap,
                                           A.d_map, A.nD, A.pv, Li.p, Li.u, Li.v, Li.w, A.lam_sh[t.lin], Lb.p, Lb.u,
                                           Lb.v, Lb.w, A.lam_sh[iU0], Kd);
                }
                if (ctx->pml.on) {
                    const double *lso = nap ? A.lam_sh[t.lout] : nullptr;
                    double *lsi = nap ? A.lam_sh[t.lin] : nullptr;
                    if (s == 2)
                        launch_pml_adj<S, R, true>(ctx, A, g, t.vin, sc, Lo, Li, pidx, lso, lsi, K, t.beta);
                    else
                        launch_pml_adj<S, R, false>(ctx, A, g, t.vin, sc, Lo, Li, pidx, lso, lsi, K, t.beta);
                }
                if (ctx->nB > 0)
                    hipLaunchKernelGGL((k_probe_gather<R>), tgrid, dim3(sb), 0, ctx->stream, A.nT, A.tgt, A.toff, A.eb,
                                       A.ew, A.lD, Li.p, Li.u, Li.v, Li.w, pidx, A.lam_sh[t.lin], nap);
            }
            std::swap(iNext, iU0);
            std::swap(ix3, ix0);
        }
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipGetLastError());
    std::copy(fexp_end, fexp_end + 4, ctx->fexp);
    // material gradients: sum the per-cell contributions by material
    if (grad_mat) {
        std::vector<double> Gh((size_t)ctx->nB * A.P);
        std::vector<int> mat((size_t)ctx->nB);
        HIPCHK(hipMemcpy(Gh.data(), A.G, Gh.size() * 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(mat.data(), ctx->d_bmat, mat.size() * 4, hipMemcpyDeviceToHost));
        std::fill(grad_mat, grad_mat + (size_t)ctx->nMat * A.P, 0.0);
        for (int b = 0; b < ctx->nB; ++b) {
            double *o = grad_mat + (size_t)(mat[(size_t)b] & 0xFF) * A.P;
            for (int p = 0; p < A.P; ++p) o[p] += Gh[(size_t)p * ctx->nB + b];
        }
    }
    if (grad_sval && ctx->n_src > 0) HIPCHK(hipMemcpy(grad_sval, A.gs, (size_t)ctx->n_src * 8, hipMemcpyDeviceToHost));
    if (grad_us && gus) HIPCHK(hipMemcpy(grad_us, gus, (size_t)gus_n * 8, hipMemcpyDeviceToHost));
    return FASES_OK;
}

}  // namespace

extern "C" {

int fases_adjoint_forward(fases_ctx *ctx, int64_t nsteps, int checkpoint_every, double *ir) {
    if (!ctx) return FASES_E_ARG;
    if (!ctx->have_geo || !ctx->have_mat)
        return set_error(ctx, FASES_E_STATE, "fases_adjoint_forward: geometry and materials must be set first");
    if (nsteps < 1 || checkpoint_every < 1)
        return set_error(ctx, FASES_E_ARG, "fases_adjoint_forward: nsteps and checkpoint_every must be >= 1");
    HIPCHK(hipSetDevice(ctx->prm.device));
    if (ctx->prm.precision == FASES_FP64) return forward_impl<double, double>(ctx, nsteps, checkpoint_every, ir);
    if (ctx->prm.precision == FASES_FP16) return forward_impl<__half, float>(ctx, nsteps, checkpoint_every, ir);
    return forward_impl<float, float>(ctx, nsteps, checkpoint_every, ir);
}

int fases_adjoint(fases_ctx *ctx, const double *dJ_dir, double *grad_mat, double *grad_sval, double *grad_us) {
    if (!ctx) return FASES_E_ARG;
    if (!ctx->adj) return set_error(ctx, FASES_E_STATE, "fases_adjoint: run fases_adjoint_forward first");
    if (ctx->n_rec > 0 && !dJ_dir) return set_error(ctx, FASES_E_ARG, "fases_adjoint: dJ_dir is NULL");
    HIPCHK(hipSetDevice(ctx->prm.device));
    if (ctx->prm.precision == FASES_FP64) return backward_impl<double, double>(ctx, dJ_dir, grad_mat, grad_sval, grad_us);
    if (ctx->prm.precision == FASES_FP16) return backward_impl<__half, float>(ctx, dJ_dir, grad_mat, grad_sval, grad_us);
    return backward_impl<float, float>(ctx, dJ_dir, grad_mat, grad_sval, grad_us);
}

}  // extern "C"
