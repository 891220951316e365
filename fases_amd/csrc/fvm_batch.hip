// Batched time loop: many independent solver contexts (jobs) on one device advanced
// together. Every stage launches each kernel kind once for all jobs (the block / thread
// index selects the job through a prefix table), so small rooms that would leave most
// CUs idle on their own fill the chip together; the per-step kernel sequence is
// captured into a hipGraph (two steps, one per rotation parity) and replayed.
//
// The per-job arithmetic is exactly the single-context path's (same device functions:
// flux_tile, face_flux_pf, boundary_cell), so a job gives the same IR batched or alone.
// Jobs may differ in grid, materials, sources, receivers and step count; jobs that
// finish drop out at phase boundaries (the step counts sorted ascending).

#include "fvm_ctx.hpp"

#include <cstring>
#include <numeric>

namespace {

// Physical field set of logical set L (0 = U^n, 1, 2) at step parity p: the rotation
// U0 <-> U1 after every step (main3D.cu:1312-1319) leaves set 2 in place.
__host__ __device__ inline int phys(int L, int p) { return L < 2 ? (L ^ p) : 2; }

// Job index of global item t given ascending offsets off[0..nj-1] (off[0] = 0).
// (the last j with off[j] <= t, so jobs without items are skipped), by bisection: a
// handful of dependent loads at the start of a block or thread instead of up to nj.
__device__ inline int job_of(const int *__restrict__ off, int nj, int t) {
    int lo = 0, hi = nj - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename S> struct FluxJob {
    int Nx, Ny, Nz, tz, nbx, nby;
    S *U[3][4];
    const uint8_t *upd;
    PmlSet pml;
    TileK<float> K;
    float dt;
    Scl<float> sc;
};

// Rows (waves) per block of the batched flux kernel: 8-wave blocks, two per CU, overlap
// each other's barriers (the faster shape on the 512x384x256 room and on the batch rooms).
constexpr int kBatchRows = 8;
// Tile width of a job in the batched flux kernel: the width its own autotuned run uses
// (62 or 64 columns; results do not depend on the width). Jobs of each width share one
// launch per stage.
constexpr int kBatchWidths[2] = {62, 64};

// PF: the cPML fused in the tile (some job of the phase has a band). Without it the tile
// is held to 80 VGPRs, three blocks per CU, like the single-context kernel.
template <typename S, int STAGE, bool PF, int TX>
__global__ void __launch_bounds__(64 * kBatchRows) __attribute__((amdgpu_waves_per_eu(PF ? 1 : 6, 8)))
k_flux_batch(const FluxJob<S> *__restrict__ jobs, const int *__restrict__ boff, int nj, int parity) {
    __shared__ TileLds<float, kBatchRows, TX> L;
    const int j = job_of(boff, nj, blockIdx.x);
    const FluxJob<S> &J = jobs[j];
    const int lb = blockIdx.x - boff[j];
    const int bx = lb % J.nbx, by = (lb / J.nbx) % J.nby, bz = lb / (J.nbx * J.nby);
    constexpr int in = STAGE == 0 ? 0 : (STAGE == 1 ? 1 : 2), out = STAGE == 1 ? 2 : 1;
    const int pi = phys(in, parity), pb = phys(0, parity), po = phys(out, parity);
    const TileFields<S> F{J.U[pi][0], J.U[pi][1], J.U[pi][2], J.U[pi][3], J.U[pb][0], J.U[pb][1],
                          J.U[pb][2], J.U[pb][3], J.U[po][0], J.U[po][1], J.U[po][2], J.U[po][3]};
    const TileK<float> K = J.K;
    const Scl<float> sc = J.sc;
    flux_tile<S, float, STAGE, kBatchRows, TX, PF>(L, J.Nx, J.Ny, J.Nz, F, J.upd, &J.pml, K, J.dt, sc, J.tz, bx, by, bz);
}

template <typename S> struct PatchJob {
    Grid g;
    S *U[3][4];
    double *sh[3];
    const int *pidx, *cells;
    int np;
    PmlSet pml;
    double rho, c, h, dt;
    Scl<double> sc;
};

// Each job's items start on a wave boundary in the patch and boundary tables (offsets
// padded to 64), so a wave's job is wave-uniform: its descriptor comes through scalar loads
// and stays in scalar registers.
constexpr int kWaveItems = 64;
inline int pad_wave(int n) { return (n + kWaveItems - 1) / kWaveItems * kWaveItems; }

template <typename S, int STAGE>
__global__ void __launch_bounds__(64)
k_patch_batch(const PatchJob<S> *__restrict__ jobs, const int *__restrict__ off, int nj, int parity) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= off[nj]) return;
    const int j = __builtin_amdgcn_readfirstlane(job_of(off, nj, t));
    const PatchJob<S> &J = jobs[j];
    const int lt = t - off[j];
    if (lt >= J.np) return;
    constexpr int in = STAGE == 0 ? 0 : (STAGE == 1 ? 1 : 2), out = STAGE == 1 ? 2 : 1;
    const int pi = phys(in, parity), pb = phys(0, parity), po = phys(out, parity);
    auto fs = [&](int q) { return FieldSet<S>{J.U[q][0], J.U[q][1], J.U[q][2], J.U[q][3]}; };
    const PatchIO<S> io{fs(pi), fs(pb), fs(po), J.pidx, J.sh[pi], J.sh[pb], J.sh[po], J.np, J.sc};
    const int C = J.cells[lt];
    const int i = C % J.g.Nx, jj = (C / J.g.Nx) % J.g.Ny, k = (int)(C / J.g.nxy);
    patch_cell<S, STAGE>(J.g, i, jj, k, io, J.pml, J.rho, J.c, J.h, J.dt);
}

template <typename S> struct SrcJob {
    S *p[3];                  // pressure of each physical set
    double *sh[3];            // FP64 shadows of each physical set (patch cells), or null
    const int *cell, *pidx;
    const double *sval, *u_s;
    long long us_stride, step0;
    double dt, c, rho, As, h;
    float s, is;
};

// functions3D.cuh:1039-1047 for every source of every job; the step index comes from the
// device counter so a captured graph replays it.
template <typename S>
__global__ void k_sources_batch(const SrcJob<S> *__restrict__ jobs, const int *__restrict__ off, int nj, int stage,
                                int parity, const long long *__restrict__ d_n) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= off[nj]) return;
    const int j = job_of(off, nj, t);
    const SrcJob<S> &J = jobs[j];
    const int s = t - off[j];
    const long long n = J.step0 + *d_n;
    const long long nn = stage == 0 ? 2 * n : (stage == 1 ? 2 * n + 2 : 2 * n + 1);
    const double w = stage == 0 ? 1.0 : (stage == 1 ? 0.25 : 2.0 / 3.0);
    const int po = phys(stage == 1 ? 2 : 1, parity);
    const double amp = J.u_s[(long long)s * J.us_stride + nn];
    const double add = w * J.dt * J.c * J.c * J.sval[s] * amp * J.rho * J.As / (J.h * J.h * J.h);
    const int C = J.cell[s];
    S *p = J.p[po];
    const int q = J.pidx ? J.pidx[C] : -1;
    if (q >= 0) {
        const double v = J.sh[po][q] + add;
        J.sh[po][q] = v;
        p[C] = Conv<S, double>::st(v, (double)J.is);
    } else {
        const float cur = Conv<S, float>::ld(p[C], J.s);
        p[C] = Conv<S, float>::st((float)((double)cur + add), J.is);
    }
}

template <typename S> struct BndJob {
    Grid g;
    S *U[3][4];
    Bnd bd;
    Ade ade;
    double *phi[3], *psi1[3], *psi2[3];
    double rho, c, dt;
    Scl<double> sc;
    AdeRk rk;                 // stage-collapsed ADE (reduced-precision storage)
    double *aux;
};

template <typename S, int STAGE>
__global__ void __launch_bounds__(256)
k_boundary_batch(const BndJob<S> *__restrict__ jobs, const int *__restrict__ off, int nj, int parity) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= off[nj]) return;
    const int j = __builtin_amdgcn_readfirstlane(job_of(off, nj, t));
    const BndJob<S> &J = jobs[j];
    if (t - off[j] >= J.bd.n) return;
    // (set in, set out, ADE a, b, c, buffers left / mid / right) per stage (main3D.cu:1290-1306)
    const int in = STAGE, out = STAGE == 1 ? 2 : 1;
    const double a = STAGE == 0 ? 0.0 : (STAGE == 1 ? 3.0 : 1.0);
    const double b = STAGE == 2 ? 2.0 : 1.0;
    const double cc = STAGE == 0 ? 1.0 : (STAGE == 1 ? 4.0 : 3.0);
    const int l = STAGE == 1 ? 2 : 0, r = STAGE == 0 ? 1 : (STAGE == 1 ? 0 : 2);
    const int pi = phys(in, parity), po = phys(out, parity);
    const int pl = phys(l, parity), pm = phys(1, parity), pr = phys(r, parity);
    const FieldSet<S> fi{J.U[pi][0], J.U[pi][1], J.U[pi][2], J.U[pi][3]};
    const FieldSet<S> fo{J.U[po][0], J.U[po][1], J.U[po][2], J.U[po][3]};
    if constexpr (std::is_same<S, double>::value) {
        const AdeBufs ab{J.phi[pl], J.phi[pm], J.phi[pr], J.psi1[pl], J.psi1[pm], J.psi1[pr],
                         J.psi2[pl], J.psi2[pm], J.psi2[pr]};
        boundary_cell<S>(t - off[j], J.g, fi, fo, J.bd, J.ade, J.rho, J.c, J.dt, a, b, cc, ab, J.sc);
    } else {
        // the single-context loop's stage-collapsed ADE: x0 = phi[mid], x3 -> phi[logical 0]
        const int p0 = phys(0, parity);
        const AdeRkBufs rb{J.phi[pm], J.psi1[pm], J.psi2[pm], J.phi[p0], J.psi1[p0], J.psi2[p0], J.aux};
        const double Z0 = J.rho * J.c, iz = 1.0 / (2.0 * Z0);
        boundary_cell_rk<S, STAGE>(t - off[j], J.g, fi, fo, J.bd, J.rk, Z0, iz, rb, J.sc);
    }
}

template <typename S> struct RecJob {
    S *U[3][4];
    const int *cell;
    double *ir;
    long long stride;
    Scl<double> sc;
};

// After step n (parity p) the new U^{n+1} is physical set phys(0, p ^ 1).
template <typename S>
__global__ void k_receivers_batch(const RecJob<S> *__restrict__ jobs, const int *__restrict__ off, int nj,
                                  int parity, const long long *__restrict__ d_n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= off[nj]) return;
    const int j = job_of(off, nj, t);
    const RecJob<S> &J = jobs[j];
    const int r = t - off[j];
    const int q = phys(0, parity ^ 1);
    J.ir[r * J.stride + *d_n] = (double)Conv<S, float>::ld(J.U[q][0][J.cell[r]], (float)J.sc.s[0]);
}

__global__ void k_step_advance(long long *d_n) {
    if (threadIdx.x == 0) *d_n += 1;
}

// Device tables of one phase (the jobs still running).
template <typename S> struct Phase {
    int npatch = 0, nsrc = 0, nbnd = 0, nrec = 0;
    bool any_pml = false;                // some job has a cPML band: the fused tile
    // flux jobs per tile width (kBatchWidths): job tables, block offsets, counts
    int fnj[2] = {0, 0}, fnblk[2] = {0, 0};
    FluxJob<S> *fjobs[2] = {nullptr, nullptr};
    int *fboff[2] = {nullptr, nullptr};
    PatchJob<S> *patch = nullptr;
    SrcJob<S> *src = nullptr;
    BndJob<S> *bnd = nullptr;
    RecJob<S> *rec = nullptr;
    int *poff = nullptr, *soff = nullptr, *doff = nullptr, *roff = nullptr;
};

struct Batch {
    fases_ctx *own;                      // error sink and allocation owner (jobs[0])
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    std::vector<hipGraphExec_t> execs;   // captured step pairs, one per phase
    std::vector<hipGraph_t> graphs;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // Every exit, an error return included, drains the stream before anything its queued
    // replays use (graphs, execs, phase tables) is released.
    ~Batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &x : execs) (void)hipGraphExecDestroy(x);
        for (auto &g : graphs) (void)hipGraphDestroy(g);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        for (void *p : allocs) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
    template <typename T> int put(const std::vector<T> &v, T **out) {
        fases_ctx *ctx = own;
        *out = nullptr;
        void *p = nullptr;
        HIPCHK(hipMalloc(&p, std::max<size_t>(1, v.size()) * sizeof(T)));
        allocs.push_back(p);
        if (!v.empty()) HIPCHK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));        // callers pass short-lived host tables
        *out = static_cast<T *>(p);
        return FASES_OK;
    }
};

template <typename S>
int build_phase(Batch &bt, const std::vector<fases_ctx *> &jobs, const std::vector<long long> &nsteps, Phase<S> &ph) {
    std::vector<FluxJob<S>> fj[2];
    std::vector<PatchJob<S>> pj;
    std::vector<SrcJob<S>> sj;
    std::vector<BndJob<S>> bj;
    std::vector<RecJob<S>> rj;
    std::vector<int> boff[2] = {{0}, {0}}, poff{0}, soff{0}, doff{0}, roff{0};
    for (size_t q = 0; q < jobs.size(); ++q) {
        fases_ctx *c = jobs[q];
        const fases_params &P = c->prm;
        const Scl<float> scf = field_scales<float>(c);
        const Scl<double> scd = field_scales<double>(c);
        const Grid g{c->Nx, c->Ny, c->Nz, (long long)c->Nx * c->Ny};
        FluxJob<S> f{};
        const int w = c->tx == kBatchWidths[0] ? 0 : 1;
        f.Nx = c->Nx; f.Ny = c->Ny; f.Nz = c->Nz; f.tz = c->tz;
        f.nbx = (c->Nx - 6 + kBatchWidths[w] - 1) / kBatchWidths[w];
        f.nby = (c->Ny - 6 + kBatchRows - 1) / kBatchRows;
        const int nbz = (c->Nz - 6 + c->tz - 1) / c->tz;
        for (int s = 0; s < 3; ++s)
            for (int v = 0; v < 4; ++v) f.U[s][v] = (S *)c->U[s][v];
        f.upd = c->d_upd; f.pml = c->pml; f.K = tile_constants<float>(P); f.dt = (float)P.dt; f.sc = scf;
        ph.any_pml = ph.any_pml || c->pml.on;
        fj[w].push_back(f);
        boff[w].push_back(boff[w].back() + f.nbx * f.nby * nbz);
        if (c->n_patch > 0) {
            PatchJob<S> p{};
            p.g = g;
            std::memcpy(p.U, f.U, sizeof p.U);
            for (int s = 0; s < 3; ++s) p.sh[s] = c->shadow[s];
            p.pidx = c->d_pidx; p.cells = c->d_patch; p.np = c->n_patch; p.pml = c->pml;
            p.rho = P.rho; p.c = P.c; p.h = P.h; p.dt = P.dt; p.sc = scd;
            pj.push_back(p);
            poff.push_back(poff.back() + pad_wave(c->n_patch));
        }
        if (c->n_src > 0) {
            SrcJob<S> s{};
            for (int k = 0; k < 3; ++k) { s.p[k] = (S *)c->U[k][0]; s.sh[k] = c->n_patch > 0 ? c->shadow[k] : nullptr; }
            s.cell = c->d_scell; s.pidx = c->n_patch > 0 ? c->d_pidx : nullptr;
            s.sval = c->d_sval; s.u_s = c->d_us;
            s.us_stride = c->src_opt == FASES_SRC_SINE ? c->n_samp : 0;
            s.step0 = c->step;
            s.dt = P.dt; s.c = P.c; s.rho = P.rho; s.As = 6.0 * P.h * P.h; s.h = P.h;
            s.s = scf.s[0]; s.is = scf.is[0];
            sj.push_back(s);
            soff.push_back(soff.back() + c->n_src);
        }
        if (c->nB > 0) {
            BndJob<S> b{};
            b.g = g;
            std::memcpy(b.U, f.U, sizeof b.U);
            b.bd = bnd_of(c);
            b.ade = Ade{c->K, c->L, c->d_Rinf, c->d_A, c->d_zeta, c->d_B, c->d_C, c->d_alpha, c->d_beta,
                        c->d_Keff, c->d_Leff};
            for (int k = 0; k < 3; ++k) { b.phi[k] = c->phi[k]; b.psi1[k] = c->psi1[k]; b.psi2[k] = c->psi2[k]; }
            b.rho = P.rho; b.c = P.c; b.dt = P.dt; b.sc = scd;
            b.rk = ade_rk_of(c);
            b.aux = c->d_ade_aux;
            bj.push_back(b);
            doff.push_back(doff.back() + pad_wave(c->nB));
        }
        if (c->n_rec > 0) {
            RecJob<S> r{};
            std::memcpy(r.U, f.U, sizeof r.U);
            r.cell = c->d_rcell; r.ir = c->d_ir; r.stride = nsteps[q]; r.sc = scd;
            rj.push_back(r);
            roff.push_back(roff.back() + c->n_rec);
        }
    }
    ph.npatch = poff.back(); ph.nsrc = soff.back(); ph.nbnd = doff.back(); ph.nrec = roff.back();
    int rc;
    for (int w = 0; w < 2; ++w) {
        ph.fnj[w] = (int)fj[w].size();
        ph.fnblk[w] = boff[w].back();
        if ((rc = bt.put(fj[w], &ph.fjobs[w])) || (rc = bt.put(boff[w], &ph.fboff[w]))) return rc;
    }
    if ((rc = bt.put(pj, &ph.patch)) || (rc = bt.put(poff, &ph.poff))) return rc;
    if ((rc = bt.put(sj, &ph.src)) || (rc = bt.put(soff, &ph.soff))) return rc;
    if ((rc = bt.put(bj, &ph.bnd)) || (rc = bt.put(doff, &ph.doff))) return rc;
    if ((rc = bt.put(rj, &ph.rec)) || (rc = bt.put(roff, &ph.roff))) return rc;
    return FASES_OK;
}

// The flux stage s of the phase's jobs of tile width TX (table w).
template <typename S, int TX>
void launch_flux_batch(hipStream_t st, const Phase<S> &ph, int w, int s, int parity) {
    const dim3 g(ph.fnblk[w]), b(64, kBatchRows);
    const FluxJob<S> *J = ph.fjobs[w];
    const int *off = ph.fboff[w];
    const int nj = ph.fnj[w];
    if (ph.any_pml) {
        if (s == 0) hipLaunchKernelGGL((k_flux_batch<S, 0, true, TX>), g, b, 0, st, J, off, nj, parity);
        else if (s == 1) hipLaunchKernelGGL((k_flux_batch<S, 1, true, TX>), g, b, 0, st, J, off, nj, parity);
        else hipLaunchKernelGGL((k_flux_batch<S, 2, true, TX>), g, b, 0, st, J, off, nj, parity);
    } else {
        if (s == 0) hipLaunchKernelGGL((k_flux_batch<S, 0, false, TX>), g, b, 0, st, J, off, nj, parity);
        else if (s == 1) hipLaunchKernelGGL((k_flux_batch<S, 1, false, TX>), g, b, 0, st, J, off, nj, parity);
        else hipLaunchKernelGGL((k_flux_batch<S, 2, false, TX>), g, b, 0, st, J, off, nj, parity);
    }
}

// Kernel sequence of one step of a phase at the given parity.
template <typename S>
void launch_batch_step(hipStream_t st, const Phase<S> &ph, int np_jobs, int ns_jobs, int nb_jobs, int nr_jobs,
                       int parity, long long *d_n) {
    for (int s = 0; s < 3; ++s) {
        if (ph.fnblk[0] > 0) launch_flux_batch<S, 62>(st, ph, 0, s, parity);
        if (ph.fnblk[1] > 0) launch_flux_batch<S, 64>(st, ph, 1, s, parity);
        if (ph.npatch > 0) {
            const dim3 g((ph.npatch + 63) / 64);
            if (s == 0)
                hipLaunchKernelGGL((k_patch_batch<S, 0>), g, dim3(64), 0, st, ph.patch, ph.poff, np_jobs, parity);
            else if (s == 1)
                hipLaunchKernelGGL((k_patch_batch<S, 1>), g, dim3(64), 0, st, ph.patch, ph.poff, np_jobs, parity);
            else
                hipLaunchKernelGGL((k_patch_batch<S, 2>), g, dim3(64), 0, st, ph.patch, ph.poff, np_jobs, parity);
        }
        if (ph.nsrc > 0)
            hipLaunchKernelGGL((k_sources_batch<S>), dim3((ph.nsrc + 63) / 64), dim3(64), 0, st, ph.src, ph.soff,
                               ns_jobs, s, parity, (const long long *)d_n);
        if (ph.nbnd > 0) {
            const dim3 g((ph.nbnd + 255) / 256);
            if (s == 0)
                hipLaunchKernelGGL((k_boundary_batch<S, 0>), g, dim3(256), 0, st, ph.bnd, ph.doff, nb_jobs, parity);
            else if (s == 1)
                hipLaunchKernelGGL((k_boundary_batch<S, 1>), g, dim3(256), 0, st, ph.bnd, ph.doff, nb_jobs, parity);
            else
                hipLaunchKernelGGL((k_boundary_batch<S, 2>), g, dim3(256), 0, st, ph.bnd, ph.doff, nb_jobs, parity);
        }
    }
    if (ph.nrec > 0)
        hipLaunchKernelGGL((k_receivers_batch<S>), dim3((ph.nrec + 63) / 64), dim3(64), 0, st, ph.rec, ph.roff,
                           nr_jobs, parity, (const long long *)d_n);
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, st, d_n);
}

template <typename S>
int batch_impl(fases_ctx **ctxs, int n, const int64_t *nsteps, double **ir, double *elapsed_ms) {
    fases_ctx *ctx = ctxs[0];
    // per-job preparation on each job's own stream: waveform tables, receiver buffers, tile
    for (int q = 0; q < n; ++q) {
        fases_ctx *c = ctxs[q];
        int rc = prepare_run(c, nsteps[q]);
        if (rc) { if (c != ctx) ctx->err = c->err; return rc; }
        if (c->tz == 0) {
            rc = autotune_tile<S, float>(c);
            if (rc) { if (c != ctx) ctx->err = c->err; return rc; }
        }
        {
            fases_ctx *ctx = c;          // HIPCHK reports into the job
            HIPCHK(hipStreamSynchronize(c->stream));
        }
    }
    Batch bt;
    bt.own = ctx;
    HIPCHK(hipStreamCreateWithFlags(&bt.stream, hipStreamNonBlocking));
    long long *d_n = nullptr;
    {
        std::vector<long long> zero{0};
        int rc = bt.put(zero, &d_n);
        if (rc) return rc;
    }
    // phases: ascending distinct step counts
    std::vector<int> order((size_t)n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return nsteps[a] < nsteps[b]; });
    HIPCHK(hipEventCreate(&bt.e0));
    HIPCHK(hipEventCreate(&bt.e1));
    HIPCHK(hipEventRecord(bt.e0, bt.stream));
    long long done = 0;
    size_t first = 0;
    while (first < order.size()) {
        const long long until = nsteps[order[first]];
        if (until > done) {
            std::vector<fases_ctx *> act;
            std::vector<long long> act_steps;
            for (size_t q = first; q < order.size(); ++q) { act.push_back(ctxs[order[q]]); act_steps.push_back(nsteps[order[q]]); }
            // keep table order = the caller's order among the active jobs
            std::vector<size_t> idx(act.size());
            std::iota(idx.begin(), idx.end(), 0);
            std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return order[first + a] < order[first + b]; });
            std::vector<fases_ctx *> jobs;
            std::vector<long long> js;
            for (size_t q : idx) { jobs.push_back(act[q]); js.push_back(act_steps[q]); }
            Phase<S> ph;
            int rc = build_phase<S>(bt, jobs, js, ph);
            if (rc) return rc;
            int np_j = 0, ns_j = 0, nb_j = 0, nr_j = 0;
            for (fases_ctx *c : jobs) {
                np_j += c->n_patch > 0; ns_j += c->n_src > 0; nb_j += c->nB > 0; nr_j += c->n_rec > 0;
            }
            const long long len = until - done;
            const int par0 = (int)(done & 1);
            // capture one step per parity, replay the pair
            hipGraph_t g = nullptr;
            hipGraphExec_t x = nullptr;
            if (len >= 2) {
                HIPCHK(hipStreamBeginCapture(bt.stream, hipStreamCaptureModeThreadLocal));
                launch_batch_step<S>(bt.stream, ph, np_j, ns_j, nb_j, nr_j, par0, d_n);
                launch_batch_step<S>(bt.stream, ph, np_j, ns_j, nb_j, nr_j, par0 ^ 1, d_n);
                HIPCHK(hipStreamEndCapture(bt.stream, &g));
                bt.graphs.push_back(g);
                HIPCHK(hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
                bt.execs.push_back(x);
                for (long long r = 0; r < len / 2; ++r) HIPCHK(hipGraphLaunch(x, bt.stream));
            }
            if (len & 1) launch_batch_step<S>(bt.stream, ph, np_j, ns_j, nb_j, nr_j, (int)((until - 1) & 1), d_n);
            done = until;
        }
        while (first < order.size() && nsteps[order[first]] <= done) ++first;
    }
    HIPCHK(hipEventRecord(bt.e1, bt.stream));
    HIPCHK(hipEventSynchronize(bt.e1));
    HIPCHK(hipGetLastError());
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, bt.e0, bt.e1));
    if (elapsed_ms) *elapsed_ms = ms;
    // results and host-side rotation state of every job
    for (int q = 0; q < n; ++q) {
        fases_ctx *c = ctxs[q];
        const long long ns = nsteps[q];
        const long long need = (long long)c->n_rec * ns;
        if (need > 0 && ir && ir[q]) {
            HIPCHK(hipMemcpy(ir[q], c->d_ir, (size_t)need * sizeof(double), hipMemcpyDeviceToHost));
            for (long long i = 0; i < need; ++i)
                if (!std::isfinite(ir[q][i])) return set_error(ctx, FASES_E_NONFINITE, "receiver sample is not finite");
        }
        if (ns & 1) {
            for (int f = 0; f < 4; ++f) std::swap(c->U[0][f], c->U[1][f]);
            std::swap(c->phi[0], c->phi[1]);
            std::swap(c->psi1[0], c->psi1[1]);
            std::swap(c->psi2[0], c->psi2[1]);
            std::swap(c->shadow[0], c->shadow[1]);
        }
        c->step += ns;
    }
    return FASES_OK;
}

}  // namespace

extern "C" int fases_batch_run(fases_ctx **ctxs, int n, const int64_t *nsteps, double **ir, double *elapsed_ms) {
    if (!ctxs || n < 1 || !nsteps || !ctxs[0]) return set_error(nullptr, FASES_E_ARG, "fases_batch_run: bad argument");
    fases_ctx *ctx = ctxs[0];
    for (int q = 0; q < n; ++q) {
        fases_ctx *c = ctxs[q];
        if (!c) return set_error(ctx, FASES_E_ARG, "fases_batch_run: null context");
        if (!c->have_geo || !c->have_mat)
            return set_error(ctx, FASES_E_STATE, "fases_batch_run: every job needs geometry and materials");
        if (c->prm.precision != FASES_FP32)
            return set_error(ctx, FASES_E_ARG, "fases_batch_run: batched jobs use FP32 storage");
        if (c->prm.device != ctx->prm.device) return set_error(ctx, FASES_E_ARG, "fases_batch_run: jobs on different devices");
        if (nsteps[q] < 0) return set_error(ctx, FASES_E_ARG, "fases_batch_run: negative step count");
        for (int r = 0; r < q; ++r)
            if (ctxs[r] == c) return set_error(ctx, FASES_E_ARG, "fases_batch_run: a context appears twice");
    }
    HIPCHK(hipSetDevice(ctx->prm.device));
    return batch_impl<float>(ctxs, n, nsteps, ir, elapsed_ms);
}
