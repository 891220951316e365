/*
 * fases_core.h -- C ABI of the MI355X-native FASES 3D-solver time-stepping core.
 *
 * Plain C: pointers, sizes and ints only. Every function returns 0 on success
 * or a negative FASES_E_* code, and never exits the process. Host arrays passed
 * in belong to the caller: they are copied and never retained. Device memory
 * belongs to the context. A context is single-threaded and owns one HIP stream
 * on an explicit device.
 *
 * Reference interfaces replaced (vitorotri/fases, paths relative to the repo):
 *   scene     <- 3D-solver/include/parsing3D.h (parse_obj3d :690, build_sdf :860,
 *                build_oid_dense :1145) and common/filterABC.h (build_boundary_map :95)
 *   formats   <- 3D-solver/core/main3D.cu (config.txt :114-128, poles.txt :592-603,
 *                src.txt :861-906, rec.txt :919-953, waveforms :963-1124,
 *                IR/VEL files :1242-1256,1397-1414, header.txt :1441-1449) and
 *                common/filterABC.h read_fit_file :31-92
 *   solver    <- 3D-solver/core/main3D.cu time loop :1280-1320 driving
 *                3D-solver/core/functions3D.cuh kernel_WENO_stage{0,1,2} :1086-1327,
 *                kernel_inject_src_SSP :1039-1047, kernel_apply_BC_immersed_sdf
 *                :462-565 (ADE_solver :426-459, cPML :681-756)
 */
#ifndef FASES_CORE_H
#define FASES_CORE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FASES_OK        0
#define FASES_E_ARG    -1   /* bad argument / size                               */
#define FASES_E_IO     -2   /* file cannot be opened / written                   */
#define FASES_E_PARSE  -3   /* malformed input file                              */
#define FASES_E_HIP    -4   /* HIP runtime error                                 */
#define FASES_E_NOMEM  -5   /* host or device allocation failed                  */
#define FASES_E_STATE  -6   /* call out of order (e.g. run before geometry)      */
#define FASES_E_RANGE  -7   /* receiver / source outside the grid                */
#define FASES_E_NONFINITE -8 /* receiver sample became NaN/Inf                   */

/* storage precision of the p, vx, vy, vz fields */
#define FASES_FP64 0        /* fields, flux math, ADE states: FP64               */
#define FASES_FP32 1        /* fields + flux math FP32; ADE states, receivers FP64 */
#define FASES_FP16 2        /* fields FP16 with per-field power-of-two scale, math FP32;
                               ADE states, receivers FP64                          */

/* source waveform selection (config line 16, main3D.cu:966) */
#define FASES_SRC_PCS    0
#define FASES_SRC_RICKER 1
#define FASES_SRC_SINE   2

typedef struct fases_ctx   fases_ctx;
typedef struct fases_scene fases_scene;

typedef struct fases_params {
    int    device;          /* HIP device index                                        */
    int    precision;       /* FASES_FP64 / FASES_FP32 / FASES_FP16                    */
    double h;               /* cell size [m]                                           */
    double dt;              /* time step [s]  (reference: CFL*h/c, main3D.cu:730)      */
    double c;               /* speed of sound [m/s]                                    */
    double rho;             /* density [kg/m^3]  (reference: 1.225, main3D.cu:454)     */
    double tol;             /* immersed band thickness (reference: 3h, main3D.cu:224)  */
    float  pml_keys[6];     /* U(z+) D(z-) R(x+) L(x-) F(y+) B(y-): config lines 7..12 */
    int    pml_cells;       /* PML depth d (reference PML_CELLS = 20)                  */
    double pml_sigma_max;   /* <= 0: reference 0.5 c (m+1) ln(1/R0)/(d h), R0=1e-7,m=2 */
    double pml_kappa_max;   /* reference 5                                             */
    double pml_alpha_max;   /* reference 20 pi rad/s                                   */
    int    fp16_rescale_every; /* FP16: steps between per-field scale updates (100)    */
    int    use_graph;       /* reserved, ignored: fases_run launches stage by stage;
                               fases_batch_run always replays captured hipGraphs       */
    int    kernel_variant;  /* 0 auto: FP64 -> per-cell kernel, FP32/FP16 -> tiled;
                               1 per-cell kernel (reference operation order, no FMA
                               contraction: bit-faithful to an FP64 CPU evaluation);
                               2 tiled kernel (faces reconstructed once, LDS/shuffles) */
    int    tile_z;          /* tiled kernel: z planes per block; 0 = autotune at the
                               first fases_run (times 16, 32, 64, 128 on the real grid) */
    int    adjoint_variant; /* flux adjoint of fases_adjoint: 0 tiled (each WENO pair VJP
                               once), 1 per cell (neighbours' pairs re-evaluated), 2 tiled
                               in the XCD-aware tile order (same sums as 0)              */
    int    tile_rows;       /* tiled kernel: waves (grid rows) per block, 8 or 16; 0 = autotune
                               together with tile_z                                        */
    int    source_patch;    /* FP32/FP16 storage: Chebyshev radius of the FP64 region around
                               each point source (forward and adjoint); 0 = 12. Drive
                               gradients of large rooms need ~32 (FP32 drive gradient of
                               the 192x160x128 room: 14 % off FP64 at 12, 8e-4 at 32)     */
    int    tile_cols;       /* tiled kernel: x columns per block, 64 or 62 (same results);
                               0 = autotune together with tile_z and tile_rows              */
} fases_params;

/* Fills *p with the reference defaults (FP64, device 0, PML depth 20, ...). */
int fases_params_default(fases_params *p);

/* ------------------------------------------------------------------------- */
/* Formats (host C++)                                                        */
/* ------------------------------------------------------------------------- */

/* config.txt: first float of every line after '#'-comments are cut (main3D.cu:114-128).
 * Writes up to maxn values into vals, the count into *n. */
int fases_read_config(const char *path, float *vals, int maxn, int *n);

/* poles.txt "K_m L_m", each floored to 1 (main3D.cu:592-603). */
int fases_read_poles(const char *path, int *K, int *L);

/* One .fit file into material slot m of arrays laid out [m*K + k] / [m*L + l]
 * (filterABC.h read_fit_file:31-92). K, L are the poles.txt maxima. */
int fases_read_fit(const char *path, int m, int K, int L,
                   double *Rinf, double *A, double *zeta,
                   double *B, double *C, double *alpha, double *beta);

/* src.txt rows "x y z Lp,1m [f0]" -> cell = (int)(x/h)+3 per axis, sources outside
 * the grid are dropped (main3D.cu:861-906). Arrays hold maxn entries. */
int fases_read_sources(const char *path, float h, int Nx, int Ny, int Nz, float default_f0,
                       int maxn, int64_t *cell, float *lp, float *f0, float *xyz, int *n);

/* rec.txt rows "x y z" -> cell; a receiver outside the grid is FASES_E_RANGE
 * (main3D.cu:919-953). */
int fases_read_receivers(const char *path, float h, int Nx, int Ny, int Nz,
                         int maxn, int64_t *cell, float *xyz, int *n);

/* Shared source waveform u_s[0..2*Nt] sampled at dt/2 (main3D.cu:963-1058):
 * FASES_SRC_PCS (f = config line 17, f_cutoff for the pulse width) or
 * FASES_SRC_RICKER (f = peak frequency). dt is the float time step. */
int fases_source_waveform(int src_opt, float dt, double f, float f_cutoff, int64_t Nt, double *u_s);

/* Analytic 1 m pressure per unit drive p_ref[0..2*Nt] and its peak (main3D.cu:1079-1094). */
int fases_source_pref(const double *u_s, int64_t Nt, float dt, float h, float rho,
                      double *p_ref, double *p_peak);

/* Little-endian float32 receiver file (IR<r>_<s>.bin / VEL, main3D.cu:1397-1414):
 * n values of x rounded to float32. */
int fases_write_f32(const char *path, const double *x, int64_t n);

/* ------------------------------------------------------------------------- */
/* Scene: OBJ -> padded grid, signed distance S, material id oid (host C++)  */
/* ------------------------------------------------------------------------- */

/* OBJ file -> mesh shifted by +3h, narrow-band SDF and dense OID
 * (parsing3D.h parse_obj3d:690, build_sdf:860 CPU path, build_oid_dense_cpu:955). */
int fases_scene_load_obj(const char *path, float h, fases_scene **out);

/* Axis-aligned shoebox of nx*ny*nz room cells of size h: the same mesh an OBJ with
 * walls west/east/south/north/floor/ceiling (x-,x+,y-,y+,z-,z+) would give. */
int fases_scene_shoebox(int nx, int ny, int nz, float h, fases_scene **out);

/* The same two scene builders with the signed distance and material id computed on
 * HIP device `device` (device < 0: on the host, as above). The result is bit-identical
 * to the host classification. Replaces the reference CUDA build's GPU classification:
 * build_sdf_gpu_narrowband (parsing3D.h:644-661, dispatched from build_sdf:896-908) and
 * build_oid_dense_gpu (parsing3D.h:1078-1139, dispatched from build_oid_dense:1145-1154).
 * A +X ray crossing more than 4096 triangles is FASES_E_ARG. */
int fases_scene_load_obj_dev(const char *path, float h, int device, fases_scene **out);
int fases_scene_shoebox_dev(int nx, int ny, int nz, float h, int device, fases_scene **out);

/* The OBJ scene with a classification the caller already holds (S float32 and oid, Nx*Ny*Nz
 * cells, x fastest): parsed and placed as fases_scene_load_obj, the grid checked against
 * Nx, Ny, Nz (FASES_E_ARG if it differs), S and oid copied in instead of computed. Replaces
 * the scene-cache reuse of 3D-solver/core/main3D.cu:334-379 (S.bin / oid.bin of a run whose
 * scene_cache.txt key matches). */
int fases_scene_load_obj_classified(const char *path, float h, const float *S, const uint8_t *oid, int Nx, int Ny,
                                    int Nz, fases_scene **out);

/* Padded grid dims, number of mesh groups (materials), padded extent [m]. */
int fases_scene_info(const fases_scene *s, int *Nx, int *Ny, int *Nz, int *n_groups, float *L3);

/* Name of mesh group g (0-based; material m = g, oid = g+1). */
int fases_scene_group_name(const fases_scene *s, int g, char *buf, int buflen);

/* Copies S [Nx*Ny*Nz float] and oid [Nx*Ny*Nz uint8] (either may be NULL). */
int fases_scene_get(const fases_scene *s, float *S, uint8_t *oid);

/* Surface area per mesh group (A[n_groups], may be NULL) and in total [m^2] of the placed
 * mesh (parsing3D.h mesh_group_areas:101-114), for the reverberation estimate. */
int fases_scene_group_areas(const fases_scene *s, float *A, float *A_tot);

void fases_scene_free(fases_scene *s);

/* ------------------------------------------------------------------------- */
/* Reverberation estimate of the preflight (main3D.cu:452-497)                */
/* ------------------------------------------------------------------------- */

/* Air attenuation [Np/m] of ISO 9613-1 at n frequencies for temperature T [K], altitude
 * alt [m] (barometric pressure) and molar water-vapour concentration hm [%]
 * (functions3D.cuh frO/frN/alpha_air/p_atm:15-46; the reference uses hm = 1.1). */
int fases_air_attenuation(float T, float alt, float hm, const float *freq, int n, float *m);

/* <layer>.mat: the first 8 octave-band absorption coefficients (63 Hz .. 8 kHz); values
 * not present are 0 (ism3D.h ism_load_impedances:136-150). A missing file is FASES_E_IO
 * with alpha all 0 (the reference reads it as all 0). */
int fases_read_mat(const char *path, float *alpha, int *n_read);

/* Eyring reverberation time per octave band with the air term (ism3D.h
 * estimate_T60s:735-751): n groups with areas A[n] and absorption alpha[n][8], air
 * attenuation m_air[8]; a band without absorption gets 1e9. f_schroeder (nullable) =
 * 2000 sqrt(T60[500 Hz] / V), or 0 where that band does not decay (main3D.cu:482). */
int fases_estimate_t60(float V, float c, float A_tot, int n, const float *A, const float *alpha, const float *m_air,
                       float *T60, float *f_schroeder);

/* Compact ADE index: bmap[c] = running count where oid not in {0, 0xFF}, else -1
 * (filterABC.h build_boundary_map:95-108). */
int fases_boundary_map(const uint8_t *oid, int64_t N, int32_t *bmap, int64_t *nB);

/* ------------------------------------------------------------------------- */
/* Solver context (HIP)                                                      */
/* ------------------------------------------------------------------------- */

int fases_create(const fases_params *p, fases_ctx **out);
int fases_destroy(fases_ctx *ctx);

/* Message of the last failing call on ctx (ctx may be NULL: last fases_create / scene /
 * format error of the calling thread). */
const char *fases_last_error(const fases_ctx *ctx);

/* Padded-grid geometry: S (signed distance, >0 in air) and oid (1-based material id,
 * 0xFF outside the band). Builds the update mask (interior, S >= h), the boundary-cell
 * list (bmap order) and the PML tables; uploads everything. */
int fases_set_geometry(fases_ctx *ctx, const float *S, const uint8_t *oid, int Nx, int Ny, int Nz);

/* ADE materials, arrays [m*K + k] and [m*L + l] (main3D.cu:575-668). Call after
 * fases_set_geometry; allocates the FP64 pole states. */
int fases_set_materials(fases_ctx *ctx, int nMat, int K, int L,
                        const double *Rinf, const double *A, const double *zeta,
                        const double *B, const double *C, const double *alpha, const double *beta);

/* Point sources (kernel_inject_src_SSP, functions3D.cuh:1039): cell indices, scaled
 * amplitude sval, option PCS/Ricker (shared u_s[n_samp], n_samp >= 2*nsteps+1 of every
 * later run) or SINE (per-source freq [Hz]). */
int fases_set_sources(fases_ctx *ctx, int n, const int64_t *cell, const double *sval,
                      int src_opt, const double *u_s, int64_t n_samp, const double *freq);

/* Receiver cells; each run records p (and vx,vy,vz) after every step. */
int fases_set_receivers(fases_ctx *ctx, int n, const int64_t *cell);

/* Sets the state U^n on every cell of all three field sets (NULL field = zero) and
 * clears ADE/PML states and the step counter. */
int fases_set_fields(fases_ctx *ctx, const double *p, const double *u, const double *v, const double *w);

/* Copies the current state U^n (padded grid, FP64 on the host). Any pointer may be NULL. */
int fases_get_fields(const fases_ctx *ctx, double *p, double *u, double *v, double *w);

/* Animation planes of the current fields (main3D.cu:1322-1334, copy_planes): out holds
 * 4 * (Ny*Nz + Nx*Nz + Nx*Ny) floats, fields p, u, v, w in turn, each as
 * [YZ plane at x = i0 (index j + Ny k) | XZ at y = j0 (i + Nx k) | XY at z = k0 (i + Nx j)].
 * Gathered on the device, one copy to the host. A plane outside the grid is FASES_E_RANGE. */
int fases_get_slices(const fases_ctx *ctx, int i0, int j0, int k0, float *out);

/* fases_get_slices without any device allocation: the planes are gathered into a static
 * device staging array of 2^21 floats (8 MiB) that each GPU the library runs on holds until
 * the process ends; grids whose 4 * (Ny*Nz + Nx*Nz + Nx*Ny) exceeds that are FASES_E_ARG.
 * No counterpart in the reference: for callers that cannot allocate device memory (code
 * under hipGraph capture, serving loops, harnesses that forbid hipMalloc). Calls on one GPU
 * share the array: calls that overlap in time on the same GPU race. */
int fases_get_slices_noalloc(const fases_ctx *ctx, int i0, int j0, int k0, float *out);

/* Zero fields, ADE and PML states, step counter 0. */
int fases_reset(fases_ctx *ctx);

/* Advances nsteps SSP-RK3 steps from the current step counter. ir [n_rec*nsteps]
 * (row r = receiver r) receives p after each step, vel [n_rec*nsteps*3] (nullable)
 * vx,vy,vz. *elapsed_ms (nullable) = HIP-event time of the step loop alone.
 * *kernel_ms (nullable) = HIP-event time of the flux-stage launches, scaled to all nsteps
 * from the steps it is measured on: when kernel_ms is non-NULL, events are recorded around
 * the flux stages of every 16th step (from the first), and only there. */
int fases_run(fases_ctx *ctx, int64_t nsteps, double *ir, double *vel,
              double *elapsed_ms, double *kernel_ms);

/* Forward run of nsteps steps from the current state that also stores a checkpoint
 * (fields, ADE states, cPML auxiliary states, FP64 patch shadows) every
 * checkpoint_every steps for fases_adjoint. ir as in fases_run. Any storage precision;
 * with FP16 the fields are rescaled at the steps fases_run rescales them and every
 * step's scale exponents are kept for the reverse sweep. Checkpoint and segment
 * buffers stay allocated in the context and are reused by the next call of the same
 * shape.
 * No counterpart in the reference (gradients there come from fitting loops outside the
 * solver, fminconFIT.py / vectFIT.py driving repeated runs). */
int fases_adjoint_forward(fases_ctx *ctx, int64_t nsteps, int checkpoint_every, double *ir);

/* Reverse sweep after fases_adjoint_forward: given dJ/d ir [n_rec * nsteps] of a scalar
 * misfit J of the receiver pressures, writes (each nullable)
 *   grad_mat  [nMat * (1 + 2K + 4L)]: per material dJ/d(R_inf, A[K], zeta[K], B[L], C[L],
 *             alpha[L], beta[L]) (poles a material does not use get 0),
 *   grad_sval [n_src]: dJ/d source drive amplitude,
 *   grad_us   [n_samp]: dJ/d shared waveform sample (Ricker / PCS sources only).
 * The context's field state is undefined afterwards: reset or set fields before running. */
int fases_adjoint(fases_ctx *ctx, const double *dJ_dir, double *grad_mat, double *grad_sval, double *grad_us);

/* Advances n independent contexts (jobs) together: nsteps[j] steps for job j, each
 * stage of all jobs in one launch per kernel kind, the step loop captured into
 * hipGraphs and replayed; jobs drop out as their step counts run out. ir[j]
 * (nullable) receives job j's receiver pressures [n_rec_j * nsteps[j]], exactly as
 * fases_run would give them. Jobs must share the device and use FP32 storage.
 * *elapsed_ms (nullable) = HIP-event time of the whole batch.
 * Replaces running the reference executable once per job (one process per source,
 * common/gui.py:11137-11168, each stepping main3D.cu:1280-1320 alone). */
int fases_batch_run(fases_ctx **ctx, int n, const int64_t *nsteps, double **ir, double *elapsed_ms);

/* cPML tables of one axis (0 x, 1 y, 2 z) as the kernels use them, each array of the
 * padded axis length (Nx, Ny or Nz; any pointer may be NULL): band = tile index in
 * [0, 2d) or -1 outside an enabled band, kc = 1/kappa - 1, b = exp(-(alpha + sigma/kappa)
 * dt), a = sigma/(kappa (sigma + kappa alpha)) (b - 1) (functions3D.cuh cPML :681-756,
 * cpml_adv :669-673: the reference evaluates these per cell and stage). */
int fases_pml_tables(const fases_ctx *ctx, int axis, int32_t *band, double *kc, double *b, double *a);

/* Tile of the tiled flux kernel in use: output columns x rows x z planes per block
   (*tile_z = 0 before the first fases_run when autotuning). No counterpart in the
   reference (one thread per cell there); reported so runs can state their launch shape. */
int fases_tile_config(const fases_ctx *ctx, int *tile_x, int *tile_y, int *tile_z);

/* Sizes: padded cells, boundary cells, room (updated) cells, device bytes in use. */
int fases_info(const fases_ctx *ctx, int64_t *n_cells, int64_t *n_bnd, int64_t *n_upd,
               int64_t *dev_bytes);

/* HBM bandwidth reference: best of `reps` timed launches of a streaming 16-byte copy
 * between two fresh buffers of `bytes` each on `device`; *gbs = (bytes read + bytes
 * written) / time [GB/s]. No counterpart in the reference; the roofline reports of the
 * time-stepping kernels are stated against it and against the published peak. */
int fases_hbm_copy(int device, int64_t bytes, int reps, double *gbs);

/* ------------------------------------------------------------------------- */
/* Geometrical acoustics: image sources + ray tracer (host C++ and HIP)       */
/* ------------------------------------------------------------------------- */

typedef struct fases_ga fases_ga;

typedef struct fases_ga_params {
    int   device;        /* HIP device of the path-validity and ray-tracing kernels; < 0:
                            image sources on the host, no ray tracer                    */
    int   ism_order;     /* image-source order (config line 13, main3D.cu:1484; <= 32),
                            reduced when the image tree would not fit in half the free
                            host memory (ism3D.h:539-559)                              */
    int   n_rays;        /* rays per source (config line 15, main3D.cu:1540); 0: none   */
    int   rt_max_order;  /* reflections per ray at most (main3D.cu:1541: 100000)        */
    float recv_radius;   /* detection-sphere radius [m] (main3D.cu:1543: 0.12)          */
    float c;             /* speed of sound [m/s]                                        */
    float fs;            /* sample rate of the echograms [Hz] (main3D.cu:1508: 44100)   */
    int   Nt;            /* samples (main3D.cu:1510: t_max fs)                          */
} fases_ga_params;

/* Fills *p with the reference's defaults (order 3, 10000 rays, radius 0.12 m, 44.1 kHz). */
int fases_ga_params_default(fases_ga_params *p);

/* Image sources and ray tracing of one scene for n_src sources and n_rec receivers given
 * in the scene's raw (OBJ) coordinates, placed like the mesh (ism3D.h ism_solve :665,
 * raytrace3D.h rt_build :229 + rt_trace_late :335, driven as main3D.cu:1479-1585 drives
 * them). src_pressure [n_src] (nullable: 1 Pa): pressure at 1 m; alpha [n_groups*8]
 * (nullable: rigid): octave-band absorption per mesh group (the .mat values); sctr
 * [n_groups] (nullable: none): scattering per group (the .sctr values); m_air [8]
 * (nullable: none): air attenuation [Np/m]. Paths of order <= the image-source order
 * belong to the image sources; the ray tracer keeps every higher specular order and all
 * scattered energy. */
int fases_ga_run(const fases_scene *scene, const fases_ga_params *p, int n_src, const float *src_xyz,
                 const float *src_pressure, int n_rec, const float *rec_xyz, const float *alpha,
                 const float *sctr, const float *m_air, fases_ga **out);
void fases_ga_free(fases_ga *g);

/* Image-source order used, image-tree nodes, valid image paths, and the host times of the
 * image sources, of the ray tracer and of its kernel [ms] (any pointer may be NULL). */
int fases_ga_info(const fases_ga *g, int *order_used, int64_t *n_nodes, int64_t *n_paths, double *ism_ms,
                  double *rt_ms, double *rt_kernel_ms);

/* Direct source-receiver distance [m], 0 when occluded (ISMResult::direct). */
int fases_ga_direct(const fases_ga *g, int s, int r, float *dist);

/* Image paths of (s, r), by increasing length (ISMResult::images): count and total
 * bounces (call with NULL arrays first), then per path the unfolded distance, the unit
 * arrival direction (receiver to image, placed frame) [3], the curvature gain and the
 * order; per bounce in source-to-receiver order the hit group, incidence cosine and angle. */
int fases_ga_images(const fases_ga *g, int s, int r, int *n_images, int *n_bounces, float *distance, float *doa,
                    float *amp, int *order, int *group, float *cos_inc, float *angle);

/* Ray-traced energy echograms e_spec, e_diff [8 * n_rec * Nt], index (b n_rec + r) Nt + t. */
int fases_ga_energy(const fases_ga *g, float *e_spec, float *e_diff);

/* Octave band `band` (0 = 63 Hz .. 7 = 8 kHz) of the geometrical IR of every receiver,
 * p_hf [n_rec * Nt] (the HIR file: image sources as windowed-sinc fractional delays plus
 * the ray-traced energy as pressure with a random sign) and p_rt [n_rec * Nt] (nullable;
 * the HRT file: the ray-traced part alone) (main3D.cu:1600-1757). */
int fases_ga_band_irs(const fases_ga *g, int band, float *p_hf, float *p_rt);

/* Discrete arrivals at receiver r (the ARR file, main3D.cu:1759-1801): *n arrivals (buf
 * NULL to count), each 12 floats: delay [s], unit direction receiver -> (image) source
 * [3], pressure gain per octave band [8]. */
int fases_ga_arrivals(const fases_ga *g, int r, int *n, float *buf);

/* Debug trace for the GUI overlay (--dump-rays, raytrace3D.h rt_dump_paths :581-640):
 * n_rays Fibonacci rays per source (raw coordinates; none given: the origin), purely
 * specular, up to max_order segments each, written to `path` as "src ray x0 y0 z0 x1 y1
 * z1 order group" in raw (OBJ) coordinates after one "# group <id> <name>" line per mesh
 * group; an escaped ray ends in a stub of a quarter of the scene diagonal (order, group -1). */
int fases_ga_dump_rays(const fases_scene *scene, int n_src, const float *src_xyz, int n_rays, int max_order,
                       const char *path);

/* Number of HIP devices visible to this process (0 without a usable runtime or device;
   the call itself succeeds). Lets a caller pick the GPU or the host classification
   before it has a context; the reference's CUDA build makes that choice at compile time
   (parsing3D.h:1145-1154). */
int fases_device_count(int *n);

#ifdef __cplusplus
}
#endif
#endif /* FASES_CORE_H */
