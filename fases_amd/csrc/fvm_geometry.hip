// Scene classification on the GPU: narrow-band |distance|, +X parity sign, dense
// material id (oid). The same results as the host classification (host_geometry.cpp),
// bit for bit: every float expression below is the host's, in the host's operand
// order, with no FMA contraction (this file is compiled with -ffp-contract=off) and
// IEEE division and square root (hipcc's default for float).
//
// Reference: the CUDA build classifies on the GPU as well (parsing3D.h:568-655
// k_sdf_scatter / k_sdf_sign_finalize / build_sdf_gpu_narrowband, dispatched from
// build_sdf :860-910; parsing3D.h:1025-1139 k_build_oid_dense / build_oid_dense_gpu,
// dispatched from build_oid_dense :1145-1154). Layout here:
//   band:  cell-major. A block owns an 8x8x4 cell tile, culls the triangles whose
//          band box (integer cell range, computed on the host) meets the tile into an
//          LDS list, one chunk of triangles at a time, and each thread takes the min
//          over the list for its cell. Min is order-free, so the result does not
//          depend on the list order.
//   sign:  one wave per (j, k) row. The lanes intersect the row's +X ray with the
//          triangles into an LDS list of crossing abscissae, then every cell counts the
//          crossings ahead of it (> x_c + 1e-6): a count, so no sort and no cap below
//          the list capacity (a longer list is an error, not a fallback).
//   oid:   the cells with 0 <= S <= 3h go to a compact list (one atomic per block);
//          one thread per listed cell searches the nearest triangle over all triangles,
//          streamed through LDS in index order with a strict < (the first of equal
//          distances wins, as on the host).
#pragma clang fp contract(off)

#include "host.hpp"
#include "../../include/fases_core.h"

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct G3 {
    float x, y, z;
};
__device__ inline G3 gmk(float x, float y, float z) { G3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ inline G3 gsub(const G3 &a, const G3 &b) { return gmk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline G3 gadd(const G3 &a, const G3 &b) { return gmk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline G3 gscale(const G3 &a, float s) { return gmk(a.x * s, a.y * s, a.z * s); }
__device__ inline float gdot(const G3 &a, const G3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline G3 gcross(const G3 &a, const G3 &b) {
    return gmk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Placed triangle corners (a, b, c), 9 floats.
struct GTri {
    G3 a, b, c;
};
// Band box of a triangle in cell indices, inclusive, clamped to the grid.
struct GBox {
    int i0, i1, j0, j1, k0, k1;
};

// Squared distance from p to triangle (a, b, c): the host's Voronoi-region walk
// (Ericson, Real-Time Collision Detection, sec. 5.1.5), same expressions.
__device__ float g_tri_dist2(const G3 &p, const G3 &a, const G3 &b, const G3 &c) {
    const G3 ab = gsub(b, a), ac = gsub(c, a), ap = gsub(p, a);
    const float d1 = gdot(ab, ap), d2 = gdot(ac, ap);
    if (d1 <= 0.f && d2 <= 0.f) return gdot(ap, ap);
    const G3 bp = gsub(p, b);
    const float d3 = gdot(ab, bp), d4 = gdot(ac, bp);
    if (d3 >= 0.f && d4 <= d3) return gdot(bp, bp);
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {
        const float t = d1 / (d1 - d3);
        const G3 q = gsub(p, gadd(a, gscale(ab, t)));
        return gdot(q, q);
    }
    const G3 cp = gsub(p, c);
    const float d5 = gdot(ab, cp), d6 = gdot(ac, cp);
    if (d6 >= 0.f && d5 <= d6) return gdot(cp, cp);
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) {
        const float t = d2 / (d2 - d6);
        const G3 q = gsub(p, gadd(a, gscale(ac, t)));
        return gdot(q, q);
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
        const G3 bc = gsub(c, b);
        const float t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        const G3 q = gsub(p, gadd(b, gscale(bc, t)));
        return gdot(q, q);
    }
    const float inv = 1.f / (va + vb + vc);
    const float v = vb * inv, w = vc * inv;
    const G3 q = gsub(p, gadd(a, gadd(gscale(ab, v), gscale(ac, w))));
    return gdot(q, q);
}

// +X ray (direction (1, 0, 0)) against a triangle: the host's Moller-Trumbore test,
// edge-inclusive, t > tmin accepted.
__device__ bool g_ray_hits(const G3 &o, const G3 &a, const G3 &b, const G3 &c, float tmin, float &t_out) {
    const float eps = 1e-8f;
    const G3 d = gmk(1.f, 0.f, 0.f);
    const G3 e1 = gsub(b, a), e2 = gsub(c, a);
    const G3 pv = gcross(d, e2);
    const float det = gdot(e1, pv);
    if (det > -eps && det < eps) return false;
    const float inv = 1.f / det;
    const G3 tv = gsub(o, a);
    const float u = gdot(tv, pv) * inv;
    if (u < 0.f || u > 1.f) return false;
    const G3 qv = gcross(tv, e1);
    const float v = gdot(d, qv) * inv;
    if (v < 0.f || u + v > 1.f) return false;
    const float t = gdot(e2, qv) * inv;
    if (t > tmin) { t_out = t; return true; }
    return false;
}

__device__ inline float centre(int i, float h) { return 0.f + ((float)i + 0.5f) * h; }

// ---- band |distance|: tile 8 x 8 x 4 cells, 256 threads
constexpr int kTX = 8, kTY = 8, kTZ = 4, kBandThreads = kTX * kTY * kTZ;
constexpr int kChunk = 1024;    // triangles culled per pass (the list holds a whole chunk)

__global__ void __launch_bounds__(kBandThreads)
k_geo_band(const GTri *__restrict__ tri, const GBox *__restrict__ box, int nt, int Nx, int Ny, int Nz, float h,
           float band, float *__restrict__ absd) {
    __shared__ int list[kChunk];
    __shared__ int cnt;
    const int nbx = (Nx + kTX - 1) / kTX, nby = (Ny + kTY - 1) / kTY;
    const int b = blockIdx.x;
    const int ti0 = (b % nbx) * kTX, tj0 = ((b / nbx) % nby) * kTY, tk0 = (b / (nbx * nby)) * kTZ;
    const int ti1 = min(ti0 + kTX, Nx) - 1, tj1 = min(tj0 + kTY, Ny) - 1, tk1 = min(tk0 + kTZ, Nz) - 1;
    const int tid = threadIdx.x;
    const int i = ti0 + tid % kTX, j = tj0 + (tid / kTX) % kTY, k = tk0 + tid / (kTX * kTY);
    const bool mine = i < Nx && j < Ny && k < Nz;
    const G3 p = gmk(centre(i, h), centre(j, h), centre(k, h));
    float best = band;
    for (int c0 = 0; c0 < nt; c0 += kChunk) {
        if (tid == 0) cnt = 0;
        __syncthreads();
        for (int t = c0 + tid; t < min(nt, c0 + kChunk); t += kBandThreads) {
            const GBox bx = box[t];
            if (bx.i0 <= ti1 && bx.i1 >= ti0 && bx.j0 <= tj1 && bx.j1 >= tj0 && bx.k0 <= tk1 && bx.k1 >= tk0)
                list[atomicAdd(&cnt, 1)] = t;
        }
        __syncthreads();
        const int n = cnt;
        if (mine) {
            for (int q = 0; q < n; ++q) {
                const int t = list[q];
                const GBox bx = box[t];
                if (i < bx.i0 || i > bx.i1 || j < bx.j0 || j > bx.j1 || k < bx.k0 || k > bx.k1) continue;
                const GTri T = tri[t];
                const float d = sqrtf(g_tri_dist2(p, T.a, T.b, T.c));
                if (d < best) best = d;
            }
        }
        __syncthreads();            // the list is rewritten by the next chunk
    }
    if (mine) absd[(size_t)k * Nx * Ny + (size_t)j * Nx + i] = best;
}

// ---- sign: one 64-lane block per (j, k) row
constexpr int kXCap = 4096;      // crossings per +X ray held in LDS

__global__ void __launch_bounds__(64)
k_geo_sign(const GTri *__restrict__ tri, int nt, int Nx, int Ny, float h, const float *__restrict__ absd,
           float *__restrict__ S, int *__restrict__ overflow) {
    __shared__ float xs[kXCap];
    __shared__ int cnt;
    const int row = blockIdx.x;           // j + Ny k
    const int j = row % Ny, k = row / Ny;
    const int lane = threadIdx.x;
    if (lane == 0) cnt = 0;
    __syncthreads();
    const float y0 = 0.f + ((float)j + 0.5f) * h + 0.0137f * h;
    const float z0 = 0.f + ((float)k + 0.5f) * h + 0.0091f * h;
    const G3 o = gmk(0.f, y0, z0);
    for (int t = lane; t < nt; t += 64) {
        float tt;
        if (g_ray_hits(o, tri[t].a, tri[t].b, tri[t].c, 0.0f, tt)) {
            const int q = atomicAdd(&cnt, 1);
            if (q < kXCap) xs[q] = o.x + tt;
        }
    }
    __syncthreads();
    const int n = cnt;
    if (n > kXCap) {                      // uniform: every lane sees the same count
        if (lane == 0) atomicMax(overflow, n);
        return;
    }
    const float tmin = 1e-6f;
    const size_t base = (size_t)row * Nx;
    for (int i = lane; i < Nx; i += 64) {
        const float lim = centre(i, h) + tmin;
        int ahead = 0;
        for (int q = 0; q < n; ++q) ahead += xs[q] > lim;
        const float s = absd[base + i];
        S[base + i] = (ahead & 1) ? s : -s;
    }
}

// Foot point of band cell (i, j, k): the host's gradient, normal and W = C - s n.
// Returns false for a cell that is not a band cell (or has a vanishing gradient).
__device__ inline bool band_foot(const float *__restrict__ S, int i, int j, int k, int Nx, int Ny, float h, float tol,
                                 G3 &W) {
    const size_t nxy = (size_t)Nx * Ny;
    const size_t id = (size_t)k * nxy + (size_t)j * Nx + i;
    const float s = S[id];
    if (!(s >= 0.0f && fabsf(s) <= tol)) return false;
    const float two_h = 2.0f * h;
    const float gx = (S[id + 1] - S[id - 1]) / two_h;
    const float gy = (S[id + Nx] - S[id - Nx]) / two_h;
    const float gz = (S[id + nxy] - S[id - nxy]) / two_h;
    const float g2 = gx * gx + gy * gy + gz * gz;
    if (g2 <= 1e-24f) return false;
    const float ig = 1.0f / sqrtf(g2);
    const G3 n = gmk(gx * ig, gy * ig, gz * ig);
    const G3 C = gmk(centre(i, h), centre(j, h), centre(k, h));
    W = gmk(C.x - s * n.x, C.y - s * n.y, C.z - s * n.z);
    return true;
}

// ---- oid, pass 1: compact list of band cells (interior 1 <= i,j,k < N-1). A block
// takes kListPer x 256 cells (coalesced rounds), counts its band cells with a block
// scan and reserves their slots with one atomic.
constexpr int kListThreads = 256, kListPer = 16;

__global__ void __launch_bounds__(kListThreads)
k_geo_band_list(const float *__restrict__ S, int Nx, int Ny, int Nz, float h, float tol, int *__restrict__ list,
                int *__restrict__ count) {
    __shared__ int wsum[kListThreads / 64];
    __shared__ int bbase;
    const long long nxy = (long long)Nx * Ny, N = nxy * Nz;
    const long long base = (long long)blockIdx.x * kListThreads * kListPer;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    unsigned mask = 0;
    for (int q = 0; q < kListPer; ++q) {
        const long long c = base + (long long)q * kListThreads + tid;
        if (c >= N) break;
        const int i = (int)(c % Nx), j = (int)((c / Nx) % Ny), k = (int)(c / nxy);
        G3 W;
        if (i >= 1 && j >= 1 && k >= 1 && i < Nx - 1 && j < Ny - 1 && k < Nz - 1 &&
            band_foot(S, i, j, k, Nx, Ny, h, tol, W))
            mask |= 1u << q;
    }
    const int mine = __popc(mask);
    int x = mine;                                   // inclusive scan over the wave
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    if (tid == 0) {
        int tot = 0;
        for (int w = 0; w < kListThreads / 64; ++w) tot += wsum[w];
        bbase = tot ? atomicAdd(count, tot) : 0;
    }
    __syncthreads();
    int off = bbase + x - mine;
    for (int w = 0; w < wave; ++w) off += wsum[w];
    for (int q = 0; q < kListPer; ++q)
        if (mask >> q & 1u) list[off++] = (int)(base + (long long)q * kListThreads + tid);
}

// ---- oid, pass 2: nearest triangle of each listed cell
constexpr int kOidThreads = 256, kOidChunk = 1024;

__global__ void __launch_bounds__(kOidThreads)
k_geo_oid(const float *__restrict__ S, const GTri *__restrict__ tri, const int *__restrict__ gid, int nt,
          const int *__restrict__ lid_of, int max_gid, int Nx, int Ny, float h, float tol,
          const int *__restrict__ list, const int *__restrict__ count, uint8_t *__restrict__ oid) {
    __shared__ GTri T[kOidChunk];
    __shared__ int G[kOidChunk];
    const int n = *count;
    const int q = blockIdx.x * kOidThreads + threadIdx.x;
    if (blockIdx.x * kOidThreads >= n) return;     // uniform per block
    const bool mine = q < n;
    int c = 0;
    G3 W = gmk(0.f, 0.f, 0.f);
    if (mine) {
        c = list[q];
        const long long nxy = (long long)Nx * Ny;
        const int i = (int)(c % Nx), j = (int)((c / Nx) % Ny), k = (int)(c / nxy);
        band_foot(S, i, j, k, Nx, Ny, h, tol, W);   // true for every listed cell
    }
    float best = FLT_MAX;
    int gbest = -1;
    for (int c0 = 0; c0 < nt; c0 += kOidChunk) {
        const int m = min(kOidChunk, nt - c0);
        for (int t = threadIdx.x; t < m; t += kOidThreads) {
            T[t] = tri[c0 + t];
            G[t] = gid[c0 + t];
        }
        __syncthreads();
        if (mine) {
            for (int t = 0; t < m; ++t) {
                const float d2 = g_tri_dist2(W, T[t].a, T[t].b, T[t].c);
                if (d2 < best) { best = d2; gbest = G[t]; }
            }
        }
        __syncthreads();
    }
    if (!mine || gbest < 0 || gbest > max_gid) return;
    const int lid = lid_of[gbest];
    if (lid >= 0 && lid <= 254) oid[c] = (uint8_t)lid;
}

__global__ void k_fill_u8(uint8_t *__restrict__ a, long long n, uint8_t v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

struct DevMem {
    std::vector<void *> p;
    ~DevMem() {
        for (void *q : p) (void)hipFree(q);
    }
    template <typename T> hipError_t get(T **out, size_t n) {
        void *q = nullptr;
        const hipError_t e = hipMalloc(&q, n * sizeof(T) + 1);
        if (e == hipSuccess) p.push_back(q);
        *out = (T *)q;
        return e;
    }
};

}  // namespace

namespace fases {

// Device part of scene_build for a scene whose placement, grid size, triangle corners,
// band boxes and group -> material table the host has already set up. Writes S and oid
// (host arrays of Nx Ny Nz). Returns FASES_OK or a negative code with `err` set.
int classify_gpu(int device, int Nx, int Ny, int Nz, float h, const std::vector<F3> &ta, const std::vector<F3> &tb,
                 const std::vector<F3> &tc, const std::vector<int> &boxes, const std::vector<int> &gid,
                 const std::vector<int> &lid_of, int max_gid, float *S_out, uint8_t *oid_out, std::string &err) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) { err = "no HIP device for the GPU classification"; return FASES_E_HIP; }
    if (device < 0 || device >= ndev) { err = "device index out of range"; return FASES_E_ARG; }
#define GCHK(x)                                                                             \
    do {                                                                                    \
        const hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) { err = std::string("GPU classification: ") + hipGetErrorString(e_); return FASES_E_HIP; } \
    } while (0)
    int prev = -1;
    GCHK(hipGetDevice(&prev));
    struct DeviceGuard {              // the caller's current device is restored on return
        int d;
        ~DeviceGuard() { if (d >= 0) (void)hipSetDevice(d); }
    } dg{prev};
    GCHK(hipSetDevice(device));
    const int nt = (int)ta.size();
    const size_t N = (size_t)Nx * Ny * Nz;
    std::vector<GTri> ht((size_t)nt);
    for (int t = 0; t < nt; ++t) {
        ht[(size_t)t].a = {ta[(size_t)t].x, ta[(size_t)t].y, ta[(size_t)t].z};
        ht[(size_t)t].b = {tb[(size_t)t].x, tb[(size_t)t].y, tb[(size_t)t].z};
        ht[(size_t)t].c = {tc[(size_t)t].x, tc[(size_t)t].y, tc[(size_t)t].z};
    }
    hipStream_t st = nullptr;
    GCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{st};
    DevMem mem;
    GTri *d_tri = nullptr;
    GBox *d_box = nullptr;
    int *d_gid = nullptr, *d_lid = nullptr, *d_list = nullptr, *d_cnt = nullptr;
    float *d_absd = nullptr, *d_S = nullptr;
    uint8_t *d_oid = nullptr;
    GCHK(mem.get(&d_tri, (size_t)nt));
    GCHK(mem.get(&d_box, (size_t)nt));
    GCHK(mem.get(&d_gid, (size_t)nt));
    GCHK(mem.get(&d_lid, lid_of.size()));
    GCHK(mem.get(&d_cnt, 2));
    GCHK(mem.get(&d_absd, N));
    GCHK(mem.get(&d_S, N));
    GCHK(mem.get(&d_oid, N));
    GCHK(mem.get(&d_list, N));
    if (nt > 0) {
        GCHK(hipMemcpyAsync(d_tri, ht.data(), (size_t)nt * sizeof(GTri), hipMemcpyHostToDevice, st));
        GCHK(hipMemcpyAsync(d_box, boxes.data(), (size_t)nt * sizeof(GBox), hipMemcpyHostToDevice, st));
        GCHK(hipMemcpyAsync(d_gid, gid.data(), (size_t)nt * sizeof(int), hipMemcpyHostToDevice, st));
    }
    if (!lid_of.empty())
        GCHK(hipMemcpyAsync(d_lid, lid_of.data(), lid_of.size() * sizeof(int), hipMemcpyHostToDevice, st));
    GCHK(hipMemsetAsync(d_cnt, 0, 2 * sizeof(int), st));
    const float band = 8.f * h, tol = 3.0f * h;
    const int nbx = (Nx + kTX - 1) / kTX, nby = (Ny + kTY - 1) / kTY, nbz = (Nz + kTZ - 1) / kTZ;
    hipLaunchKernelGGL(k_geo_band, dim3((unsigned)(nbx * nby * nbz)), dim3(kBandThreads), 0, st, d_tri, d_box, nt,
                       Nx, Ny, Nz, h, band, d_absd);
    GCHK(hipGetLastError());
    hipLaunchKernelGGL(k_geo_sign, dim3((unsigned)(Ny * Nz)), dim3(64), 0, st, d_tri, nt, Nx, Ny, h, d_absd, d_S,
                       d_cnt + 1);
    GCHK(hipGetLastError());
    const unsigned nfill = (unsigned)((N + 255) / 256);
    hipLaunchKernelGGL(k_fill_u8, dim3(nfill), dim3(256), 0, st, d_oid, (long long)N, (uint8_t)0xFF);
    GCHK(hipGetLastError());
    if (max_gid >= 0 && nt > 0) {
        const unsigned nlist = (unsigned)((N + kListThreads * kListPer - 1) / (kListThreads * kListPer));
        hipLaunchKernelGGL(k_geo_band_list, dim3(nlist), dim3(kListThreads), 0, st, d_S, Nx, Ny, Nz, h, tol, d_list,
                           d_cnt);
        GCHK(hipGetLastError());
        int n_band = 0;
        GCHK(hipMemcpyAsync(&n_band, d_cnt, sizeof(int), hipMemcpyDeviceToHost, st));
        GCHK(hipStreamSynchronize(st));
        if (n_band > 0) {
            hipLaunchKernelGGL(k_geo_oid, dim3((unsigned)((n_band + kOidThreads - 1) / kOidThreads)), dim3(kOidThreads),
                               0, st, d_S, d_tri, d_gid, nt, d_lid, max_gid, Nx, Ny, h, tol, d_list, d_cnt, d_oid);
            GCHK(hipGetLastError());
        }
    }
    int ovf = 0;
    GCHK(hipMemcpyAsync(&ovf, d_cnt + 1, sizeof(int), hipMemcpyDeviceToHost, st));
    GCHK(hipMemcpyAsync(S_out, d_S, N * sizeof(float), hipMemcpyDeviceToHost, st));
    GCHK(hipMemcpyAsync(oid_out, d_oid, N, hipMemcpyDeviceToHost, st));
    GCHK(hipStreamSynchronize(st));
#undef GCHK
    if (ovf > 0) {
        err = "GPU classification: a +X ray crosses " + std::to_string(ovf) + " triangles (more than " +
              std::to_string(kXCap) + ")";
        return FASES_E_ARG;
    }
    return FASES_OK;
}

}  // namespace fases

extern "C" int fases_device_count(int *n) {
    if (!n) return FASES_E_ARG;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
    return FASES_OK;
}

