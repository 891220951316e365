// Device arithmetic of the FVM step shared by the HIP kernels: WENO5-Z face
// reconstruction, Godunov acoustic flux, trilinear probe, ADE recursion.
//
// WENO5-Z: 3D-solver/core/functions3D.cuh:763-977 (Borges, Carmona, Costa & Don 2008,
// q = 1, eps = 1e-10); Godunov flux: functions3D.cuh:56-78; trilinear: :411-422;
// ADE_solver: :426-459; chi_layered: :263-271.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace fasesdev {

// Storage <-> compute conversion. FP16 fields carry a per-field power-of-two scale:
// stored = value * inv_scale, value = stored * scale (exact for powers of two).
template <typename S, typename R> struct Conv;
template <> struct Conv<double, double> {
    __device__ static inline double ld(double x, double) { return x; }
    __device__ static inline double st(double x, double) { return x; }
};
template <> struct Conv<float, float> {
    __device__ static inline float ld(float x, float) { return x; }
    __device__ static inline float st(float x, float) { return x; }
};
template <> struct Conv<float, double> {
    __device__ static inline double ld(float x, double) { return (double)x; }
    __device__ static inline float st(double x, double) { return (float)x; }
};
template <> struct Conv<__half, double> {
    __device__ static inline double ld(__half x, double scale) { return (double)__half2float(x) * scale; }
    __device__ static inline __half st(double x, double inv_scale) { return __float2half_rn((float)(x * inv_scale)); }
};
template <> struct Conv<__half, float> {
    __device__ static inline float ld(__half x, float scale) { return __half2float(x) * scale; }
    __device__ static inline __half st(float x, float inv_scale) { return __float2half_rn(x * inv_scale); }
};

// Storage scales of the four fields (p, vx, vy, vz), shared by the three field sets:
// value = stored * s[f], stored = value * is[f]. Powers of two; 1 for FP32 / FP64.
template <typename R> struct Scl {
    R s[4], is[4];
};

// Literal WENO5-Z of the reference for the two states of face i+1/2, from the six
// cells u[i-2..i+3] = (a,b,c,d,e,f):
//   left  = V^-_{i+1/2}, reconstruction of cell i   (idx = 1 in functions3D.cuh)
//   right = V^+_{i+1/2}, reconstruction of cell i+1 (idx = 0)
// Candidate rows c[rr] = {11/6,-7/6,1/3}, {1/3,5/6,-1/6}, {-1/6,5/6,1/3}, {1/3,-7/6,11/6}
// (main3D.cu:834-837), beta rows {3,-4,1}, {1,0,-1}, {1,-4,3} (:839-841), linear
// weights d = {3/10, 3/5, 1/10} (:843) mirrored for the right state.
// The literal-reference functions below evaluate without FMA contraction so that the
// FP64 per-cell kernel rounds exactly like the FP64 CPU restatement of the reference.
template <typename R>
__device__ inline R weno_smooth(R x, R y, R z, R k0, R k1, R k2) {
#pragma clang fp contract(off)
    const R s = x - R(2) * y + z;
    const R t = k0 * x + k1 * y + k2 * z;
    return (R(13) * s * s / R(12)) + (R(0.25) * t * t);
}

template <typename R>
__device__ inline R weno_combine(R v0, R v1, R v2, R b0, R b1, R b2, R d0, R d1, R d2) {
#pragma clang fp contract(off)
    const R eps = R(1e-10);
    const R tau = fabs(b0 - b2);
    const R a0 = d0 * (R(1) + tau / (b0 + eps));
    const R a1 = d1 * (R(1) + tau / (b1 + eps));
    const R a2 = d2 * (R(1) + tau / (b2 + eps));
    const R sum = a0 + a1 + a2;
    R v = R(0);
    v += (a0 / sum) * v0;
    v += (a1 / sum) * v1;
    v += (a2 / sum) * v2;
    return v;
}

template <typename R>
__device__ inline void weno_face(R a, R b, R c, R d, R e, R f, R &left, R &right) {
#pragma clang fp contract(off)
    const R c3 = R(1) / R(3), c56 = R(5) / R(6), c16 = R(1) / R(6), c116 = R(11) / R(6), c76 = R(7) / R(6);
    // cell i (centre c): candidates r = 0,1,2 on (c,d,e), (b,c,d), (a,b,c)
    {
        R v0 = R(0); v0 += c3 * c; v0 += c56 * d; v0 += -c16 * e;
        R v1 = R(0); v1 += -c16 * b; v1 += c56 * c; v1 += c3 * d;
        R v2 = R(0); v2 += c3 * a; v2 += -c76 * b; v2 += c116 * c;
        const R b0 = weno_smooth(c, d, e, R(3), R(-4), R(1));
        const R b1 = weno_smooth(b, c, d, R(1), R(0), R(-1));
        const R b2 = weno_smooth(a, b, c, R(1), R(-4), R(3));
        left = weno_combine(v0, v1, v2, b0, b1, b2, R(0.3), R(0.6), R(0.1));
    }
    // cell i+1 (centre d): candidates r = 0,1,2 on (d,e,f), (c,d,e), (b,c,d)
    {
        R v0 = R(0); v0 += c116 * d; v0 += -c76 * e; v0 += c3 * f;
        R v1 = R(0); v1 += c3 * c; v1 += c56 * d; v1 += -c16 * e;
        R v2 = R(0); v2 += -c16 * b; v2 += c56 * c; v2 += c3 * d;
        const R b0 = weno_smooth(d, e, f, R(3), R(-4), R(1));
        const R b1 = weno_smooth(c, d, e, R(1), R(0), R(-1));
        const R b2 = weno_smooth(b, c, d, R(1), R(-4), R(3));
        right = weno_combine(v0, v1, v2, b0, b1, b2, R(0.1), R(0.6), R(0.3));
    }
}

// Godunov (exact upwind Riemann) flux of the linear acoustic system across a face
// with left state (pL,uL) and right state (pR,uR); u = the face-normal velocity.
template <typename R>
__device__ inline void godunov(R pL, R pR, R uL, R uR, R rho, R c, R &Fp, R &Fu) {
#pragma clang fp contract(off)
    Fp = R(0.5) * (rho * c * c * (uR + uL) - c * (pR - pL));
    Fu = R(0.5) * ((pR + pL) / rho - c * (uR - uL));
}

// Immersed-boundary blend weight: 1 for s <= h, smoothstep down to 0 at s = tol.
template <typename R>
__host__ __device__ inline R chi_band(R s, R h, R tol) {
#pragma clang fp contract(off)
    if (s >= tol) return R(0);
    if (s <= h) return R(1);
    const R r = (s - h) / (tol - h);
    return R(1) - r * r * (R(3) - R(2) * r);
}

// cPML of one updated cell (functions3D.cuh:681-756): the frozen psi stretches the
// stage divergence at the stage's RK weight (wdt = weight * dt); the last stage
// (ADVANCE) then advances psi once per step with its own divergence. Per-axis 1-D
// tables hold band index, 1/kappa - 1, b = exp(-(alpha + sigma/kappa) dt) and
// a = sigma/(kappa(sigma + kappa alpha)) (b - 1).
struct PmlAxis {
    const int *band;       // -1 outside the band, else tile index in [0, 2d)
    const double *kc;
    const double *b;
    const double *a;
};

struct PmlSet {
    int on;
    int d;
    // per axis (x, y, z): the indices in [clear0, clear1) lie in no enabled band
    int clear0[3], clear1[3];
    PmlAxis x, y, z;
    double *psi_px, *psi_py, *psi_pz, *psi_ux, *psi_vy, *psi_wz;
};

template <typename R, bool ADVANCE>
__device__ inline void pml_apply(const PmlSet &pml, int i, int j, int k, int Nx, int Ny, R &po, R &uo, R &vo, R &wo,
                                 R Lpx, R Lpy, R Lpz, R Lux, R Lvy, R Lwz, R wdt) {
#pragma clang fp contract(off)
    const int bx = pml.x.band[i], by = pml.y.band[j], bz = pml.z.band[k];
    if (bx >= 0) {
        const long long X = bx + (long long)(2 * pml.d) * (j + (long long)Ny * k);
        const R kc = (R)pml.x.kc[i];
        po = po + wdt * (kc * Lpx + (R)pml.psi_px[X]);
        uo = uo + wdt * (kc * Lux + (R)pml.psi_ux[X]);
        if (ADVANCE) {
            const R b = (R)pml.x.b[i], a = (R)pml.x.a[i];
            pml.psi_px[X] = (double)(b * (R)pml.psi_px[X] + a * Lpx);
            pml.psi_ux[X] = (double)(b * (R)pml.psi_ux[X] + a * Lux);
        }
    }
    if (by >= 0) {
        const long long Y = i + (long long)Nx * (by + (long long)(2 * pml.d) * k);
        const R kc = (R)pml.y.kc[j];
        po = po + wdt * (kc * Lpy + (R)pml.psi_py[Y]);
        vo = vo + wdt * (kc * Lvy + (R)pml.psi_vy[Y]);
        if (ADVANCE) {
            const R b = (R)pml.y.b[j], a = (R)pml.y.a[j];
            pml.psi_py[Y] = (double)(b * (R)pml.psi_py[Y] + a * Lpy);
            pml.psi_vy[Y] = (double)(b * (R)pml.psi_vy[Y] + a * Lvy);
        }
    }
    if (bz >= 0) {
        const long long Z = i + (long long)Nx * (j + (long long)Ny * bz);
        const R kc = (R)pml.z.kc[k];
        po = po + wdt * (kc * Lpz + (R)pml.psi_pz[Z]);
        wo = wo + wdt * (kc * Lwz + (R)pml.psi_wz[Z]);
        if (ADVANCE) {
            const R b = (R)pml.z.b[k], a = (R)pml.z.a[k];
            pml.psi_pz[Z] = (double)(b * (R)pml.psi_pz[Z] + a * Lpz);
            pml.psi_wz[Z] = (double)(b * (R)pml.psi_wz[Z] + a * Lwz);
        }
    }
}

}  // namespace fasesdev
