// ORACLE -- test infrastructure only. FP64 CPU restatement of the FASES 3D-solver
// time step, used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// as the checker of the HIP path. Never linked into, or called by, the product.
//
// Follows, operation by operation (all arithmetic in double):
//   3D-solver/core/main3D.cu:1280-1320    time loop, stage order, buffer rotation
//   3D-solver/core/functions3D.cuh:1086-1327  kernel_WENO_stage{0,1,2}
//   functions3D.cuh:763-1001              construct_V*, compute_beta_*, compute_alpha
//                                          (WENO-Z), compute_w, WENO_*, compute_Vs_*
//   main3D.cu:830-843                     WENO tables c_rl, beta_rl, dd, eps = 1e-10
//   functions3D.cuh:56-78                 Godunov fluxes
//   functions3D.cuh:660-756               cPML (corr each stage, psi advance in stage 2)
//   functions3D.cuh:1039-1047             kernel_inject_src_SSP
//   functions3D.cuh:462-565, 411-459, 263-271  immersed BC, trilinear, ADE, chi
//   common/filterABC.h:95-108             build_boundary_map
// The reference's tables are read as their semantic 4x3 / 3x3 layout (its flat()
// indexing of 12- and 9-float allocations runs past them; see main3D.cu:830-841).

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>

extern "C" {

typedef struct oracle_cfg {
    int Nx, Ny, Nz;
    double h, dt, c, rho, tol;
    float pml_keys[6];      // U D R L F B
    int pml_cells;
    double pml_sigma_max;   // <= 0: reference default
    double pml_kappa_max, pml_alpha_max;
} oracle_cfg;

}

namespace {

const double CRL[4][3] = {{11.0 / 6.0, -7.0 / 6.0, 1.0 / 3.0},
                          {1.0 / 3.0, 5.0 / 6.0, -1.0 / 6.0},
                          {-1.0 / 6.0, 5.0 / 6.0, 1.0 / 3.0},
                          {1.0 / 3.0, -7.0 / 6.0, 11.0 / 6.0}};
const double BRL[3][3] = {{3.0, -4.0, 1.0}, {1.0, 0.0, -1.0}, {1.0, -4.0, 3.0}};
const double DD[3] = {3.0 / 10.0, 3.0 / 5.0, 1.0 / 10.0};
const double EPS = 1e-10;
const int NG = 3;

struct G {
    int Nx, Ny, Nz;
    long nxy;
    long at(int i, int j, int k) const { return i + (long)Nx * (j + (long)k * Ny); }
};

// u along one axis around cell C with stride st: value at offset q
inline double A(const double *u, long C, long st, int q) { return u[C + q * st]; }

double construct_V(const double *u, long C, long st, int r, int idx) {
    double ret = 0.0;
    for (int l = 0; l <= 2; ++l) ret += CRL[r + idx][l] * A(u, C, st, -r + l);
    return ret;
}

double beta_r(const double *u, long C, long st, int r) {
    const double b1 = A(u, C, st, -r) - 2 * A(u, C, st, -r + 1) + A(u, C, st, -r + 2);
    const double b2 = BRL[r][0] * A(u, C, st, -r) + BRL[r][1] * A(u, C, st, -r + 1) + BRL[r][2] * A(u, C, st, -r + 2);
    return (13.0 * b1 * b1 / 12.0) + (0.25 * b2 * b2);
}

double alpha_r(const double *beta, int r, int SS) {
    const double tau5 = std::fabs(beta[0] - beta[2]);
    const double ratio = tau5 / (beta[r] + EPS);
    const int rr = SS * 3 - SS * 1 - r;
    return DD[std::abs(rr)] * (1.0 + ratio);
}

double weno(const double *u, long C, long st, int idx) {
    double V[3], beta[3], alpha[3], w[3];
    for (int r = 0; r < 3; ++r) V[r] = construct_V(u, C, st, r, idx);
    for (int r = 0; r < 3; ++r) beta[r] = beta_r(u, C, st, r);
    for (int r = 0; r < 3; ++r) alpha[r] = alpha_r(beta, r, std::abs(idx - 1));
    for (int r = 0; r < 3; ++r) {
        const double sum = alpha[0] + alpha[1] + alpha[2];
        w[r] = alpha[r] / sum;
    }
    double v = 0.0;
    for (int r = 0; r < 3; ++r) v += w[r] * V[r];
    return v;
}

void compute_Vs(double *QV, const double *u, long C, long st) {
    QV[0] = weno(u, C, st, 1);
    QV[1] = weno(u, C + st, st, 0);
    QV[2] = weno(u, C - st, st, 1);
    QV[3] = weno(u, C, st, 0);
}

inline double Fp(double pM, double pP, double uM, double uP, double rho, double c) {
    return 0.5 * (rho * c * c * (uP + uM) - c * (pP - pM));
}
inline double Fu(double pM, double pP, double uM, double uP, double rho, double c) {
    return 0.5 * ((pP + pM) / rho - c * (uP - uM));
}

struct Pml {
    int on, d;
    double smax, kmax, amax, amin;
    float U, D, R, L, F, B;
};

inline double cpml_corr(double D, double kappa, double psi) { return (1.0 / kappa - 1.0) * D + psi; }
// Recursive-convolution coefficients of the psi update (functions3D.cuh:669-673).
inline void cpml_coef(double kappa, double sigma, double alpha, double dt, double &b, double &ca) {
    b = std::exp(-(alpha + sigma / kappa) * dt);
    ca = (sigma > 0.0) ? (sigma / (kappa * (sigma + kappa * alpha))) * (b - 1.0) : 0.0;
}
inline void cpml_adv(double D, double kappa, double sigma, double alpha, double dt, double &psi) {
    double b, ca;
    cpml_coef(kappa, sigma, alpha, dt, b, ca);
    psi = b * psi + ca * D;
}

struct Fields {
    std::vector<double> p, u, v, w;
    void init(long N) { p.assign(N, 0.0); u.assign(N, 0.0); v.assign(N, 0.0); w.assign(N, 0.0); }
};

// cPML profile of one axis at index i (functions3D.cuh:692-719): depth x in [0, 1] from
// the interior edge (NG + d, or N-1-NG-d) to the outermost updated cell, for each enabled
// face (key_lo at the low end, key_hi at the high end), then
//   sigma = sigma_max x^m, kappa = 1 + (kappa_max - 1) x^m, m = 2,
//   alpha = alpha_min + (alpha_max - alpha_min)(1 - x), alpha_min = alpha_max / 10.
// Returns whether i lies in an enabled band (functions3D.cuh:721-723).
bool pml_profile(const Pml &P, int N, float key_lo, float key_hi, int i, double &sig, double &kap, double &alp) {
    const int d = P.d;
    const double m = 2.0;
    const double xm = std::fmin(1.0, std::fmax(0.0, ((double)(NG + d) - (double)i) / (double)d));
    const double xp = std::fmin(1.0, std::fmax(0.0, ((double)i - ((double)(N - 1 - NG) - (double)d)) / (double)d));
    sig = (double)key_lo * P.smax * std::pow(xm, m) + (double)key_hi * P.smax * std::pow(xp, m);
    const double dep = (double)key_lo * xm + (double)key_hi * xp;
    kap = 1.0 + (P.kmax - 1.0) * std::pow(dep, m);
    alp = P.amin + (P.amax - P.amin) * (1.0 - dep);
    return (key_lo > 0.f && i >= NG && i < NG + d) || (key_hi > 0.f && i >= N - NG - d && i <= N - 1 - NG);
}

// Face keys and profile constants (functions3D.cuh:687-719; sigma_max defaults to
// 0.5 c (m + 1) ln(1/R0) / (d h) with R0 = 1e-7, m = 2).
Pml make_pml(const oracle_cfg &cf) {
    Pml P;
    P.d = cf.pml_cells;
    P.U = cf.pml_keys[0]; P.D = cf.pml_keys[1]; P.R = cf.pml_keys[2];
    P.L = cf.pml_keys[3]; P.F = cf.pml_keys[4]; P.B = cf.pml_keys[5];
    P.on = (P.U || P.D || P.R || P.L || P.F || P.B) ? 1 : 0;
    P.smax = cf.pml_sigma_max;
    if (!(P.smax > 0.0)) P.smax = 0.5 * cf.c * (2.0 + 1.0) * std::log(1.0 / 1e-7) / ((double)P.d * cf.h);
    P.kmax = cf.pml_kappa_max;
    P.amax = cf.pml_alpha_max;
    P.amin = 0.1 * P.amax;
    return P;
}

void cpml_cell(const G &g, const Pml &P, std::vector<double> *psi, double &p, double &u, double &v, double &w,
               double Lpx, double Lpy, double Lpz, double Lux, double Lvy, double Lwz, int i, int j, int k,
               double dt, double wssp, bool advance) {
    const int d = P.d;
    double sig_x, kx, ax, sig_y, ky, ay, sig_z, kz, az;
    const bool bx = pml_profile(P, g.Nx, P.L, P.R, i, sig_x, kx, ax);
    const bool by = pml_profile(P, g.Ny, P.B, P.F, j, sig_y, ky, ay);
    const bool bz = pml_profile(P, g.Nz, P.D, P.U, k, sig_z, kz, az);
    const double wdt = wssp * dt;
    if (bx) {
        const int ix = (i < NG + d) ? (i - NG) : (i - (g.Nx - NG - d) + d);
        const long X = ix + (2L * d) * (j + (long)g.Ny * k);
        p += wdt * cpml_corr(Lpx, kx, psi[0][X]);
        u += wdt * cpml_corr(Lux, kx, psi[3][X]);
        if (advance) { cpml_adv(Lpx, kx, sig_x, ax, dt, psi[0][X]); cpml_adv(Lux, kx, sig_x, ax, dt, psi[3][X]); }
    }
    if (by) {
        const int jy = (j < NG + d) ? (j - NG) : (j - (g.Ny - NG - d) + d);
        const long Y = i + (long)g.Nx * (jy + (2L * d) * k);
        p += wdt * cpml_corr(Lpy, ky, psi[1][Y]);
        v += wdt * cpml_corr(Lvy, ky, psi[4][Y]);
        if (advance) { cpml_adv(Lpy, ky, sig_y, ay, dt, psi[1][Y]); cpml_adv(Lvy, ky, sig_y, ay, dt, psi[4][Y]); }
    }
    if (bz) {
        const int kz_ = (k < NG + d) ? (k - NG) : (k - (g.Nz - NG - d) + d);
        const long Z = i + (long)g.Nx * (j + (long)g.Ny * kz_);
        p += wdt * cpml_corr(Lpz, kz, psi[2][Z]);
        w += wdt * cpml_corr(Lwz, kz, psi[5][Z]);
        if (advance) { cpml_adv(Lpz, kz, sig_z, az, dt, psi[2][Z]); cpml_adv(Lwz, kz, sig_z, az, dt, psi[5][Z]); }
    }
}

// Stage s: in = stage input, base = U^n, out = stage output (functions3D.cuh:1086-1327).
void flux_stage(const G &g, const oracle_cfg &cf, const Pml &P, std::vector<double> *psi, const float *S, int s,
                const Fields &in, const Fields &base, Fields &out) {
    const double rho = cf.rho, c = cf.c, h = cf.h, dt = cf.dt;
    const float hf = (float)cf.h;
#pragma omp parallel for schedule(static)
    for (int k = NG; k < g.Nz - NG; ++k) {
        for (int j = NG; j < g.Ny - NG; ++j) {
            for (int i = NG; i < g.Nx - NG; ++i) {
                const long C = g.at(i, j, k);
                if (S[C] < hf) continue;
                double pV[4], uV[4], vV[4], wV[4];
                compute_Vs(pV, in.p.data(), C, 1);
                compute_Vs(uV, in.u.data(), C, 1);
                const double F_pP = Fp(pV[0], pV[1], uV[0], uV[1], rho, c);
                const double F_pM = Fp(pV[2], pV[3], uV[2], uV[3], rho, c);
                const double F_uP = Fu(pV[0], pV[1], uV[0], uV[1], rho, c);
                const double F_uM = Fu(pV[2], pV[3], uV[2], uV[3], rho, c);
                compute_Vs(pV, in.p.data(), C, g.Nx);
                compute_Vs(vV, in.v.data(), C, g.Nx);
                const double G_pP = Fp(pV[0], pV[1], vV[0], vV[1], rho, c);
                const double G_pM = Fp(pV[2], pV[3], vV[2], vV[3], rho, c);
                const double G_vP = Fu(pV[0], pV[1], vV[0], vV[1], rho, c);
                const double G_vM = Fu(pV[2], pV[3], vV[2], vV[3], rho, c);
                compute_Vs(pV, in.p.data(), C, g.nxy);
                compute_Vs(wV, in.w.data(), C, g.nxy);
                const double H_pP = Fp(pV[0], pV[1], wV[0], wV[1], rho, c);
                const double H_pM = Fp(pV[2], pV[3], wV[2], wV[3], rho, c);
                const double H_wP = Fu(pV[0], pV[1], wV[0], wV[1], rho, c);
                const double H_wM = Fu(pV[2], pV[3], wV[2], wV[3], rho, c);
                double uo, vo, wo, po, wssp;
                if (s == 0) {
                    uo = in.u[C] + dt * (-F_uP + F_uM) / h;
                    vo = in.v[C] + dt * (-G_vP + G_vM) / h;
                    wo = in.w[C] + dt * (-H_wP + H_wM) / h;
                    po = in.p[C] + dt * (-F_pP + F_pM - G_pP + G_pM - H_pP + H_pM) / h;
                    wssp = 1.0;
                } else if (s == 1) {
                    uo = 0.75 * base.u[C] + 0.25 * (in.u[C] + dt * ((-F_uP + F_uM) / h));
                    vo = 0.75 * base.v[C] + 0.25 * (in.v[C] + dt * ((-G_vP + G_vM) / h));
                    wo = 0.75 * base.w[C] + 0.25 * (in.w[C] + dt * ((-H_wP + H_wM) / h));
                    po = 0.75 * base.p[C] + 0.25 * (in.p[C] + dt * ((-F_pP + F_pM - G_pP + G_pM - H_pP + H_pM) / h));
                    wssp = 0.25;
                } else {
                    uo = (base.u[C] + 2.0 * (in.u[C] + dt * ((-F_uP + F_uM) / h))) / 3.0;
                    vo = (base.v[C] + 2.0 * (in.v[C] + dt * ((-G_vP + G_vM) / h))) / 3.0;
                    wo = (base.w[C] + 2.0 * (in.w[C] + dt * ((-H_wP + H_wM) / h))) / 3.0;
                    po = (base.p[C] + 2.0 * (in.p[C] + dt * ((-F_pP + F_pM - G_pP + G_pM - H_pP + H_pM) / h))) / 3.0;
                    wssp = 2.0 / 3.0;
                }
                if (P.on)
                    cpml_cell(g, P, psi, po, uo, vo, wo, (-F_pP + F_pM) / h, (-G_pP + G_pM) / h, (-H_pP + H_pM) / h,
                              (-F_uP + F_uM) / h, (-G_vP + G_vM) / h, (-H_wP + H_wM) / h, i, j, k, dt, wssp, s == 2);
                out.u[C] = uo; out.v[C] = vo; out.w[C] = wo; out.p[C] = po;
            }
        }
    }
}

double trilinear(const double *f, const G &g, double Dx, double Dy, double Dz, double h) {
    const double gx = Dx / h - 0.5, gy = Dy / h - 0.5, gz = Dz / h - 0.5;
    int i = (int)std::floor(gx), j = (int)std::floor(gy), k = (int)std::floor(gz);
    i = std::max(0, std::min(i, g.Nx - 2));
    j = std::max(0, std::min(j, g.Ny - 2));
    k = std::max(0, std::min(k, g.Nz - 2));
    const double sx = gx - (double)i, sy = gy - (double)j, sz = gz - (double)k;
    const double c00 = f[g.at(i, j, k)] * (1 - sx) + f[g.at(i + 1, j, k)] * sx;
    const double c10 = f[g.at(i, j + 1, k)] * (1 - sx) + f[g.at(i + 1, j + 1, k)] * sx;
    const double c01 = f[g.at(i, j, k + 1)] * (1 - sx) + f[g.at(i + 1, j, k + 1)] * sx;
    const double c11 = f[g.at(i, j + 1, k + 1)] * (1 - sx) + f[g.at(i + 1, j + 1, k + 1)] * sx;
    const double c0 = c00 * (1 - sy) + c10 * sy, c1 = c01 * (1 - sy) + c11 * sy;
    return c0 * (1 - sz) + c1 * sz;
}

double chi_layered(double s, double h, double tol) {
    if (s >= tol) return 0.0;
    if (s <= h) return 1.0;
    const double r = (s - h) / (tol - h);
    const double smooth = r * r * (3.0 - 2.0 * r);
    return 1.0 - smooth;
}

struct Mat {
    int K, L;
    const double *Rinf, *A, *zeta, *B, *C, *alpha, *beta;
};

double ade(int b, int m, double win, double dt, const Mat &M, double as, double bs, double cs, double *phi_l,
           const double *phi_m, const double *phi_r, double *p1_l, const double *p1_m, const double *p1_r,
           double *p2_l, const double *p2_m, const double *p2_r) {
    double sumA = 0.0, sumBC = 0.0;
    for (int k = 0; k < M.K; ++k) {
        const int mi = m * M.K + k;
        const long s = (long)b * M.K + k;
        const double L_phi = win - M.zeta[mi] * phi_r[s];
        phi_l[s] = as * phi_m[s] + bs * (phi_r[s] + dt * L_phi);
        phi_l[s] /= cs;
        sumA += M.A[mi] * phi_l[s];
    }
    for (int l = 0; l < M.L; ++l) {
        const int mi = m * M.L + l;
        const long s = (long)b * M.L + l;
        const double L1 = win - (M.alpha[mi] * p1_r[s] + M.beta[mi] * p2_r[s]);
        const double L2 = M.beta[mi] * p1_r[s] - M.alpha[mi] * p2_r[s];
        p1_l[s] = as * p1_m[s] + bs * (p1_r[s] + dt * L1);
        p1_l[s] /= cs;
        p2_l[s] = as * p2_m[s] + bs * (p2_r[s] + dt * L2);
        p2_l[s] /= cs;
        sumBC += M.B[mi] * p1_l[s] + M.C[mi] * p2_l[s];
    }
    return M.Rinf[m] * win + sumA + sumBC;
}

// kernel_apply_BC_immersed_sdf over the whole interior (functions3D.cuh:462-565)
void boundary(const G &g, const oracle_cfg &cf, const float *S, const uint8_t *oid, const int *bmap, const Mat &M,
              Fields &out, const Fields &old, double as, double bs, double cs, double *phi_l, const double *phi_m,
              const double *phi_r, double *p1_l, const double *p1_m, const double *p1_r, double *p2_l,
              const double *p2_m, const double *p2_r) {
    const double h = cf.h, rho = cf.rho, c = cf.c, dt = cf.dt, tol = cf.tol;
#pragma omp parallel for schedule(static)
    for (int k = NG; k < g.Nz - NG; ++k) {
        for (int j = NG; j < g.Ny - NG; ++j) {
            for (int i = NG; i < g.Nx - NG; ++i) {
                const long C = g.at(i, j, k);
                const double s = S[C];
                const uint8_t o = oid[C];
                const int b = bmap[C];
                if (o == 0 || o == 0xFF || b < 0) continue;
                const double gx = ((double)S[g.at(i + 1, j, k)] - (double)S[g.at(i - 1, j, k)]) * (0.5 / h);
                const double gy = ((double)S[g.at(i, j + 1, k)] - (double)S[g.at(i, j - 1, k)]) * (0.5 / h);
                const double gz = ((double)S[g.at(i, j, k + 1)] - (double)S[g.at(i, j, k - 1)]) * (0.5 / h);
                double gn = std::sqrt(gx * gx + gy * gy + gz * gz);
                gn = std::fmax(gn, 1e-12);
                const double nx = -gx / gn, ny = -gy / gn, nz = -gz / gn;
                const double Cx = (i + 0.5) * h, Cy = (j + 0.5) * h, Cz = (k + 0.5) * h;
                const double Wx = Cx + s * nx, Wy = Cy + s * ny, Wz = Cz + s * nz;
                const double l = s + 2.0 * h;
                const double Dx = Wx - l * nx, Dy = Wy - l * ny, Dz = Wz - l * nz;
                const double pD = trilinear(old.p.data(), g, Dx, Dy, Dz, h);
                const double uD = trilinear(old.u.data(), g, Dx, Dy, Dz, h);
                const double vD = trilinear(old.v.data(), g, Dx, Dy, Dz, h);
                const double wD = trilinear(old.w.data(), g, Dx, Dy, Dz, h);
                const double un = uD * nx + vD * ny + wD * nz;
                const double utx = uD - un * nx, uty = vD - un * ny, utz = wD - un * nz;
                const double Z0 = rho * c;
                const double wout = pD + Z0 * un;
                const int m = o - 1;
                const double win = ade(b, m, wout, dt, M, as, bs, cs, phi_l, phi_m, phi_r, p1_l, p1_m, p1_r, p2_l,
                                       p2_m, p2_r);
                const double pb = 0.5 * (win + wout);
                const double unb = (wout - win) / (2.0 * Z0);
                const double ub = unb * nx + utx, vb = unb * ny + uty, wb = unb * nz + utz;
                const double chi = chi_layered(s, h, tol);
                out.u[C] = (1.0 - chi) * out.u[C] + chi * ub;
                out.v[C] = (1.0 - chi) * out.v[C] + chi * vb;
                out.w[C] = (1.0 - chi) * out.w[C] + chi * wb;
                out.p[C] = (1.0 - chi) * out.p[C] + chi * pb;
            }
        }
    }
}

}  // namespace

extern "C" {

// Runs nsteps steps from step index step0 (source sample 2n, 2n+2, 2n+1 per stage).
// IC arrays (nullable) initialise all three field sets. ir[n_rec*nsteps], vel[n_rec*nsteps*3]
// and the final U^n (p_out.. nullable) are returned.
int oracle_run(const oracle_cfg *cf, const float *S, const uint8_t *oid, int nMat, int K, int L, const double *Rinf,
               const double *A, const double *zeta, const double *B, const double *Cc, const double *alpha,
               const double *beta, int n_src, const int64_t *scell, const double *sval, int src_opt,
               const double *u_s, const double *freq, int n_rec, const int64_t *rcell, const double *p0,
               const double *u0, const double *v0, const double *w0, int64_t step0, int64_t nsteps, double *ir,
               double *vel, double *p_out, double *u_out, double *v_out, double *w_out) {
    if (!cf || !S || !oid || nMat < 1 || K < 1 || L < 1) return -1;
    G g{cf->Nx, cf->Ny, cf->Nz, (long)cf->Nx * cf->Ny};
    const long N = (long)g.Nx * g.Ny * g.Nz;
    Fields U[3];
    for (auto &f : U) f.init(N);
    const double *ic[4] = {p0, u0, v0, w0};
    for (int s = 0; s < 3; ++s) {
        std::vector<double> *dst[4] = {&U[s].p, &U[s].u, &U[s].v, &U[s].w};
        for (int f = 0; f < 4; ++f)
            if (ic[f]) std::memcpy(dst[f]->data(), ic[f], sizeof(double) * N);
    }
    std::vector<int> bmap(N);
    int nB = 0;
    for (long c = 0; c < N; ++c) bmap[c] = (oid[c] != 0xFF && oid[c] != 0) ? nB++ : -1;
    std::vector<double> phi[3], p1[3], p2[3];   // phi, phi_old, phi_nu
    for (int q = 0; q < 3; ++q) {
        phi[q].assign((size_t)nB * K, 0.0);
        p1[q].assign((size_t)nB * L, 0.0);
        p2[q].assign((size_t)nB * L, 0.0);
    }
    const Pml P = make_pml(*cf);
    std::vector<double> psi[6];
    const long d2 = 2L * P.d;
    psi[0].assign(d2 * g.Ny * g.Nz, 0.0); psi[3].assign(d2 * g.Ny * g.Nz, 0.0);
    psi[1].assign((long)g.Nx * d2 * g.Nz, 0.0); psi[4].assign((long)g.Nx * d2 * g.Nz, 0.0);
    psi[2].assign((long)g.Nx * g.Ny * d2, 0.0); psi[5].assign((long)g.Nx * g.Ny * d2, 0.0);
    Mat M{K, L, Rinf, A, zeta, B, Cc, alpha, beta};
    const double As = 6.0 * cf->h * cf->h;
    const double h = cf->h, c = cf->c, rho = cf->rho, dt = cf->dt;
    auto inject = [&](std::vector<double> &p, double wssp, long nn) {
        for (int s = 0; s < n_src; ++s) {
            const double amp = (src_opt == 2) ? std::sin(2.0 * M_PI * freq[s] * (double)nn * (0.5 * dt)) : u_s[nn];
            p[scell[s]] += wssp * dt * c * c * sval[s] * amp * rho * As / (h * h * h);
        }
    };
    int i0 = 0, i1 = 1, i2 = 2;          // U0, U1, U2 (swapped by index)
    int ph = 0, po = 1, pn = 2;          // phi, phi_old, phi_nu
    for (int64_t q = 0; q < nsteps; ++q) {
        const long n = (long)(step0 + q);
        flux_stage(g, *cf, P, psi, S, 0, U[i0], U[i0], U[i1]);
        inject(U[i1].p, 1.0, 2 * n);
        boundary(g, *cf, S, oid, bmap.data(), M, U[i1], U[i0], 0.0, 1.0, 1.0, phi[ph].data(), phi[po].data(),
                 phi[po].data(), p1[ph].data(), p1[po].data(), p1[po].data(), p2[ph].data(), p2[po].data(),
                 p2[po].data());
        flux_stage(g, *cf, P, psi, S, 1, U[i1], U[i0], U[i2]);
        inject(U[i2].p, 0.25, 2 * n + 2);
        boundary(g, *cf, S, oid, bmap.data(), M, U[i2], U[i1], 3.0, 1.0, 4.0, phi[pn].data(), phi[po].data(),
                 phi[ph].data(), p1[pn].data(), p1[po].data(), p1[ph].data(), p2[pn].data(), p2[po].data(),
                 p2[ph].data());
        flux_stage(g, *cf, P, psi, S, 2, U[i2], U[i0], U[i1]);
        inject(U[i1].p, 2.0 / 3.0, 2 * n + 1);
        boundary(g, *cf, S, oid, bmap.data(), M, U[i1], U[i2], 1.0, 2.0, 3.0, phi[ph].data(), phi[po].data(),
                 phi[pn].data(), p1[ph].data(), p1[po].data(), p1[pn].data(), p2[ph].data(), p2[po].data(),
                 p2[pn].data());
        std::swap(i0, i1);
        std::swap(po, ph);
        for (int r = 0; r < n_rec; ++r) {
            if (ir) ir[r * nsteps + q] = U[i0].p[rcell[r]];
            if (vel) {
                double *o = vel + 3 * (r * nsteps + q);
                o[0] = U[i0].u[rcell[r]]; o[1] = U[i0].v[rcell[r]]; o[2] = U[i0].w[rcell[r]];
            }
        }
    }
    if (p_out) std::memcpy(p_out, U[i0].p.data(), sizeof(double) * N);
    if (u_out) std::memcpy(u_out, U[i0].u.data(), sizeof(double) * N);
    if (v_out) std::memcpy(v_out, U[i0].v.data(), sizeof(double) * N);
    if (w_out) std::memcpy(w_out, U[i0].w.data(), sizeof(double) * N);
    return 0;
}

// Single reference WENO5-Z reconstruction of the 5-cell window u[0..4] centred on u[2]:
// idx = 1 gives V^-_{i+1/2}, idx = 0 gives V^+_{i-1/2} (functions3D.cuh:863-907).
double oracle_weno5(const double *u5, int idx) { return weno(u5 + 2, 0, 1, idx); }

// cPML profile the run uses on axis `axis` (0 x: keys L/R, 1 y: B/F, 2 z: D/U) of an
// N-cell padded axis at index i: out = {sigma, kappa, alpha, b, a} (b, a: the psi
// update coefficients). Returns 1 in an enabled band, else 0.
int oracle_pml_profile(const oracle_cfg *cf, int axis, int N, int i, double *out) {
    const Pml P = make_pml(*cf);
    const float lo = axis == 0 ? P.L : (axis == 1 ? P.B : P.D);
    const float hi = axis == 0 ? P.R : (axis == 1 ? P.F : P.U);
    double sig, kap, alp, b, ca;
    const bool in = pml_profile(P, N, lo, hi, i, sig, kap, alp);
    cpml_coef(kap, sig, alp, cf->dt, b, ca);
    out[0] = sig; out[1] = kap; out[2] = alp; out[3] = b; out[4] = ca;
    return in ? 1 : 0;
}

// Godunov flux pair of one face (functions3D.cuh:56-62).
void oracle_godunov(double pM, double pP, double uM, double uP, double rho, double c, double *Fp_out,
                    double *Fu_out) {
    *Fp_out = Fp(pM, pP, uM, uP, rho, c);
    *Fu_out = Fu(pM, pP, uM, uP, rho, c);
}

// ADE of one boundary cell driven by w(t) sampled at t, t+dt, t+dt/2 per step, exactly as
// the boundary kernel runs it (three SSP-RK3 stages with the (a,b,c) of main3D.cu:1290-1306).
// w3[3*n + s] is the stage-s drive of step n; out[3*n + s] the stage-s admittance output.
void oracle_ade_series(int K, int L, double Rinf, const double *A, const double *zeta, const double *B,
                       const double *C, const double *alpha, const double *beta, double dt, const double *w3,
                       int64_t nsteps, double *out) {
    Mat M{K, L, &Rinf, A, zeta, B, C, alpha, beta};
    std::vector<double> phi[3], p1[3], p2[3];
    for (int q = 0; q < 3; ++q) { phi[q].assign(K, 0.0); p1[q].assign(L, 0.0); p2[q].assign(L, 0.0); }
    int ph = 0, po = 1, pn = 2;
    for (int64_t n = 0; n < nsteps; ++n) {
        out[3 * n + 0] = ade(0, 0, w3[3 * n + 0], dt, M, 0.0, 1.0, 1.0, phi[ph].data(), phi[po].data(), phi[po].data(),
                             p1[ph].data(), p1[po].data(), p1[po].data(), p2[ph].data(), p2[po].data(), p2[po].data());
        out[3 * n + 1] = ade(0, 0, w3[3 * n + 1], dt, M, 3.0, 1.0, 4.0, phi[pn].data(), phi[po].data(), phi[ph].data(),
                             p1[pn].data(), p1[po].data(), p1[ph].data(), p2[pn].data(), p2[po].data(), p2[ph].data());
        out[3 * n + 2] = ade(0, 0, w3[3 * n + 2], dt, M, 1.0, 2.0, 3.0, phi[ph].data(), phi[po].data(), phi[pn].data(),
                             p1[ph].data(), p1[po].data(), p1[pn].data(), p2[ph].data(), p2[po].data(), p2[pn].data());
        std::swap(po, ph);
    }
}

}  // extern "C"
