// Project file formats of the 3D solver and the shared source waveform.
//
// config.txt      3D-solver/core/main3D.cu:114-128 (first float per line, '#' comments)
// poles.txt       main3D.cu:592-603 ("K_m L_m", floored to 1)
// <layer>.fit     common/filterABC.h read_fit_file:31-92 (section grammar, '#' lines skipped)
// src.txt/rec.txt main3D.cu:861-906 / 919-953 (cell = (int)(x/h) + 3)
// waveforms       main3D.cu:963-1058 (PCS, Ricker), amplitude scaling main3D.cu:1063-1110
// IR/VEL files    main3D.cu:1242-1256, 1397-1414 (little-endian float32, no header)
// <layer>.mat     octave-band absorption, first 8 floats ('#' comments; 3D-solver/include/
//                 ism3D.h ism_load_impedances:98-152, missing file -> 0)
// T60 estimate    Eyring with air absorption (ism3D.h estimate_T60s:735-751); air
//                 attenuation of ISO 9613-1 (3D-solver/core/functions3D.cuh frO/frN/
//                 alpha_air/p_atm:15-46, main3D.cu:452-484), evaluated with the same
//                 float/double mixing as the reference's host build

#include "host.hpp"
#include "../../include/fases_core.h"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

using namespace fases;

namespace {

// Skips blanks, newlines and whole '#' lines before the next datum of a .fit file.
void next_datum(std::istream &in) {
    for (;;) {
        in >> std::ws;
        const int ch = in.peek();
        if (ch == '#') { in.ignore(std::numeric_limits<std::streamsize>::max(), '\n'); continue; }
        if (ch == '\n') { in.get(); continue; }
        return;
    }
}

inline int grid_cell(float x, float h) { return (int)(x / h) + 3; }

}  // namespace

extern "C" {

int fases_params_default(fases_params *p) {
    if (!p) return FASES_E_ARG;
    p->device = 0;
    p->precision = FASES_FP64;
    p->h = 0.0;
    p->dt = 0.0;
    p->c = 343.0;
    p->rho = 1.225;
    p->tol = 0.0;
    for (int i = 0; i < 6; ++i) p->pml_keys[i] = 0.f;
    p->pml_cells = 20;
    p->pml_sigma_max = 0.0;
    p->pml_kappa_max = 5.0;
    p->pml_alpha_max = M_PI * 20.0;
    p->fp16_rescale_every = 100;
    p->use_graph = 0;
    p->kernel_variant = 0;
    p->tile_z = 0;
    p->adjoint_variant = 0;
    p->tile_rows = 0;
    p->tile_cols = 0;
    p->source_patch = 0;
    return FASES_OK;
}

int fases_read_config(const char *path, float *vals, int maxn, int *n) {
    if (!path || !n || (maxn > 0 && !vals)) { set_global_error("fases_read_config: null argument"); return FASES_E_ARG; }
    std::ifstream f(path);
    if (!f) { set_global_error(std::string("cannot open config: ") + path); return FASES_E_IO; }
    std::string line;
    int cnt = 0;
    while (std::getline(f, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.erase(hash);
        std::istringstream ss(line);
        float v;
        if (ss >> v) {
            if (cnt < maxn) vals[cnt] = v;
            ++cnt;
        }
    }
    *n = cnt;
    return FASES_OK;
}

int fases_read_poles(const char *path, int *K, int *L) {
    if (!path || !K || !L) return FASES_E_ARG;
    std::ifstream f(path);
    if (!f) { set_global_error(std::string("Cannot open poles file: ") + path); return FASES_E_IO; }
    int k = 0, l = 0;
    f >> k >> l;
    if (f.fail()) { set_global_error(std::string("Parse failed while reading K_m L_m from: ") + path); return FASES_E_PARSE; }
    *K = k < 1 ? 1 : k;
    *L = l < 1 ? 1 : l;
    return FASES_OK;
}

int fases_read_fit(const char *path, int m, int K, int L,
                   double *Rinf, double *A, double *zeta,
                   double *B, double *C, double *alpha, double *beta) {
    if (!path || !Rinf || !A || !zeta || !B || !C || !alpha || !beta || m < 0 || K < 0 || L < 0) return FASES_E_ARG;
    std::ifstream in(path);
    if (!in.is_open()) { set_global_error(std::string("Cannot open file: ") + path); return FASES_E_IO; }
    auto fail = [&](const std::string &what) {
        set_global_error("Parse failed while reading " + what + " from file: " + path);
        return FASES_E_PARSE;
    };
    next_datum(in);
    in >> Rinf[m];
    if (!in) return fail("R_inf");
    int kl = 0, ll = 0;
    next_datum(in);
    in >> kl >> ll;
    if (!in) return fail("K L");
    if (kl < 0 || kl > K) {
        set_global_error("Invalid K_local=" + std::to_string(kl) + " (K_m=" + std::to_string(K) + ") in file: " + path);
        return FASES_E_PARSE;
    }
    if (ll < 0 || ll > L) {
        set_global_error("Invalid L_local=" + std::to_string(ll) + " (L_m=" + std::to_string(L) + ") in file: " + path);
        return FASES_E_PARSE;
    }
    next_datum(in);
    for (int k = 0; k < kl; ++k) {
        in >> A[(size_t)m * K + k] >> zeta[(size_t)m * K + k];
        if (!in) return fail("A_k zeta_k row " + std::to_string(k));
    }
    next_datum(in);
    for (int l = 0; l < ll; ++l) {
        const size_t q = (size_t)m * L + l;
        in >> B[q] >> C[q] >> alpha[q] >> beta[q];
        if (!in) return fail("B_l C_l alpha_l beta_l row " + std::to_string(l));
    }
    return FASES_OK;
}

int fases_read_sources(const char *path, float h, int Nx, int Ny, int Nz, float default_f0,
                       int maxn, int64_t *cell, float *lp, float *f0, float *xyz, int *n) {
    if (!path || !n) return FASES_E_ARG;
    std::ifstream f(path);
    if (!f) { set_global_error(std::string("Failed to open sources: ") + path); return FASES_E_IO; }
    std::string line;
    int cnt = 0;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        float x, y, z, s, fr;
        if (!(ss >> x >> y >> z >> s)) continue;
        if (!(ss >> fr)) fr = default_f0;
        const int i = grid_cell(x, h), j = grid_cell(y, h), k = grid_cell(z, h);
        if (!(i >= 0 && i < Nx && j >= 0 && j < Ny && k >= 0 && k < Nz)) continue;
        if (cnt < maxn) {
            if (cell) cell[cnt] = (int64_t)i + (int64_t)Nx * ((int64_t)j + (int64_t)k * Ny);
            if (lp) lp[cnt] = s;
            if (f0) f0[cnt] = fr;
            if (xyz) { xyz[3 * cnt] = x; xyz[3 * cnt + 1] = y; xyz[3 * cnt + 2] = z; }
        }
        ++cnt;
    }
    *n = cnt;
    return FASES_OK;
}

int fases_read_receivers(const char *path, float h, int Nx, int Ny, int Nz,
                         int maxn, int64_t *cell, float *xyz, int *n) {
    if (!path || !n) return FASES_E_ARG;
    std::ifstream f(path);
    if (!f) { set_global_error(std::string("Failed to open receivers: ") + path); return FASES_E_IO; }
    std::string line;
    int cnt = 0;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        float x, y, z;
        if (!(ss >> x >> y >> z)) continue;
        const int i = grid_cell(x, h), j = grid_cell(y, h), k = grid_cell(z, h);
        if (!(i >= 0 && i < Nx && j >= 0 && j < Ny && k >= 0 && k < Nz)) {
            set_global_error("Receiver(s) outside domain");
            return FASES_E_RANGE;
        }
        if (cnt < maxn) {
            if (cell) cell[cnt] = (int64_t)i + (int64_t)Nx * ((int64_t)j + (int64_t)k * Ny);
            if (xyz) { xyz[3 * cnt] = x; xyz[3 * cnt + 1] = y; xyz[3 * cnt + 2] = z; }
        }
        ++cnt;
    }
    *n = cnt;
    return FASES_OK;
}

int fases_source_waveform(int src_opt, float dt, double f, float f_cutoff, int64_t Nt, double *u_s) {
    if (!u_s || Nt < 1 || !(dt > 0.f)) return FASES_E_ARG;
    const int64_t ns = 2 * Nt + 1;
    for (int64_t t = 0; t < ns; ++t) u_s[t] = 0.0;
    const double dth = 0.5 * (double)dt;
    if (src_opt == FASES_SRC_PCS) {
        // Sheaffer, van Walstijn & Fazenda (2012) physically constrained source: a Gaussian
        // force through the bilinear-transformed mass-spring-damper (M, R, K) of a sphere.
        const double Q = 1.0, w0 = 2.0 * M_PI * f;
        const double bs = w0 / std::tan(0.5 * w0 * dth);
        const double M = 0.025, R = w0 * M / Q, Ks = M * w0 * w0;
        const double den = M * bs * bs + R * bs + Ks;
        const double b0 = bs / den, b2 = -b0;
        const double a1 = 2.0 * (Ks - M * bs * bs) / den;
        const double a2 = 1.0 - 2.0 * R * bs / den;
        const double sig = 0.425 / (double)f_cutoff;
        const int lag = (int)std::ceil(8.5 * sig / dth);
        std::vector<double> F((size_t)ns, 0.0);
        for (int64_t t = 0; t < ns; ++t) {
            const double td = (double)(t - lag) * dth;
            F[(size_t)t] = std::exp(-td * td / (2.0 * sig * sig));
        }
        for (int64_t t = 2; t < ns; ++t)
            u_s[t] = b0 * F[(size_t)t] + b2 * F[(size_t)t - 2] - a1 * u_s[t - 1] - a2 * u_s[t - 2];
        return FASES_OK;
    }
    if (src_opt == FASES_SRC_RICKER) {
        // Surface velocity = Gaussian first derivative x e^{-x^2/2}; its time derivative
        // (what radiates) is the Ricker wavelet peaking at f.
        const double sig = 1.0 / (std::sqrt(2.0) * M_PI * f);
        const double lag = (double)(int)std::ceil(5.0 * sig / dth);
        for (int64_t t = 0; t < ns; ++t) {
            const double x = ((double)t - lag) * dth / sig;
            u_s[t] = x * std::exp(-0.5 * x * x);
        }
        return FASES_OK;
    }
    set_global_error("fases_source_waveform: unknown source option");
    return FASES_E_ARG;
}

int fases_source_pref(const double *u_s, int64_t Nt, float dt, float h, float rho,
                      double *p_ref, double *p_peak) {
    if (!u_s || Nt < 1) return FASES_E_ARG;
    const double As = 6.0 * (double)h * (double)h;
    const double dth = 0.5 * (double)dt;
    double peak = 0.0;
    const int64_t ns = 2 * Nt + 1;
    if (p_ref) for (int64_t t = 0; t < ns; ++t) p_ref[t] = 0.0;
    for (int64_t t = 1; t < 2 * Nt; ++t) {
        const double dq = As * (u_s[t + 1] - u_s[t - 1]) / (2.0 * dth);
        const double pr = (double)rho * dq / (4.0 * M_PI);
        if (p_ref) p_ref[t] = pr;
        peak = std::fmax(peak, std::fabs(pr));
    }
    if (p_peak) *p_peak = peak;
    return FASES_OK;
}

int fases_write_f32(const char *path, const double *x, int64_t n) {
    if (!path || (n > 0 && !x)) return FASES_E_ARG;
    std::FILE *fp = std::fopen(path, "wb");
    if (!fp) { set_global_error(std::string("cannot write: ") + path); return FASES_E_IO; }
    std::vector<float> buf((size_t)n);
    for (int64_t i = 0; i < n; ++i) buf[(size_t)i] = (float)x[i];
    const size_t w = n > 0 ? std::fwrite(buf.data(), sizeof(float), (size_t)n, fp) : 0;
    std::fclose(fp);
    if (w != (size_t)n) { set_global_error(std::string("short write: ") + path); return FASES_E_IO; }
    return FASES_OK;
}

// Air attenuation m [Np/m] at frequency f of ISO 9613-1 for temperature T [K],
// atmospheric pressure p_a [Pa] and molar water-vapour concentration hm [%]: relaxation
// frequencies of oxygen and nitrogen, then the classical + rotational term A and the two
// vibrational terms, scaled from dB to Np (x 8.686 x 0.115). Each intermediate is
// rounded to float where the reference stores one.
static float iso_relax_o2(float p_a, float hm) {
    return p_a * (24.0 + 4.04 * 1e4 * hm * (0.02 + hm) / (0.391 + hm)) / 101325;
}
static float iso_relax_n2(float p_a, float hm, float T) {
    const float a = p_a * std::sqrt(293.15 / T) / 101325;
    const float b = 9.0 + 280.0 * hm * std::exp(-4.17 * (std::pow(T / 293.15, -1.0 / 3.0) - 1.0));
    return a * b;
}
static float iso_attenuation(float p_a, float hm, float T, float f) {
    const float a = 1.84 * 1e-11 * 101325 * std::sqrt(T / 293.15) / p_a;
    const float b = std::pow(T / 293.15, -5.0 / 2.0);
    const float fo = iso_relax_o2(p_a, hm), fn = iso_relax_n2(p_a, hm, T);
    const float co = 0.01275 * std::exp(-2239.1 / T) * (1.0 / (fo + (f * f / fo)));
    const float cn = 0.10680 * std::exp(-3352.0 / T) * (1.0 / (fn + (f * f / fn)));
    return 0.115 * 8.686 * f * f * (a + b * (co + cn));
}
// Barometric pressure at altitude alt [m] and temperature T [K].
static float barometric(float alt, float T) {
    const float M_air = 0.02897, g = 9.8067, R = 8.314;
    return 101325 * std::exp(-M_air * g * alt / (R * T));
}

int fases_air_attenuation(float T, float alt, float hm, const float *freq, int n, float *m) {
    if (n < 0 || (n > 0 && (!freq || !m))) { set_global_error("fases_air_attenuation: bad argument"); return FASES_E_ARG; }
    const float p_a = barometric(alt, T);
    for (int b = 0; b < n; ++b) m[b] = iso_attenuation(p_a, hm, T, freq[b]);
    return FASES_OK;
}

int fases_read_mat(const char *path, float *alpha, int *n_read) {
    if (!path || !alpha) { set_global_error("fases_read_mat: null argument"); return FASES_E_ARG; }
    for (int b = 0; b < 8; ++b) alpha[b] = 0.f;
    if (n_read) *n_read = 0;
    std::ifstream f(path);
    if (!f) { set_global_error(std::string("cannot open material file: ") + path); return FASES_E_IO; }
    std::string line;
    int n = 0;
    while (n < 8 && std::getline(f, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.erase(hash);
        std::istringstream ss(line);
        float v;
        while (n < 8 && (ss >> v)) alpha[n++] = v;
    }
    if (n_read) *n_read = n;
    return FASES_OK;
}

int fases_estimate_t60(float V, float c, float A_tot, int n, const float *A, const float *alpha, const float *m_air,
                       float *T60, float *f_schroeder) {
    if (n < 0 || (n > 0 && (!A || !alpha)) || !m_air || !T60) {
        set_global_error("fases_estimate_t60: bad argument");
        return FASES_E_ARG;
    }
    // Eyring with the air term: T = 55.25 V / (-c S ln(1 - a_mean) + 4 (2 m) V c), the mean
    // absorption area-weighted; a band without absorption (a_mean < 1e-6) does not decay
    const float num = 55.25 * V;
    for (int b = 0; b < 8; ++b) {
        float am = 0.0;
        for (int i = 0; i < n; ++i) am += A[i] * alpha[(size_t)i * 8 + b];
        am /= A_tot;
        if (am < 1e-6f) { T60[b] = 1e9f; continue; }
        const float den = -c * A_tot * std::log(1.0 - am) + 4.0 * (2.0f * m_air[b]) * V * c;
        T60[b] = num / den;
    }
    // Schroeder frequency from the 500 Hz band; 0 for a band that does not decay
    if (f_schroeder) *f_schroeder = (T60[3] < 1e6f) ? 2000.0f * std::sqrt(T60[3] / V) : 0.0f;
    return FASES_OK;
}

}  // extern "C"
