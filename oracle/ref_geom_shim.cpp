// ORACLE -- test infrastructure only. Exposes the reference's own CPU geometry path
// (3D-solver/include/parsing3D.h parse_obj3d / build_sdf / build_oid_dense /
// mesh_group_areas, common/filterABC.h read_fit_file / build_boundary_map, the
// reverberation estimate of 3D-solver/include/ism3D.h ism_load_impedances_layered /
// estimate_T60s, the image-source method ism_solve and the ray tracer of
// 3D-solver/include/raytrace3D.h rt_build / rt_trace_late) through a C ABI so tests can
// check the product's restatement bit for bit. Built only by oracle/build_ref.sh from
// the sources where they lie under the reference checkout, into oracle/_ref/.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include "parsing3D.h"
#include "filterABC.h"
#include "ism3D.h"
#include "raytrace3D.h"

extern "C" {

// OBJ -> padded grid dims, S and oid (malloc'd, release with ref_free).
int ref_scene(const char *obj, float h, int *dims, float **S_out, uint8_t **oid_out) {
    Scene3D sc;
    if (!parse_obj3d(obj, &sc, h)) return -1;
    int Nx = 0, Ny = 0, Nz = 0;
    float Lx = 0, Ly = 0, Lz = 0;
    float *S = nullptr;
    if (!build_sdf(&sc, h, &Nx, &Ny, &Nz, &Lx, &Ly, &Lz, &S)) return -2;
    uint8_t *oid = nullptr;
    if (!build_oid_dense(sc, S, Nx, Ny, Nz, h, 3.0f * h, &oid)) { std::free(S); return -3; }
    dims[0] = Nx; dims[1] = Ny; dims[2] = Nz;
    *S_out = S;
    *oid_out = oid;
    return 0;
}

void ref_free(void *p) { std::free(p); }

// Group areas and total area of the placed mesh; returns the group count.
int ref_group_areas(const char *obj, float h, float *A, int maxn, float *A_tot) {
    Scene3D sc;
    if (!parse_obj3d(obj, &sc, h)) return -1;
    float tot = 0.f;
    const std::vector<float> a = mesh_group_areas(sc, &tot);
    for (size_t i = 0; i < a.size() && (int)i < maxn; ++i) A[i] = a[i];
    *A_tot = tot;
    return (int)a.size();
}

// T60 per band from the scene's group areas, the .mat files of mat_dir named after the
// groups (layer names with BoundingBox first, as main3D.cu:434 builds them) and m_air[8].
int ref_t60(const char *obj, float h, const char *mat_dir, float V, float c, const float *m_air, float *T60,
            float *alpha_out, int maxn) {
    Scene3D sc;
    if (!parse_obj3d(obj, &sc, h)) return -1;
    float tot = 0.f;
    const std::vector<float> A = mesh_group_areas(sc, &tot);
    std::vector<std::string> names;
    scene_layer_names_3d(sc, names);
    const std::vector<MaterialImpedance> Z = ism_load_impedances_layered("", names, mat_dir);
    std::vector<float> T(8, 0.5f), m(m_air, m_air + 8);
    estimate_T60s(T, V, c, tot, A, Z, m);
    for (int b = 0; b < 8; ++b) T60[b] = T[(size_t)b];
    for (size_t i = 0; i < Z.size() && (int)i < maxn; ++i)
        for (int b = 0; b < 8; ++b) alpha_out[i * 8 + b] = Z[i].alpha[(size_t)b];
    return (int)Z.size();
}

int ref_boundary_map(uint8_t *oid, int N, int *bmap) { return build_boundary_map(oid, N, bmap); }

// ism_solve (object-space endpoints, as main3D.cu:1481-1488 calls it, no impedances): per
// (s, r) in row-major order the direct distance and n_img[s nr + r] image paths, packed one
// after another: distance, doa[3], amp, order, then per bounce group / cos / angle. Returns
// the order used, or -1 when the OBJ does not parse, -2 when a capacity is exceeded.
int ref_ism(const char *obj, float h, const float *src, int ns, const float *rec, int nr, int order, int cap_img,
            int cap_b, float *direct, int *n_img, float *dist, float *doa, float *amp, int *ord, int *group,
            float *cosv, float *angle) {
    Scene3D sc;
    if (!parse_obj3d(obj, &sc, h)) return -1;
    ISMConfig cfg;
    cfg.frame = ISMFrame::OBJECT_SPACE;
    std::vector<float> sa(src, src + 3 * ns), ra(rec, rec + 3 * nr);
    const ISMResult Q = ism_solve(sc, sa, ra, {}, {}, order, {}, cfg);
    int ni = 0, nb = 0;
    for (int s = 0; s < ns; ++s)
        for (int r = 0; r < nr; ++r) {
            direct[s * nr + r] = Q.direct_at(s, r);
            const auto &im = Q.images_at(s, r);
            n_img[s * nr + r] = (int)im.size();
            for (const ISMImagePath &p : im) {
                if (ni >= cap_img || nb + (int)p.cos_inc.size() > cap_b) return -2;
                dist[ni] = p.distance;
                for (int k = 0; k < 3; ++k) doa[3 * ni + k] = p.doa[(size_t)k];
                amp[ni] = p.amp;
                ord[ni] = (int)p.cos_inc.size();
                for (size_t k = 0; k < p.cos_inc.size(); ++k, ++nb) {
                    group[nb] = p.group[k];
                    cosv[nb] = p.cos_inc[k];
                    angle[nb] = p.angle_rad[k];
                }
                ++ni;
            }
        }
    return Q.order_used;
}

// rt_build + rt_trace_late of the scene (main3D.cu:1496, 1583): energy echograms
// e_spec / e_diff [8 nr Nt]. This build has no OpenMP, so the sums run in ray order.
int ref_rt(const char *obj, float h, const float *src, int ns, const float *src_p, const float *rec, int nr,
           const float *alpha_gb, const float *m_air, const float *sctr_g, float c, float fs, int Nt, int n_rays,
           int max_order, float radius, int ism_order, float *e_spec, float *e_diff) {
    Scene3D sc;
    if (!parse_obj3d(obj, &sc, h)) return -1;
    const RTScene S = rt_build(sc);
    std::vector<float> sa(src, src + 3 * ns), sp(src_p, src_p + ns), ra(rec, rec + 3 * nr), es, ed;
    rt_trace_late(S, sa, sp, ra, alpha_gb, m_air, sctr_g, c, fs, Nt, n_rays, max_order, radius, es, ed, ism_order);
    std::copy(es.begin(), es.end(), e_spec);
    std::copy(ed.begin(), ed.end(), e_diff);
    return 0;
}

// rt_dump_paths of the scene (main3D.cu:310-328): the --dump-rays file for all sources.
int ref_dump_rays(const char *obj, float h, const float *src, int ns, int n_rays, int max_order, const char *file) {
    Scene3D sc;
    if (!parse_obj3d(obj, &sc, h)) return -1;
    const RTScene S = rt_build(sc);
    std::vector<vec3> srcs;
    for (int q = 0; q < ns; ++q) srcs.push_back(v3(src[3 * q], src[3 * q + 1], src[3 * q + 2]));
    if (srcs.empty()) srcs.push_back(v3(0, 0, 0));
    rt_dump_paths(S, srcs, n_rays, max_order, file, &sc.group_names);
    return 0;
}

// The reference's oblique-incidence reflection factor (parsing3D.h:1178-1186) and the
// ray tracer's hash sign (raytrace3D.h:69-72), for the IR-synthesis restatement's pins.
float ref_r_oblique(float alpha, float cos_theta) { return R_oblique(alpha, cos_theta); }
float ref_rt_sign(unsigned s) { return rt_sign(s); }

// read_fit_file into float arrays (exits the process on a malformed file, as the reference does).
void ref_read_fit(const char *path, int m, int K, int L, float *Rinf, float *A, float *zeta, float *B, float *C,
                  float *alpha, float *beta) {
    read_fit_file(path, m, K, L, Rinf, A, zeta, B, C, alpha, beta);
}

}
