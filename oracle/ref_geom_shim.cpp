// ORACLE -- test infrastructure only. Exposes the reference's own CPU geometry path
// (3D-solver/include/parsing3D.h parse_obj3d / build_sdf / build_oid_dense /
// mesh_group_areas, common/filterABC.h read_fit_file / build_boundary_map, and the
// reverberation estimate of 3D-solver/include/ism3D.h ism_load_impedances_layered /
// estimate_T60s) through a C ABI so tests can
// check the product's restatement bit for bit. Built only by oracle/build_ref.sh from
// the sources where they lie under the reference checkout, into oracle/_ref/.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include "parsing3D.h"
#include "filterABC.h"
#include "ism3D.h"

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

// read_fit_file into float arrays (exits the process on a malformed file, as the reference does).
void ref_read_fit(const char *path, int m, int K, int L, float *Rinf, float *A, float *zeta, float *B, float *C,
                  float *alpha, float *beta) {
    read_fit_file(path, m, K, L, Rinf, A, zeta, B, C, alpha, beta);
}

}
