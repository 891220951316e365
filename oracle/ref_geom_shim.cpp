// ORACLE -- test infrastructure only. Exposes the reference's own CPU geometry path
// (3D-solver/include/parsing3D.h parse_obj3d / build_sdf / build_oid_dense, and
// common/filterABC.h read_fit_file / build_boundary_map) through a C ABI so tests can
// check the product's restatement bit for bit. Built only by oracle/build_ref.sh from
// the sources where they lie under the reference checkout, into oracle/_ref/.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include "parsing3D.h"
#include "filterABC.h"

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

int ref_boundary_map(uint8_t *oid, int N, int *bmap) { return build_boundary_map(oid, N, bmap); }

// read_fit_file into float arrays (exits the process on a malformed file, as the reference does).
void ref_read_fit(const char *path, int m, int K, int L, float *Rinf, float *A, float *zeta, float *B, float *C,
                  float *alpha, float *beta) {
    read_fit_file(path, m, K, L, Rinf, A, zeta, B, C, alpha, beta);
}

}
