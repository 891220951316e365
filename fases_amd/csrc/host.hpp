// Internal host-side declarations shared by the C++ translation units of libfases_core.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace fases {

// Thread-local message of the last failing call made without a context.
void set_global_error(const std::string &msg);
const char *global_error();

// Single-precision 3-vector. The geometry pipeline is evaluated in float32 with the
// operand order of 3D-solver/include/parsing3D.h so that S and oid come out bit-identical.
struct F3 {
    float x, y, z;
};

struct Triangle {
    int a, b, c;   // vertex indices
    int group;     // mesh group (material) index
};

struct Mesh {
    std::vector<F3> verts;
    std::vector<Triangle> tris;
    std::vector<std::string> groups;   // group id == index
    F3 lo{0, 0, 0}, hi{0, 0, 0};       // raw bounding box
    F3 offset{0, 0, 0};                // placement shift (+3h padding)
};

struct Scene {
    Mesh mesh;
    float h = 0.f;
    int Nx = 0, Ny = 0, Nz = 0;
    float Lx = 0.f, Ly = 0.f, Lz = 0.f;   // padded extent
    std::vector<float> S;
    std::vector<uint8_t> oid;
    // a classification supplied by the caller (a cached S / oid): scene_build places the
    // mesh, checks the grid and adopts these instead of classifying
    const float *adopt_S = nullptr;
    const uint8_t *adopt_oid = nullptr;
    int adopt_dims[3] = {0, 0, 0};
};

bool mesh_from_obj(const std::string &path, Mesh &m, std::string &err);
void mesh_shoebox(int nx, int ny, int nz, float h, Mesh &m);
// Placement, grid dims, SDF, OID. device < 0: on the host; otherwise steps SDF and OID
// run on that HIP device (same results bit for bit). On failure *code (if given) is
// the FASES_E_* code.
bool scene_build(Scene &sc, std::string &err, int device = -1, int *code = nullptr);

// Device part of scene_build (fvm_geometry.hip): triangle corners, band boxes
// (i0, i1, j0, j1, k0, k1 per triangle), group per triangle and the group -> material
// id table in; S and oid (Nx Ny Nz each) out.
int classify_gpu(int device, int Nx, int Ny, int Nz, float h, const std::vector<F3> &ta, const std::vector<F3> &tb,
                 const std::vector<F3> &tc, const std::vector<int> &boxes, const std::vector<int> &gid,
                 const std::vector<int> &lid_of, int max_gid, float *S_out, uint8_t *oid_out, std::string &err);

}  // namespace fases
