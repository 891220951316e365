// Geometrical acoustics (image sources + ray tracer): data shared by the host part
// (host_ga.cpp, g++) and the HIP kernels (fvm_ga.hip). Plain structs of floats and ints
// so one layout serves both compilers.
//
// Follows 3D-solver/include/ism3D.h (ism_build_reflectors :268-393, ism_build_images
// :398-418, ism_validate_path :424-503, ism_paths_for_pairs :521-621, ism_solve :665-723)
// and 3D-solver/include/raytrace3D.h (rt_build :229-261, rt_closest_hit :276-299,
// rt_fib_ray :141-149, rt_trace_late :335-568). Every float expression keeps the
// reference's operand order and both files are compiled without FMA contraction, so the
// GPU's decisions (path valid or not, which triangle a ray hits) are the host's.
#pragma once

#include "host.hpp"

#include <cstdint>
#include <string>
#include <vector>

namespace fases {
namespace ga {

constexpr int kBands = 8;            // octave bands 63 Hz .. 8 kHz
constexpr int kMaxIsmOrder = 32;     // image-tree depth the validity kernel supports
constexpr int kBvhStack = 64;        // traversal stack of the ray tracer (the reference's)

// One reflector of the image tree: a merged plane of coplanar facets, or a proxy patch of
// a dense facet cluster (with a curvature radius when a sphere fits it).
struct Reflector {
    F3 n{0, 0, 1};
    float d = 0.f;
    float area = 0.f;
    bool curved = false;
    float radius = 0.f;              // valid when curved; flat reflectors carry FLT_MAX
    int convex = 1;                  // +1 convex, -1 concave
    float hit_band = 1e-4f;          // in-facet tolerance
    std::vector<int> tris;           // mesh triangles of the reflector
};

// Tuning of the image-source method (the reference's ISMConfig defaults).
struct IsmCfg {
    int max_order = 3;
    float coplanar_cos = 0.9998f;
    float coplanar_dist = 1e-3f;
    float min_plane_area = 1e-4f;
    int max_planes_per_group = 12;
    float facet_tol = 1e-4f;
    float seg_eps = 1e-5f;
    float proxy_normal_cos = 0.906f;
    float min_proxy_area = 1e-4f;
    int proxy_max_per_group = 512;
    float flat_radius_factor = 50.f;
    int max_order_curved = 1;
    float curved_amp_cap = 4.0f;
    int convex_unknown_as_convex = 1;
};

// Image-tree node: image position, the reflector that made it, parent node.
struct Node {
    F3 image;
    int plane, parent, order, curved_count;
};

// Geometry of the validity kernel in flat arrays (device copies made by fvm_ga.hip).
struct IsmGeom {
    std::vector<F3> ta, tb, tc;              // placed triangles
    std::vector<int> tri2plane;              // reflector of each triangle (-1: occluder only)
    std::vector<F3> pn;                      // reflector normals
    std::vector<float> pd, pband;            // plane offsets, hit bands
    std::vector<int> poff, ptri;             // reflector -> triangles (CSR)
};

// Flat BVH node (raytrace3D.h BVHNode): leaf when count > 0.
struct BvhNode {
    F3 bmin, bmax;
    int left, right, first, count;
};

// Ray-tracer scene: triangles in BVH-leaf order, their unit normals and groups.
struct RtScene {
    std::vector<F3> a, b, c, nrm;
    std::vector<int> group;
    std::vector<BvhNode> nodes;
    int ngroups = 0;
};

// Per-launch constants of the ray tracer.
struct RtParams {
    int n_src, n_rec, n_rays, max_order, Nt, ism_order;
    int s_base;                    // index of this launch's first source among all sources
    float c, fs, recv_radius;
    float Kdiff[64], Kspec[64];    // per source (at most 64 sources per launch)
};

// Reflection factor of one (group, band) with the angle-independent part precomputed:
// kind 0 = anechoic (0), 1 = near-rigid (r0), 2 = locally reacting (xi), as R_oblique
// (parsing3D.h:1178-1186) evaluates it.
struct Refl {
    int kind;
    float r0, xi;
};

// GPU parts (fvm_ga.hip). Return 0 or a FASES_E_* code with err set.
int ism_valid_gpu(int device, const IsmGeom &g, const std::vector<Node> &nodes, const F3 &src,
                  const std::vector<F3> &recs, float seg_eps, std::vector<uint8_t> &valid, std::string &err);
int rt_trace_gpu(int device, const RtScene &S, const RtParams &P, const std::vector<F3> &src,
                 const std::vector<F3> &rec, const std::vector<F3> &dirs, const std::vector<Refl> &refl,
                 const std::vector<float> &sctr, const std::vector<float> &m_air, float *e_spec, float *e_diff,
                 double *ms, std::string &err);

}  // namespace ga
}  // namespace fases
