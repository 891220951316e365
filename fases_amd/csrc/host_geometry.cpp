// Scene classification on the host: OBJ mesh -> padded grid -> narrow-band signed
// distance S -> dense material id (oid) -> compact boundary map.
//
// Follows 3D-solver/include/parsing3D.h (load_obj_mesh3d :169-288, parse_obj_3d :672-688,
// compute_sdf_cpu :389-452, build_sdf :860-910, grid_dims_checked :792-856,
// build_oid_dense_cpu :955-1015) and common/filterABC.h (build_boundary_map :95-108).
// Classification must be bit-identical to the reference, so every float expression is
// evaluated in float32 with the reference's operand order (this file is compiled with
// -ffp-contract=off, like the reference's default g++ host build which has no FMA).
// The distance routine is Ericson's closest-point-on-triangle (Real-Time Collision
// Detection, 2005, sec. 5.1.5); the crossing test is Moller-Trumbore (1997).

#include "host.hpp"
#include "../../include/fases_core.h"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <unordered_map>

namespace fases {

namespace {

thread_local std::string g_err;

inline F3 mk(float x, float y, float z) { F3 r; r.x = x; r.y = y; r.z = z; return r; }
inline F3 vsub(const F3 &a, const F3 &b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline F3 vadd(const F3 &a, const F3 &b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline F3 vscale(const F3 &a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline float vdot(const F3 &a, const F3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 vcross(const F3 &a, const F3 &b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Squared distance from p to triangle (a,b,c): Voronoi-region walk (Ericson 5.1.5).
float tri_dist2(const F3 &p, const F3 &a, const F3 &b, const F3 &c) {
    const F3 ab = vsub(b, a), ac = vsub(c, a), ap = vsub(p, a);
    const float d1 = vdot(ab, ap), d2 = vdot(ac, ap);
    if (d1 <= 0.f && d2 <= 0.f) return vdot(ap, ap);                    // vertex a
    const F3 bp = vsub(p, b);
    const float d3 = vdot(ab, bp), d4 = vdot(ac, bp);
    if (d3 >= 0.f && d4 <= d3) return vdot(bp, bp);                      // vertex b
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {                           // edge ab
        const float t = d1 / (d1 - d3);
        const F3 q = vsub(p, vadd(a, vscale(ab, t)));
        return vdot(q, q);
    }
    const F3 cp = vsub(p, c);
    const float d5 = vdot(ab, cp), d6 = vdot(ac, cp);
    if (d6 >= 0.f && d5 <= d6) return vdot(cp, cp);                      // vertex c
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) {                           // edge ac
        const float t = d2 / (d2 - d6);
        const F3 q = vsub(p, vadd(a, vscale(ac, t)));
        return vdot(q, q);
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {            // edge bc
        const F3 bc = vsub(c, b);
        const float t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        const F3 q = vsub(p, vadd(b, vscale(bc, t)));
        return vdot(q, q);
    }
    const float inv = 1.f / (va + vb + vc);                              // face interior
    const float v = vb * inv, w = vc * inv;
    const F3 q = vsub(p, vadd(a, vadd(vscale(ab, v), vscale(ac, w))));
    return vdot(q, q);
}

// Ray/triangle crossing (Moller-Trumbore), edge-inclusive; t > tmin accepted.
bool ray_hits(const F3 &o, const F3 &d, const F3 &a, const F3 &b, const F3 &c, float tmin, float &t_out) {
    const float eps = 1e-8f;
    const F3 e1 = vsub(b, a), e2 = vsub(c, a);
    const F3 pv = vcross(d, e2);
    const float det = vdot(e1, pv);
    if (det > -eps && det < eps) return false;
    const float inv = 1.f / det;
    const F3 tv = vsub(o, a);
    const float u = vdot(tv, pv) * inv;
    if (u < 0.f || u > 1.f) return false;
    const F3 qv = vcross(tv, e1);
    const float v = vdot(d, qv) * inv;
    if (v < 0.f || u + v > 1.f) return false;
    const float t = vdot(e2, qv) * inv;
    if (t > tmin) { t_out = t; return true; }
    return false;
}

inline F3 placed(const Mesh &m, int i) {
    const F3 &v = m.verts[(size_t)i];
    // (v + shift) * scale with scale == 1 (parsing3D.h:89-92); the *1 is exact
    return mk((v.x + m.offset.x) * 1.f, (v.y + m.offset.y) * 1.f, (v.z + m.offset.z) * 1.f);
}

void strip(std::string &s) {
    auto ws = [](unsigned char ch) { return ch == ' ' || ch == '\t' || ch == '\r' || ch == '\n'; };
    size_t b = 0, e = s.size();
    while (b < e && ws((unsigned char)s[b])) ++b;
    while (e > b && ws((unsigned char)s[e - 1])) --e;
    s = s.substr(b, e - b);
}

// Group label of an "o"/"g" line: second word after the tag if present, else the first.
std::string group_label(const std::string &line) {
    std::istringstream ss(line);
    std::string tag, w;
    ss >> tag;
    std::vector<std::string> words;
    while (ss >> w) words.push_back(w);
    if (words.size() >= 2) return words[1];
    if (words.size() == 1) return words[0];
    return "unnamed";
}

}  // namespace

void set_global_error(const std::string &msg) { g_err = msg; }
const char *global_error() { return g_err.c_str(); }

bool mesh_from_obj(const std::string &path, Mesh &m, std::string &err) {
    std::ifstream in(path);
    if (!in) { err = "cannot open OBJ: " + path; return false; }
    m = Mesh();
    std::unordered_map<std::string, int> by_name;
    int cur = -1;
    std::string cur_name = "unnamed";
    bool named = false;
    std::string line;
    while (std::getline(in, line)) {
        strip(line);
        if (line.empty() || line[0] == '#') continue;
        if (line.rfind("v ", 0) == 0) {
            std::istringstream ss(line);
            char tag; float x, y, z;
            ss >> tag >> x >> y >> z;
            if (!ss.fail()) m.verts.push_back(mk(x, y, z));
        } else if (line.rfind("o ", 0) == 0 || line.rfind("g ", 0) == 0) {
            const std::string nm = group_label(line);
            if (!named || nm != cur_name) { cur_name = nm; named = true; cur = -1; }
        } else if (line.rfind("f ", 0) == 0) {
            std::istringstream ss(line);
            char tag; ss >> tag;
            std::vector<int> idx;
            std::string tok;
            const int nv = (int)m.verts.size();
            while (ss >> tok) {
                const size_t sl = tok.find('/');
                const std::string head = (sl == std::string::npos) ? tok : tok.substr(0, sl);
                if (head.empty()) { err = "invalid face index token: '" + tok + "'"; return false; }
                int vi = std::atoi(head.c_str());
                if (vi > 0) vi -= 1;
                else if (vi < 0) vi += nv;
                else { err = "invalid face index token: '" + tok + "'"; return false; }
                if (vi < 0 || vi >= nv) { err = "invalid face index token: '" + tok + "'"; return false; }
                idx.push_back(vi);
            }
            if (idx.size() < 3) continue;
            if (cur < 0) {   // groups are created lazily at their first face and reused by name
                const std::string nm = named ? cur_name : std::string("unnamed");
                auto it = by_name.find(nm);
                if (it != by_name.end()) cur = it->second;
                else { cur = (int)m.groups.size(); m.groups.push_back(nm); by_name[nm] = cur; }
            }
            for (size_t q = 1; q + 1 < idx.size(); ++q) m.tris.push_back({idx[0], idx[q], idx[q + 1], cur});
        }
    }
    if (m.verts.empty() || m.tris.empty()) { err = "OBJ contained no vertices or no faces."; return false; }
    m.lo = m.hi = m.verts[0];
    for (const F3 &v : m.verts) {
        m.lo.x = std::min(m.lo.x, v.x); m.lo.y = std::min(m.lo.y, v.y); m.lo.z = std::min(m.lo.z, v.z);
        m.hi.x = std::max(m.hi.x, v.x); m.hi.y = std::max(m.hi.y, v.y); m.hi.z = std::max(m.hi.z, v.z);
    }
    return true;
}

void mesh_shoebox(int nx, int ny, int nz, float h, Mesh &m) {
    // Same vertex numbering and quad winding as a SketchUp-style shoebox OBJ
    // (3D-solver/tests/Room_Modes_benchmark/SimpleShoeBoxRoom.obj), fan-triangulated.
    m = Mesh();
    const float X = (float)nx * h, Y = (float)ny * h, Z = (float)nz * h;
    m.verts = {mk(0, 0, 0), mk(X, 0, 0), mk(X, Y, 0), mk(0, Y, 0),
               mk(0, 0, Z), mk(X, 0, Z), mk(X, Y, Z), mk(0, Y, Z)};
    const int quads[6][4] = {{3, 7, 4, 0}, {1, 5, 6, 2}, {0, 4, 5, 1},
                             {2, 6, 7, 3}, {0, 1, 2, 3}, {4, 7, 6, 5}};
    const char *names[6] = {"west", "east", "south", "north", "floor", "ceiling"};
    for (int g = 0; g < 6; ++g) {
        m.groups.push_back(names[g]);
        m.tris.push_back({quads[g][0], quads[g][1], quads[g][2], g});
        m.tris.push_back({quads[g][0], quads[g][2], quads[g][3], g});
    }
    m.lo = mk(0, 0, 0);
    m.hi = mk(X, Y, Z);
}

bool scene_build(Scene &sc, std::string &err, int device, int *code) {
    Mesh &m = sc.mesh;
    const float h = sc.h;
    if (code) *code = FASES_E_ARG;
    if (!(h > 0.f)) { err = "grid spacing h must be > 0"; return false; }
    const float pad = 3 * h;   // PARSE3D_GHOST cells of ghost padding per side
    m.offset = mk(-m.lo.x + pad, -m.lo.y + pad, -m.lo.z + pad);

    F3 bmax = placed(m, 0);
    for (size_t i = 1; i < m.verts.size(); ++i) {
        const F3 v = placed(m, (int)i);
        bmax.x = std::max(bmax.x, v.x); bmax.y = std::max(bmax.y, v.y); bmax.z = std::max(bmax.z, v.z);
    }
    sc.Lx = (bmax.x + pad) - 0.f;
    sc.Ly = (bmax.y + pad) - 0.f;
    sc.Lz = (bmax.z + pad) - 0.f;
    const double nxd = std::ceil((double)sc.Lx / (double)h);
    const double nyd = std::ceil((double)sc.Ly / (double)h);
    const double nzd = std::ceil((double)sc.Lz / (double)h);
    if (nxd < 2.0 || nyd < 2.0 || nzd < 2.0) { err = "degenerate grid (domain too small for this h)"; return false; }
    if (nxd * nyd * nzd > 4.0e9) { err = "grid too large for this h"; return false; }
    const int Nx = (int)nxd, Ny = (int)nyd, Nz = (int)nzd;
    sc.Nx = Nx; sc.Ny = Ny; sc.Nz = Nz;
    const size_t N = (size_t)Nx * Ny * Nz;
    const size_t nxy = (size_t)Nx * Ny;
    if (sc.adopt_S) {                 // classification supplied (scene cache)
        if (sc.adopt_dims[0] != Nx || sc.adopt_dims[1] != Ny || sc.adopt_dims[2] != Nz) {
            err = "supplied classification is for a " + std::to_string(sc.adopt_dims[0]) + "x" +
                  std::to_string(sc.adopt_dims[1]) + "x" + std::to_string(sc.adopt_dims[2]) + " grid, the scene gives " +
                  std::to_string(Nx) + "x" + std::to_string(Ny) + "x" + std::to_string(Nz);
            return false;
        }
        sc.S.assign(sc.adopt_S, sc.adopt_S + N);
        sc.oid.assign(sc.adopt_oid, sc.adopt_oid + N);
        sc.adopt_S = nullptr;
        sc.adopt_oid = nullptr;
        return true;
    }

    // Placed triangle corners, once.
    const size_t nt = m.tris.size();
    std::vector<F3> ta(nt), tb(nt), tc(nt);
    for (size_t t = 0; t < nt; ++t) {
        ta[t] = placed(m, m.tris[t].a); tb[t] = placed(m, m.tris[t].b); tc[t] = placed(m, m.tris[t].c);
    }

    // Material id of each mesh group: 1-based over the groups that have faces, in group
    // order; at most 254 ids (parsing3D.h:966-977).
    int max_gid = -1;
    for (const Triangle &t : m.tris) max_gid = std::max(max_gid, t.group);
    std::vector<int> lid_of((size_t)std::max(max_gid + 1, 0), -1);
    {
        int next = 0;
        for (int g = 0; g <= max_gid; ++g) {
            bool used = false;
            for (const Triangle &t : m.tris) if (t.group == g) { used = true; break; }
            if (!used) continue;
            if (next <= 254) lid_of[(size_t)g] = 1 + next++;
        }
    }

    // 1) |distance| clamped at the band radius 8h, min over triangles whose expanded
    //    bounding box covers the cell centre. Min is order-free, so z-slabs run in parallel.
    const float band = 8.f * h;
    std::vector<float> absd(N, band);
    std::vector<int> bi0(nt), bi1(nt), bj0(nt), bj1(nt), bk0(nt), bk1(nt);
    for (size_t t = 0; t < nt; ++t) {
        const F3 &a = ta[t], &b = tb[t], &c = tc[t];
        const float lox = std::min(a.x, std::min(b.x, c.x)) - band, hix = std::max(a.x, std::max(b.x, c.x)) + band;
        const float loy = std::min(a.y, std::min(b.y, c.y)) - band, hiy = std::max(a.y, std::max(b.y, c.y)) + band;
        const float loz = std::min(a.z, std::min(b.z, c.z)) - band, hiz = std::max(a.z, std::max(b.z, c.z)) + band;
        int i0 = (int)std::floor((lox - 0.f) / h - 0.5f), i1 = (int)std::ceil((hix - 0.f) / h - 0.5f);
        int j0 = (int)std::floor((loy - 0.f) / h - 0.5f), j1 = (int)std::ceil((hiy - 0.f) / h - 0.5f);
        int k0 = (int)std::floor((loz - 0.f) / h - 0.5f), k1 = (int)std::ceil((hiz - 0.f) / h - 0.5f);
        bi0[t] = std::max(i0, 0); bj0[t] = std::max(j0, 0); bk0[t] = std::max(k0, 0);
        bi1[t] = std::min(i1, Nx - 1); bj1[t] = std::min(j1, Ny - 1); bk1[t] = std::min(k1, Nz - 1);
    }
    if (device >= 0) {                     // steps 1-3 on the GPU
        std::vector<int> boxes(6 * nt), gid(nt);
        for (size_t t = 0; t < nt; ++t) {
            const int b6[6] = {bi0[t], bi1[t], bj0[t], bj1[t], bk0[t], bk1[t]};
            std::copy(b6, b6 + 6, boxes.begin() + 6 * t);
            gid[t] = m.tris[t].group;
        }
        sc.S.assign(N, 0.f);
        sc.oid.assign(N, (uint8_t)0xFF);
        const int rc = classify_gpu(device, Nx, Ny, Nz, h, ta, tb, tc, boxes, gid, lid_of, max_gid, sc.S.data(),
                                    sc.oid.data(), err);
        if (rc) { if (code) *code = rc; return false; }
        return true;
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < Nz; ++k) {
        const float pz = 0.f + ((float)k + 0.5f) * h;
        for (size_t t = 0; t < nt; ++t) {
            if (k < bk0[t] || k > bk1[t]) continue;
            for (int j = bj0[t]; j <= bj1[t]; ++j) {
                const float py = 0.f + ((float)j + 0.5f) * h;
                for (int i = bi0[t]; i <= bi1[t]; ++i) {
                    const F3 p = mk(0.f + ((float)i + 0.5f) * h, py, pz);
                    const float d = std::sqrt(tri_dist2(p, ta[t], tb[t], tc[t]));
                    float &slot = absd[(size_t)k * nxy + (size_t)j * Nx + i];
                    if (d < slot) slot = d;
                }
            }
        }
    }

    // 2) Sign: one +X ray per (j,k) row, nudged off mesh edges by (0.0137h, 0.0091h);
    //    an odd number of crossings ahead of the cell centre means inside (S > 0).
    sc.S.assign(N, 0.f);
    const F3 dir = mk(1.f, 0.f, 0.f);
    const float tmin = 1e-6f;
#pragma omp parallel
    {
        std::vector<float> xs;
        xs.reserve(64);
#pragma omp for schedule(static) collapse(2)
        for (int k = 0; k < Nz; ++k) {
            for (int j = 0; j < Ny; ++j) {
                const float y0 = 0.f + ((float)j + 0.5f) * h + 0.0137f * h;
                const float z0 = 0.f + ((float)k + 0.5f) * h + 0.0091f * h;
                const F3 o = mk(0.f, y0, z0);
                xs.clear();
                for (size_t t = 0; t < nt; ++t) {
                    float tt;
                    if (ray_hits(o, dir, ta[t], tb[t], tc[t], 0.0f, tt)) xs.push_back(o.x + tt);
                }
                std::sort(xs.begin(), xs.end());
                for (int i = 0; i < Nx; ++i) {
                    const float xc = 0.f + ((float)i + 0.5f) * h;
                    const long ahead = (long)(xs.end() - std::upper_bound(xs.begin(), xs.end(), xc + tmin));
                    const size_t id = (size_t)k * nxy + (size_t)j * Nx + i;
                    const float s = absd[id];
                    sc.S[id] = (ahead & 1) ? s : -s;
                }
            }
        }
    }

    // 3) OID: for band cells 0 <= s <= tol (tol = 3h), foot point W = C - s*n with
    //    n = grad S / |grad S| (central differences), then the nearest triangle
    //    (strictly smaller distance wins, so the first triangle keeps ties).
    const float tol = 3.0f * h;
    sc.oid.assign(N, (uint8_t)0xFF);
    if (max_gid < 0) return true;
    const float two_h = 2.0f * h;
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 1; k < Nz - 1; ++k) {
        for (int j = 1; j < Ny - 1; ++j) {
            for (int i = 1; i < Nx - 1; ++i) {
                const size_t id = (size_t)k * nxy + (size_t)j * Nx + i;
                const float s = sc.S[id];
                if (!(s >= 0.0f && std::fabs(s) <= tol)) continue;
                const float gx = (sc.S[id + 1] - sc.S[id - 1]) / two_h;
                const float gy = (sc.S[id + Nx] - sc.S[id - Nx]) / two_h;
                const float gz = (sc.S[id + nxy] - sc.S[id - nxy]) / two_h;
                const float g2 = gx * gx + gy * gy + gz * gz;
                if (g2 <= 1e-24f) continue;
                const float ig = 1.0f / std::sqrt(g2);
                const F3 n = mk(gx * ig, gy * ig, gz * ig);
                const F3 C = mk(0.f + ((float)i + 0.5f) * h, 0.f + ((float)j + 0.5f) * h, 0.f + ((float)k + 0.5f) * h);
                const F3 W = mk(C.x - s * n.x, C.y - s * n.y, C.z - s * n.z);
                float best = FLT_MAX;
                int gbest = -1;
                for (size_t t = 0; t < nt; ++t) {
                    const float d2 = tri_dist2(W, ta[t], tb[t], tc[t]);
                    if (d2 < best) { best = d2; gbest = m.tris[t].group; }
                }
                if (gbest < 0 || gbest > max_gid) continue;
                const int lid = lid_of[(size_t)gbest];
                if (lid >= 0 && lid <= 254) sc.oid[id] = (uint8_t)lid;
            }
        }
    }
    return true;
}

}  // namespace fases

using namespace fases;

extern "C" {

int fases_scene_load_obj(const char *path, float h, fases_scene **out) {
    return fases_scene_load_obj_dev(path, h, -1, out);
}

int fases_scene_shoebox(int nx, int ny, int nz, float h, fases_scene **out) {
    return fases_scene_shoebox_dev(nx, ny, nz, h, -1, out);
}

int fases_scene_load_obj_dev(const char *path, float h, int device, fases_scene **out) {
    if (!path || !out) { set_global_error("fases_scene_load_obj: null argument"); return FASES_E_ARG; }
    *out = nullptr;
    Scene *sc = new (std::nothrow) Scene();
    if (!sc) return FASES_E_NOMEM;
    std::string err;
    if (!mesh_from_obj(path, sc->mesh, err)) {
        set_global_error(err); delete sc;
        return (err.rfind("cannot open", 0) == 0) ? FASES_E_IO : FASES_E_PARSE;
    }
    sc->h = h;
    try {
        int code = FASES_E_ARG;
        if (!scene_build(*sc, err, device, &code)) { set_global_error(err); delete sc; return code; }
    } catch (const std::bad_alloc &) {
        set_global_error("out of host memory building the scene"); delete sc; return FASES_E_NOMEM;
    }
    *out = reinterpret_cast<fases_scene *>(sc);
    return FASES_OK;
}

int fases_scene_load_obj_classified(const char *path, float h, const float *S, const uint8_t *oid, int Nx, int Ny,
                                    int Nz, fases_scene **out) {
    if (!path || !out || !S || !oid || Nx < 1 || Ny < 1 || Nz < 1) {
        set_global_error("fases_scene_load_obj_classified: bad argument");
        return FASES_E_ARG;
    }
    *out = nullptr;
    Scene *sc = new (std::nothrow) Scene();
    if (!sc) return FASES_E_NOMEM;
    std::string err;
    if (!mesh_from_obj(path, sc->mesh, err)) {
        set_global_error(err); delete sc;
        return (err.rfind("cannot open", 0) == 0) ? FASES_E_IO : FASES_E_PARSE;
    }
    sc->h = h;
    sc->adopt_S = S;
    sc->adopt_oid = oid;
    sc->adopt_dims[0] = Nx; sc->adopt_dims[1] = Ny; sc->adopt_dims[2] = Nz;
    try {
        int code = FASES_E_ARG;
        if (!scene_build(*sc, err, -1, &code)) { set_global_error(err); delete sc; return code; }
    } catch (const std::bad_alloc &) {
        set_global_error("out of host memory building the scene"); delete sc; return FASES_E_NOMEM;
    }
    *out = reinterpret_cast<fases_scene *>(sc);
    return FASES_OK;
}

int fases_scene_shoebox_dev(int nx, int ny, int nz, float h, int device, fases_scene **out) {
    if (!out || nx < 1 || ny < 1 || nz < 1) { set_global_error("fases_scene_shoebox: bad argument"); return FASES_E_ARG; }
    *out = nullptr;
    Scene *sc = new (std::nothrow) Scene();
    if (!sc) return FASES_E_NOMEM;
    mesh_shoebox(nx, ny, nz, h, sc->mesh);
    sc->h = h;
    std::string err;
    try {
        int code = FASES_E_ARG;
        if (!scene_build(*sc, err, device, &code)) { set_global_error(err); delete sc; return code; }
    } catch (const std::bad_alloc &) {
        set_global_error("out of host memory building the scene"); delete sc; return FASES_E_NOMEM;
    }
    *out = reinterpret_cast<fases_scene *>(sc);
    return FASES_OK;
}

int fases_scene_info(const fases_scene *s, int *Nx, int *Ny, int *Nz, int *n_groups, float *L3) {
    if (!s) return FASES_E_ARG;
    const Scene *sc = reinterpret_cast<const Scene *>(s);
    if (Nx) *Nx = sc->Nx;
    if (Ny) *Ny = sc->Ny;
    if (Nz) *Nz = sc->Nz;
    if (n_groups) *n_groups = (int)sc->mesh.groups.size();
    if (L3) { L3[0] = sc->Lx; L3[1] = sc->Ly; L3[2] = sc->Lz; }
    return FASES_OK;
}

int fases_scene_group_name(const fases_scene *s, int g, char *buf, int buflen) {
    if (!s || !buf || buflen < 1) return FASES_E_ARG;
    const Scene *sc = reinterpret_cast<const Scene *>(s);
    if (g < 0 || g >= (int)sc->mesh.groups.size()) return FASES_E_ARG;
    std::snprintf(buf, (size_t)buflen, "%s", sc->mesh.groups[(size_t)g].c_str());
    return FASES_OK;
}

int fases_scene_get(const fases_scene *s, float *S, uint8_t *oid) {
    if (!s) return FASES_E_ARG;
    const Scene *sc = reinterpret_cast<const Scene *>(s);
    if (S) std::memcpy(S, sc->S.data(), sc->S.size() * sizeof(float));
    if (oid) std::memcpy(oid, sc->oid.data(), sc->oid.size());
    return FASES_OK;
}

int fases_scene_group_areas(const fases_scene *s, float *A, float *A_tot) {
    if (!s) return FASES_E_ARG;
    const Scene *sc = reinterpret_cast<const Scene *>(s);
    const Mesh &m = sc->mesh;
    std::vector<float> acc(m.groups.size(), 0.f);
    float tot = 0.f;
    // half the cross-product norm of each placed triangle, summed in triangle order per
    // group and in total (parsing3D.h mesh_group_areas:101-114)
    for (const Triangle &t : m.tris) {
        const F3 a = placed(m, t.a), b = placed(m, t.b), c = placed(m, t.c);
        const F3 n = vcross(vsub(b, a), vsub(c, a));
        const float ar = 0.5f * std::sqrt(vdot(n, n));
        if (t.group >= 0 && t.group < (int)acc.size()) acc[(size_t)t.group] += ar;
        tot += ar;
    }
    if (A) std::copy(acc.begin(), acc.end(), A);
    if (A_tot) *A_tot = tot;
    return FASES_OK;
}

void fases_scene_free(fases_scene *s) { delete reinterpret_cast<Scene *>(s); }

int fases_boundary_map(const uint8_t *oid, int64_t N, int32_t *bmap, int64_t *nB) {
    if (!oid || !bmap || N < 0) return FASES_E_ARG;
    int64_t n = 0;
    for (int64_t c = 0; c < N; ++c) bmap[c] = (oid[c] != 0xFF && oid[c] != 0) ? (int32_t)(n++) : -1;
    if (nB) *nB = n;
    return FASES_OK;
}

}  // extern "C"
