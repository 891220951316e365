// Geometrical acoustics on the host side: reflectors and image tree of the image-source
// method, path details, the ray tracer's BVH, the band impulse responses and the arrival
// lists of the hybrid IR, and the C ABI (fases_ga_*). The GPU validates every image path
// (occlusion tests, fvm_ga.hip) and traces the rays.
//
// Follows 3D-solver/include/ism3D.h (ism_build_reflectors :268-393, ism_build_images
// :398-418, ism_validate_path :424-503, ism_paths_for_pairs :521-621, ism_solve :665-723),
// 3D-solver/include/raytrace3D.h (rt_build :229-261, rt_fib_ray :141-149) and the GA part
// of 3D-solver/core/main3D.cu (:1479-1805: band IRs, HIR/HRT/ARR). Float expressions keep
// the reference's operand order (no FMA contraction), so image paths, lengths, angles and
// gains are the reference's bit for bit.

#include "ga.hpp"
#include "ga_math.hpp"
#include "../../include/fases_core.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <set>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

namespace fases {
namespace ga {
namespace {

// placed vertex: (v + shift) * scale with scale 1
inline F3 vtx(const Mesh &m, int i) {
    const F3 &v = m.verts[(size_t)i];
    return f3((v.x + m.offset.x) * 1.f, (v.y + m.offset.y) * 1.f, (v.z + m.offset.z) * 1.f);
}

F3 tri_normal_area(const Mesh &m, const Triangle &t, float &area) {
    const F3 a = vtx(m, t.a), b = vtx(m, t.b), c = vtx(m, t.c);
    F3 nc = cross(sub(b, a), sub(c, a));
    const float len = sqrtf(len2(nc));
    area = 0.5f * len;
    if (len > 1e-20f) nc = scl(nc, 1.f / len);
    else nc = f3(0, 0, 1);
    return nc;
}

// sign convention of a plane normal: its largest component positive
F3 canon(F3 n) {
    const float ax = std::fabs(n.x), ay = std::fabs(n.y), az = std::fabs(n.z);
    float s = 1.f;
    if (az >= ax && az >= ay) s = (n.z < 0.f) ? -1.f : 1.f;
    else if (ay >= ax) s = (n.y < 0.f) ? -1.f : 1.f;
    else s = (n.x < 0.f) ? -1.f : 1.f;
    return scl(n, s);
}

// 4x4 linear solve, Gaussian elimination with partial pivoting (false if singular)
bool solve4(double A[4][4], double b[4], double x[4]) {
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        double mx = std::fabs(A[c][c]);
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(A[r][c]) > mx) { mx = std::fabs(A[r][c]); piv = r; }
        if (mx < 1e-18) return false;
        if (piv != c) {
            for (int k = 0; k < 4; ++k) std::swap(A[piv][k], A[c][k]);
            std::swap(b[piv], b[c]);
        }
        for (int r = c + 1; r < 4; ++r) {
            const double f = A[r][c] / A[c][c];
            for (int k = c; k < 4; ++k) A[r][k] -= f * A[c][k];
            b[r] -= f * b[c];
        }
    }
    for (int r = 3; r >= 0; --r) {
        double s = b[r];
        for (int k = r + 1; k < 4; ++k) s -= A[r][k] * x[k];
        x[r] = s / A[r][r];
    }
    return true;
}

// least-squares sphere through points: |p|^2 = 2 c.p + (R^2 - |c|^2)
bool fit_sphere(const std::vector<F3> &P, F3 &c, float &R) {
    if (P.size() < 4) return false;
    double A[4][4] = {{0}}, b[4] = {0};
    for (const F3 &p : P) {
        const double row[4] = {p.x, p.y, p.z, 1.0};
        const double bb = (double)p.x * p.x + (double)p.y * p.y + (double)p.z * p.z;
        for (int i = 0; i < 4; ++i) {
            for (int j = 0; j < 4; ++j) A[i][j] += row[i] * row[j];
            b[i] += row[i] * bb;
        }
    }
    double u[4];
    if (!solve4(A, b, u)) return false;
    c = f3((float)(u[0] * 0.5), (float)(u[1] * 0.5), (float)(u[2] * 0.5));
    const double R2 = u[3] + (double)c.x * c.x + (double)c.y * c.y + (double)c.z * c.z;
    if (R2 <= 0) return false;
    R = (float)std::sqrt(R2);
    return true;
}

}  // namespace

// Reflectors of the image tree. A group with at most max_planes_per_group distinct planes
// contributes its planes (merged later with every coplanar plane of any group); a denser
// group (curved or free-form) is region-grown over shared vertices by normal coherence into
// proxy patches, each flat or, when a sphere fits it, curved with that radius.
void build_reflectors(const Mesh &M, const IsmCfg &cfg, std::vector<Reflector> &planes, std::vector<int> &tri2plane) {
    planes.clear();
    tri2plane.assign(M.tris.size(), -1);
    int max_gid = -1;
    for (const Triangle &t : M.tris) max_gid = std::max(max_gid, t.group);
    if (max_gid < 0) return;
    struct Flat { F3 n; float d; float area; std::vector<int> tris; };
    std::vector<Flat> flats;
    for (int gid = 0; gid <= max_gid; ++gid) {
        std::vector<Flat> local;
        std::vector<int> members_of_group;
        for (int t = 0; t < (int)M.tris.size(); ++t) {
            const Triangle &tr = M.tris[(size_t)t];
            if (tr.group != gid) continue;
            members_of_group.push_back(t);
            float area;
            const F3 n = canon(tri_normal_area(M, tr, area));
            const float d = dot(n, vtx(M, tr.a));
            int hit = -1;
            for (int p = 0; p < (int)local.size(); ++p)
                if (dot(local[(size_t)p].n, n) > cfg.coplanar_cos && std::fabs(local[(size_t)p].d - d) < cfg.coplanar_dist) {
                    hit = p;
                    break;
                }
            if (hit < 0) local.push_back({n, d, area, {t}});
            else { local[(size_t)hit].area += area; local[(size_t)hit].tris.push_back(t); }
        }
        if (members_of_group.empty()) continue;
        if ((int)local.size() <= cfg.max_planes_per_group) {
            for (Flat &L : local)
                if (L.area >= cfg.min_plane_area) flats.push_back(std::move(L));
            continue;
        }
        // dense group: proxy patches grown from each unassigned facet
        const int G = (int)members_of_group.size();
        std::unordered_map<int, std::vector<int>> facets_at;   // vertex -> facets of the group
        facets_at.reserve((size_t)G * 2);
        std::vector<F3> fn((size_t)G), fc((size_t)G);
        std::vector<float> fa((size_t)G);
        for (int li = 0; li < G; ++li) {
            const Triangle &tr = M.tris[(size_t)members_of_group[(size_t)li]];
            float ar;
            fn[(size_t)li] = tri_normal_area(M, tr, ar);
            fa[(size_t)li] = ar;
            const F3 a = vtx(M, tr.a), b = vtx(M, tr.b), c = vtx(M, tr.c);
            fc[(size_t)li] = scl(add(add(a, b), c), 1.f / 3.f);
            facets_at[tr.a].push_back(li);
            facets_at[tr.b].push_back(li);
            facets_at[tr.c].push_back(li);
        }
        std::vector<int> taken((size_t)G, 0);
        int n_proxies = 0;
        for (int seed = 0; seed < G; ++seed) {
            if (taken[(size_t)seed]) continue;
            std::vector<int> members{seed}, stack{seed};
            taken[(size_t)seed] = 1;
            F3 sum_an = scl(fn[(size_t)seed], fa[(size_t)seed]);
            while (!stack.empty()) {
                const int g = stack.back();
                stack.pop_back();
                const Triangle &tg = M.tris[(size_t)members_of_group[(size_t)g]];
                const int vs[3] = {tg.a, tg.b, tg.c};
                const F3 avg = unit(sum_an);
                for (int vi = 0; vi < 3; ++vi) {
                    auto it = facets_at.find(vs[vi]);
                    if (it == facets_at.end()) continue;
                    for (int nb : it->second) {
                        if (taken[(size_t)nb]) continue;
                        if (dot(fn[(size_t)nb], avg) > cfg.proxy_normal_cos) {
                            taken[(size_t)nb] = 1;
                            sum_an = add(sum_an, scl(fn[(size_t)nb], fa[(size_t)nb]));
                            stack.push_back(nb);
                            members.push_back(nb);
                        }
                    }
                }
            }
            const F3 n = unit(sum_an);
            float area = 0.f;
            F3 cw = f3(0, 0, 0);
            std::vector<int> tg;
            std::vector<F3> pts;
            pts.reserve(members.size());
            for (int li : members) {
                area += fa[(size_t)li];
                cw = add(cw, scl(fc[(size_t)li], fa[(size_t)li]));
                tg.push_back(members_of_group[(size_t)li]);
                pts.push_back(fc[(size_t)li]);
            }
            if (area <= 0.f) continue;
            const F3 C = scl(cw, 1.f / area);
            const float d = dot(n, C);
            ++n_proxies;
            if (area < cfg.min_proxy_area) continue;   // a sliver stays an occluder only
            Reflector R;
            R.n = n; R.d = d; R.area = area; R.tris = tg;
            float band = cfg.facet_tol;                // the facets' stand-off from the chord plane
            for (int li : members) {
                const Triangle &tr = M.tris[(size_t)members_of_group[(size_t)li]];
                const int vv[3] = {tr.a, tr.b, tr.c};
                for (int k = 0; k < 3; ++k) band = std::max(band, std::fabs(dot(n, vtx(M, vv[k])) - d));
            }
            R.hit_band = band + cfg.facet_tol;
            F3 sc;
            float sr;
            float extent = 0.f;
            for (const F3 &p : pts) extent = std::max(extent, sqrtf(len2(sub(p, C))));
            if (fit_sphere(pts, sc, sr) && sr < cfg.flat_radius_factor * std::max(extent, 1e-6f)) {
                R.curved = true;
                R.radius = sr;
                double cx = 0.0;
                for (int li : members) cx += (double)dot(fn[(size_t)li], sub(fc[(size_t)li], sc));
                R.convex = (cx > 0.0) ? +1 : (cx < 0.0) ? -1 : (cfg.convex_unknown_as_convex ? +1 : -1);
            } else {
                R.curved = false;
                R.radius = FLT_MAX;
            }
            const int pidx = (int)planes.size();
            for (int t : R.tris) tri2plane[(size_t)t] = pidx;
            planes.push_back(std::move(R));
        }
        if (n_proxies > cfg.proxy_max_per_group)
            std::fprintf(stderr, "[GA] group %d has %d proxy patches (more than %d)\n", gid, n_proxies,
                         cfg.proxy_max_per_group);
    }
    // flat planes of the sparse groups: coplanar ones share one reflector, also with a flat proxy
    for (Flat &P : flats) {
        int hit = -1;
        for (int r = 0; r < (int)planes.size(); ++r)
            if (!planes[(size_t)r].curved && dot(planes[(size_t)r].n, P.n) > cfg.coplanar_cos &&
                std::fabs(planes[(size_t)r].d - P.d) < cfg.coplanar_dist && planes[(size_t)r].radius == FLT_MAX) {
                hit = r;
                break;
            }
        if (hit < 0) {
            Reflector R;
            R.n = P.n; R.d = P.d; R.area = P.area; R.hit_band = cfg.facet_tol; R.radius = FLT_MAX;
            R.tris = std::move(P.tris);
            planes.push_back(std::move(R));
        } else {
            planes[(size_t)hit].area += P.area;
            auto &dst = planes[(size_t)hit].tris;
            dst.insert(dst.end(), P.tris.begin(), P.tris.end());
        }
    }
    for (int r = 0; r < (int)planes.size(); ++r)
        for (int t : planes[(size_t)r].tris) tri2plane[(size_t)t] = r;
}

// Image tree, breadth-first: every reflector mirrors every image of the previous order
// except its own (and those lying on it); curved reflectors at most max_order_curved
// times per path.
void build_images(const std::vector<Reflector> &planes, const F3 &source, const IsmCfg &cfg, std::vector<Node> &nodes) {
    nodes.clear();
    nodes.push_back(Node{source, -1, -1, 0, 0});
    size_t begin = 0;
    for (int order = 1; order <= cfg.max_order; ++order) {
        const size_t end = nodes.size();
        for (size_t pi = begin; pi < end; ++pi) {
            const Node parent = nodes[pi];
            for (int p = 0; p < (int)planes.size(); ++p) {
                if (p == parent.plane) continue;
                const Reflector &R = planes[(size_t)p];
                if (R.curved && parent.curved_count >= cfg.max_order_curved) continue;
                if (std::fabs(dot(R.n, parent.image) - R.d) < 1e-6f) continue;
                nodes.push_back(Node{mirror(parent.image, R.n, R.d), p, (int)pi, order,
                                     parent.curved_count + (R.curved ? 1 : 0)});
            }
        }
        begin = end;
    }
}

namespace {

int hit_tri(const IsmGeom &g, const Reflector &R, const F3 &x) {
    float bd2 = R.hit_band * R.hit_band;
    int best = -1;
    for (int t : R.tris) {
        const float d2 = tri_dist2(x, g.ta[(size_t)t], g.tb[(size_t)t], g.tc[(size_t)t]);
        if (d2 <= bd2) { bd2 = d2; best = t; }
    }
    return best;
}

bool occluded(const IsmGeom &g, const F3 &p0, const F3 &p1, int ignA, int ignB, float eps) {
    const F3 dir = sub(p1, p0);
    for (size_t t = 0; t < g.ta.size(); ++t) {
        const int pl = g.tri2plane[t];
        if (pl >= 0 && (pl == ignA || pl == ignB)) continue;
        float tt;
        if (ray_tri(p0, dir, g.ta[t], g.tb[t], g.tc[t], eps, tt) && tt < 1.f - eps) return true;
    }
    return false;
}

}  // namespace

// One image path in full (the GPU has already decided it is valid): reflection points
// from the receiver back to the source, hit triangles, incidence cosines, the unfolded
// length and the curvature gain of curved bounces.
struct PathOut {
    float length = 0.f, amp = 1.f;
    bool approx = false;
    std::vector<int> hit;
    std::vector<float> cosv;
};

bool path_details(const IsmGeom &g, const std::vector<Reflector> &planes, const std::vector<Node> &nodes, int leaf,
                  const F3 &source, const F3 &receiver, const IsmCfg &cfg, PathOut &out) {
    std::vector<int> seq, hits;
    std::vector<F3> pts;
    F3 point = receiver;
    int node = leaf;
    while (nodes[(size_t)node].plane >= 0) {
        const Node &nd = nodes[(size_t)node];
        const Reflector &R = planes[(size_t)nd.plane];
        const F3 dir = sub(nd.image, point);
        const float denom = dot(R.n, dir);
        if (std::fabs(denom) < 1e-12f) return false;
        const float t = (R.d - dot(R.n, point)) / denom;
        if (t < -cfg.seg_eps || t >= 1.f - cfg.seg_eps) return false;
        const F3 refl = add(point, scl(dir, t));
        const int h = hit_tri(g, R, refl);
        if (h < 0) return false;
        seq.push_back(nd.plane);
        hits.push_back(h);
        pts.push_back(refl);
        point = refl;
        node = nd.parent;
    }
    const int m = (int)pts.size();
    if (occluded(g, receiver, m > 0 ? pts[0] : source, m > 0 ? seq[0] : -1, -1, cfg.seg_eps)) return false;
    for (int s = 0; s < m; ++s)
        if (occluded(g, pts[(size_t)s], (s + 1 < m) ? pts[(size_t)(s + 1)] : source, seq[(size_t)s],
                     (s + 1 < m) ? seq[(size_t)(s + 1)] : -1, cfg.seg_eps))
            return false;
    out.length = sqrtf(len2(sub(receiver, nodes[(size_t)leaf].image)));
    std::vector<float> cos_ls((size_t)m, 1.f);
    for (int s = 0; s < m; ++s) {
        const F3 P = pts[(size_t)s];
        const F3 prev = (s == 0) ? receiver : pts[(size_t)(s - 1)];
        const F3 din = unit(sub(P, prev));
        cos_ls[(size_t)s] = std::min(1.f, std::max(0.f, std::fabs(dot(din, planes[(size_t)seq[(size_t)s]].n))));
    }
    // the physical polyline source -> bounces -> receiver and its cumulative lengths
    std::vector<F3> V;
    V.reserve((size_t)m + 2);
    V.push_back(source);
    for (int k = m - 1; k >= 0; --k) V.push_back(pts[(size_t)k]);
    V.push_back(receiver);
    std::vector<float> cum(V.size(), 0.f);
    for (size_t i = 1; i < V.size(); ++i) cum[i] = cum[i - 1] + sqrtf(len2(sub(V[i], V[i - 1])));
    const float total = cum.back();
    // spherical-mirror gain of each curved bounce (paraxial image distance)
    float amp = 1.f;
    bool approx = false;
    for (int k = 1; k <= m; ++k) {
        const Reflector &R = planes[(size_t)seq[(size_t)(m - k)]];
        if (!R.curved || R.radius == FLT_MAX) continue;
        const double di = cum[(size_t)k], dr = (double)total - cum[(size_t)k];
        const double ct = std::max(1e-3, (double)cos_ls[(size_t)(m - k)]);
        const double sgn = (R.convex >= 0) ? -1.0 : +1.0;
        const double inv = 1.0 / di + sgn * 2.0 / ((double)R.radius * ct);
        const double di_p = (std::fabs(inv) > 1e-12) ? 1.0 / inv : 1e12;
        double a = (di + dr) / (std::fabs(di_p) + dr);
        if (!(a > 0.0) || a > (double)cfg.curved_amp_cap) {
            a = std::min(std::max(a, 1e-3), (double)cfg.curved_amp_cap);
            approx = true;
        }
        amp *= (float)a;
    }
    out.amp = amp;
    out.approx = approx;
    out.hit.resize((size_t)m);
    out.cosv.resize((size_t)m);
    for (int s = 0; s < m; ++s) {
        out.hit[(size_t)s] = hits[(size_t)(m - 1 - s)];
        out.cosv[(size_t)s] = cos_ls[(size_t)(m - 1 - s)];
    }
    return true;
}

// MemAvailable of /proc/meminfo [bytes], 0 when unknown
double free_ram_bytes() {
    std::ifstream f("/proc/meminfo");
    std::string key;
    double kb;
    while (f >> key >> kb) {
        if (key == "MemAvailable:") return kb * 1024.0;
        f.ignore(256, '\n');
    }
    return 0.0;
}

// Bounding-volume hierarchy: median split on the longest centroid axis, leaves of at most
// 4 triangles; triangles end up in leaf order.
namespace {
struct Prim { F3 a, b, c, cent, bmin, bmax; int group; };
inline void grow(F3 &mn, F3 &mx, const F3 &p) {
    mn.x = std::min(mn.x, p.x); mn.y = std::min(mn.y, p.y); mn.z = std::min(mn.z, p.z);
    mx.x = std::max(mx.x, p.x); mx.y = std::max(mx.y, p.y); mx.z = std::max(mx.z, p.z);
}
int bvh_node(std::vector<BvhNode> &nodes, std::vector<int> &idx, const std::vector<Prim> &prim, int first, int last) {
    const int self = (int)nodes.size();
    nodes.push_back(BvhNode{});
    F3 bmin = f3(FLT_MAX, FLT_MAX, FLT_MAX), bmax = f3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    F3 cmin = bmin, cmax = bmax;
    for (int i = first; i < last; ++i) {
        grow(bmin, bmax, prim[(size_t)idx[(size_t)i]].bmin);
        grow(bmin, bmax, prim[(size_t)idx[(size_t)i]].bmax);
        grow(cmin, cmax, prim[(size_t)idx[(size_t)i]].cent);
    }
    nodes[(size_t)self].bmin = bmin;
    nodes[(size_t)self].bmax = bmax;
    const int n = last - first;
    if (n <= 4) {
        nodes[(size_t)self].left = nodes[(size_t)self].right = -1;
        nodes[(size_t)self].first = first;
        nodes[(size_t)self].count = n;
        return self;
    }
    const float ex = cmax.x - cmin.x, ey = cmax.y - cmin.y, ez = cmax.z - cmin.z;
    const int axis = (ex >= ey && ex >= ez) ? 0 : (ey >= ez ? 1 : 2);
    const int mid = first + n / 2;
    std::nth_element(idx.begin() + first, idx.begin() + mid, idx.begin() + last, [&](int u, int v) {
        const F3 &cu = prim[(size_t)u].cent, &cv = prim[(size_t)v].cent;
        const float ku = (axis == 0 ? cu.x : axis == 1 ? cu.y : cu.z);
        const float kv = (axis == 0 ? cv.x : axis == 1 ? cv.y : cv.z);
        return ku < kv;
    });
    const int L = bvh_node(nodes, idx, prim, first, mid);
    const int R = bvh_node(nodes, idx, prim, mid, last);
    nodes[(size_t)self].left = L;
    nodes[(size_t)self].right = R;
    nodes[(size_t)self].count = 0;
    return self;
}
}  // namespace

RtScene rt_scene(const Mesh &M) {
    RtScene S;
    int max_gid = -1;
    std::vector<Prim> prim;
    prim.reserve(M.tris.size());
    for (const Triangle &t : M.tris) {
        Prim p;
        p.a = vtx(M, t.a); p.b = vtx(M, t.b); p.c = vtx(M, t.c);
        p.cent = scl(add(add(p.a, p.b), p.c), 1.f / 3.f);
        p.bmin = p.a; p.bmax = p.a;
        grow(p.bmin, p.bmax, p.b);
        grow(p.bmin, p.bmax, p.c);
        p.group = t.group;
        max_gid = std::max(max_gid, t.group);
        prim.push_back(p);
    }
    S.ngroups = max_gid + 1;
    if (prim.empty()) return S;
    std::vector<int> idx(prim.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
    S.nodes.reserve(2 * prim.size());
    bvh_node(S.nodes, idx, prim, 0, (int)prim.size());
    const size_t n = prim.size();
    S.a.resize(n); S.b.resize(n); S.c.resize(n); S.nrm.resize(n); S.group.resize(n);
    for (size_t i = 0; i < n; ++i) {
        const Prim &p = prim[(size_t)idx[i]];
        S.a[i] = p.a; S.b[i] = p.b; S.c[i] = p.c; S.group[i] = p.group;
        const F3 nc = cross(sub(p.b, p.a), sub(p.c, p.a));     // rt_tri_normal
        const float l = sqrtf(len2(nc));
        S.nrm[i] = (l > 1e-20f) ? scl(nc, 1.f / l) : f3(0, 0, 1);
    }
    return S;
}

// Fibonacci-sphere direction i of N (rt_fib_ray): evaluated on the host, so the launch
// directions are the reference's to the bit.
F3 fib_dir(int i, int N) {
    const float ga = 2.399963229728653f;
    const float z = 1.f - (2.f * (float)i + 1.f) / (float)N;
    const float rho = std::sqrt(std::max(0.f, 1.f - z * z));
    const float phi = ga * (float)i;
    return f3(rho * std::cos(phi), rho * std::sin(phi), z);
}

}  // namespace ga
}  // namespace fases

// ---------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------

using namespace fases;
using namespace fases::ga;

struct IsmImage {
    float distance, amp;
    float doa[3];
    std::vector<float> cos_inc, angle;
    std::vector<int> group;
};

struct fases_ga {
    int n_src = 0, n_rec = 0, order_used = 0, n_groups = 0, Nt = 0, n_rays = 0;
    float c = 343.f, fs = 44100.f;
    std::vector<F3> src_raw, rec_raw;
    std::vector<float> src_p;
    std::vector<float> direct;                    // [s * n_rec + r], 0 when occluded
    std::vector<std::vector<IsmImage>> images;    // [s * n_rec + r]
    std::vector<float> e_spec, e_diff;            // [(b n_rec + r) Nt + t]
    std::vector<float> alpha, sctr, m_air;        // [n_groups * 8], [n_groups], [8]
    long long n_nodes = 0, n_valid = 0;
    double ism_ms = 0.0, rt_ms = 0.0, rt_kernel_ms = 0.0;
};

namespace {

int ga_fail(int code, const std::string &msg) {
    set_global_error(msg);
    return code;
}

F3 placed_point(const Mesh &M, const float *xyz) {
    return f3((xyz[0] + M.offset.x) * 1.f, (xyz[1] + M.offset.y) * 1.f, (xyz[2] + M.offset.z) * 1.f);
}

}  // namespace

extern "C" {

int fases_ga_params_default(fases_ga_params *p) {
    if (!p) return FASES_E_ARG;
    std::memset(p, 0, sizeof(*p));
    p->device = 0;
    p->ism_order = 3;
    p->n_rays = 10000;
    p->rt_max_order = 100000;
    p->recv_radius = 0.12f;
    p->c = 343.f;
    p->fs = 44100.f;
    p->Nt = 0;
    return FASES_OK;
}

int fases_ga_run(const fases_scene *scene, const fases_ga_params *prm, int n_src, const float *src_xyz,
                 const float *src_pressure, int n_rec, const float *rec_xyz, const float *alpha, const float *sctr,
                 const float *m_air, fases_ga **out) {
    if (!scene || !prm || !out || n_src < 1 || n_rec < 1 || !src_xyz || !rec_xyz)
        return ga_fail(FASES_E_ARG, "fases_ga_run: bad argument");
    if (prm->Nt < 1 || !(prm->fs > 0.f) || !(prm->c > 0.f) || prm->n_rays < 0 || !(prm->recv_radius >= 0.f))
        return ga_fail(FASES_E_ARG, "fases_ga_run: Nt, fs, c must be positive");
    if (prm->ism_order > kMaxIsmOrder)
        return ga_fail(FASES_E_ARG, "fases_ga_run: image-source order above " + std::to_string(kMaxIsmOrder));
    *out = nullptr;
    const Scene &sc = *reinterpret_cast<const Scene *>(scene);
    const Mesh &M = sc.mesh;
    std::unique_ptr<fases_ga> G(new (std::nothrow) fases_ga());
    if (!G) return ga_fail(FASES_E_NOMEM, "fases_ga_run: out of host memory");
    G->n_src = n_src; G->n_rec = n_rec; G->Nt = prm->Nt; G->c = prm->c; G->fs = prm->fs; G->n_rays = prm->n_rays;
    int max_gid = -1;
    for (const Triangle &t : M.tris) max_gid = std::max(max_gid, t.group);
    G->n_groups = max_gid + 1;
    const int ng = G->n_groups;
    G->alpha.assign((size_t)ng * kBands, 0.f);
    G->sctr.assign((size_t)ng, 0.f);
    G->m_air.assign(kBands, 0.f);
    if (alpha) std::copy(alpha, alpha + (size_t)ng * kBands, G->alpha.begin());
    if (sctr) std::copy(sctr, sctr + ng, G->sctr.begin());
    if (m_air) std::copy(m_air, m_air + kBands, G->m_air.begin());
    for (int s = 0; s < n_src; ++s) {
        G->src_raw.push_back(f3(src_xyz[3 * s], src_xyz[3 * s + 1], src_xyz[3 * s + 2]));
        G->src_p.push_back(src_pressure ? src_pressure[s] : 1.0f);
    }
    for (int r = 0; r < n_rec; ++r) G->rec_raw.push_back(f3(rec_xyz[3 * r], rec_xyz[3 * r + 1], rec_xyz[3 * r + 2]));

    // ---- image sources
    const auto t0 = std::chrono::steady_clock::now();
    IsmCfg cfg;
    cfg.max_order = std::max(0, prm->ism_order);
    std::vector<Reflector> planes;
    IsmGeom geo;
    build_reflectors(M, cfg, planes, geo.tri2plane);
    {   // the tree must fit in half the free host memory (28 bytes a node, x3 for vector growth)
        const double ram = free_ram_bytes();
        const double cap = (ram > 0.0) ? 0.5 * ram / (3.0 * (double)sizeof(Node)) : DBL_MAX;
        const double P = (double)planes.size();
        double total = 1.0, level = 1.0;
        int ord = 0;
        while (ord < cfg.max_order) {
            level *= (ord == 0) ? P : std::max(P - 1.0, 0.0);
            if (total + level > cap) break;
            total += level;
            ++ord;
        }
        if (ord < cfg.max_order) {
            std::fprintf(stderr, "\n[!] image-source order reduced %d -> %d (the image tree does not fit in host memory)\n",
                        cfg.max_order, ord);
            cfg.max_order = ord;
        }
    }
    G->order_used = cfg.max_order;
    for (const Triangle &t : M.tris) {
        geo.ta.push_back(vtx(M, t.a));
        geo.tb.push_back(vtx(M, t.b));
        geo.tc.push_back(vtx(M, t.c));
    }
    for (const Reflector &R : planes) {
        geo.pn.push_back(R.n);
        geo.pd.push_back(R.d);
        geo.pband.push_back(R.hit_band);
        geo.poff.push_back((int)geo.ptri.size());
        for (int t : R.tris) geo.ptri.push_back(t);
    }
    geo.poff.push_back((int)geo.ptri.size());
    std::vector<F3> recs;
    for (int r = 0; r < n_rec; ++r) recs.push_back(placed_point(M, rec_xyz + 3 * r));
    G->direct.assign((size_t)n_src * n_rec, 0.f);
    G->images.assign((size_t)n_src * n_rec, {});
    for (int s = 0; s < n_src; ++s) {
        const F3 source = placed_point(M, src_xyz + 3 * s);
        std::vector<Node> nodes;
        build_images(planes, source, cfg, nodes);
        G->n_nodes += (long long)nodes.size();
        std::vector<uint8_t> valid;
        std::string err;
        int rc = FASES_OK;
        if (nodes.size() > 1) {
            if (prm->device >= 0) {
                rc = ism_valid_gpu(prm->device, geo, nodes, source, recs, cfg.seg_eps, valid, err);
            } else {
                valid.assign((nodes.size() - 1) * (size_t)n_rec, 0);
                PathOut po;
                for (size_t leaf = 1; leaf < nodes.size(); ++leaf)
                    for (int r = 0; r < n_rec; ++r)
                        valid[(leaf - 1) * (size_t)n_rec + r] =
                            path_details(geo, planes, nodes, (int)leaf, source, recs[(size_t)r], cfg, po) ? 1 : 0;
            }
            if (rc) return ga_fail(rc, err);
        }
        for (int r = 0; r < n_rec; ++r) {
            const F3 receiver = recs[(size_t)r];
            struct Path { float length = 0.f; int leaf = 0; PathOut d; };
            std::vector<Path> dst;
            if (!occluded(geo, source, receiver, -1, -1, cfg.seg_eps)) {
                Path p0;
                p0.length = sqrtf(len2(sub(receiver, source)));
                p0.leaf = 0;
                dst.push_back(p0);
            }
            // images at one position (commuting reflections) count once: the first valid
            std::set<std::tuple<int, int, int>> seen;
            auto key = [](const F3 &v) {
                return std::make_tuple((int)std::round(v.x * 1e4f), (int)std::round(v.y * 1e4f),
                                       (int)std::round(v.z * 1e4f));
            };
            for (size_t leaf = 1; leaf < nodes.size(); ++leaf) {
                if (!valid[(leaf - 1) * (size_t)n_rec + r]) continue;
                const auto k = key(nodes[leaf].image);
                if (seen.count(k)) continue;
                Path p;
                if (!path_details(geo, planes, nodes, (int)leaf, source, receiver, cfg, p.d))
                    return ga_fail(FASES_E_STATE, "fases_ga_run: GPU and host disagree on an image path");
                seen.insert(k);
                p.length = p.d.length;
                p.leaf = (int)leaf;
                dst.push_back(std::move(p));
                ++G->n_valid;
            }
            std::sort(dst.begin(), dst.end(), [](const Path &a, const Path &b) { return a.length < b.length; });
            auto &imgs = G->images[(size_t)s * n_rec + r];
            for (const Path &p : dst) {
                if (p.leaf == 0) { G->direct[(size_t)s * n_rec + r] = p.length; continue; }
                IsmImage im;
                im.distance = p.length;
                im.amp = p.d.amp;
                const F3 dd = unit(sub(nodes[(size_t)p.leaf].image, receiver));
                im.doa[0] = dd.x; im.doa[1] = dd.y; im.doa[2] = dd.z;
                for (size_t k = 0; k < p.d.hit.size(); ++k) {
                    const float ct = p.d.cosv[k];
                    im.cos_inc.push_back(ct);
                    im.angle.push_back(std::acos(std::min(1.f, std::max(0.f, ct))));
                    im.group.push_back(M.tris[(size_t)p.d.hit[k]].group);
                }
                imgs.push_back(std::move(im));
            }
        }
    }
    G->ism_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();

    // ---- ray tracer: specular orders above the image sources' and all scattered energy
    G->e_spec.assign((size_t)kBands * n_rec * G->Nt, 0.f);
    G->e_diff.assign((size_t)kBands * n_rec * G->Nt, 0.f);
    const auto t1 = std::chrono::steady_clock::now();
    if (prm->n_rays > 0 && prm->recv_radius > 0.f && !M.tris.empty()) {
        if (prm->device < 0) return ga_fail(FASES_E_ARG, "fases_ga_run: the ray tracer runs on a HIP device");
        const RtScene S = rt_scene(M);
        std::vector<F3> dirs((size_t)prm->n_rays);
        for (int i = 0; i < prm->n_rays; ++i) dirs[(size_t)i] = fib_dir(i, prm->n_rays);
        std::vector<Refl> refl((size_t)S.ngroups * kBands);
        std::vector<float> sg((size_t)S.ngroups, 0.f);
        for (int g = 0; g < S.ngroups; ++g) {
            for (int b = 0; b < kBands; ++b) {
                float al = (g < ng) ? G->alpha[(size_t)g * kBands + b] : 0.f;
                al = std::max(0.f, std::min(0.999f, al));
                refl[(size_t)g * kBands + b] = refl_of(al);
            }
            sg[(size_t)g] = (sctr && g < ng) ? std::max(0.f, std::min(1.f, G->sctr[(size_t)g])) : 0.f;
        }
        const float R2 = prm->recv_radius * prm->recv_radius;
        for (int s0 = 0; s0 < n_src; s0 += 64) {
            RtParams P;
            std::memset(&P, 0, sizeof(P));
            P.n_src = std::min(64, n_src - s0);
            P.s_base = s0;
            P.n_rec = n_rec;
            P.n_rays = prm->n_rays;
            P.max_order = prm->rt_max_order;
            P.Nt = G->Nt;
            P.ism_order = G->order_used;
            P.c = prm->c;
            P.fs = prm->fs;
            P.recv_radius = prm->recv_radius;
            std::vector<F3> srcs;
            for (int s = 0; s < P.n_src; ++s) {
                const float ps = G->src_p[(size_t)(s0 + s)];
                P.Kdiff[s] = ps * ps / (4.f * M_PI * (float)prm->n_rays);   // power of one ray
                P.Kspec[s] = P.Kdiff[s] * 0.318309886183791f / R2;
                srcs.push_back(placed_point(M, src_xyz + 3 * (s0 + s)));
            }
            std::string err;
            double kms = 0.0;
            const int rc = rt_trace_gpu(prm->device, S, P, srcs, recs, dirs, refl, sg, G->m_air, G->e_spec.data(),
                                        G->e_diff.data(), &kms, err);
            if (rc) return ga_fail(rc, err);
            G->rt_kernel_ms += kms;
        }
    }
    G->rt_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
    *out = G.release();
    return FASES_OK;
}

void fases_ga_free(fases_ga *g) { delete g; }

int fases_ga_info(const fases_ga *g, int *order_used, int64_t *n_nodes, int64_t *n_paths, double *ism_ms,
                  double *rt_ms, double *rt_kernel_ms) {
    if (!g) return FASES_E_ARG;
    if (order_used) *order_used = g->order_used;
    if (n_nodes) *n_nodes = g->n_nodes;
    if (n_paths) *n_paths = g->n_valid;
    if (ism_ms) *ism_ms = g->ism_ms;
    if (rt_ms) *rt_ms = g->rt_ms;
    if (rt_kernel_ms) *rt_kernel_ms = g->rt_kernel_ms;
    return FASES_OK;
}

int fases_ga_direct(const fases_ga *g, int s, int r, float *dist) {
    if (!g || !dist || s < 0 || s >= g->n_src || r < 0 || r >= g->n_rec) return FASES_E_ARG;
    *dist = g->direct[(size_t)s * g->n_rec + r];
    return FASES_OK;
}

int fases_ga_images(const fases_ga *g, int s, int r, int *n_images, int *n_bounces, float *distance, float *doa,
                    float *amp, int *order, int *group, float *cos_inc, float *angle) {
    if (!g || s < 0 || s >= g->n_src || r < 0 || r >= g->n_rec) return FASES_E_ARG;
    const auto &v = g->images[(size_t)s * g->n_rec + r];
    int nb = 0;
    for (size_t q = 0; q < v.size(); ++q) {
        const IsmImage &im = v[q];
        if (distance) distance[q] = im.distance;
        if (amp) amp[q] = im.amp;
        if (doa) for (int k = 0; k < 3; ++k) doa[3 * q + (size_t)k] = im.doa[k];
        if (order) order[q] = (int)im.group.size();
        for (size_t k = 0; k < im.group.size(); ++k, ++nb) {
            if (group) group[nb] = im.group[k];
            if (cos_inc) cos_inc[nb] = im.cos_inc[k];
            if (angle) angle[nb] = im.angle[k];
        }
    }
    if (n_images) *n_images = (int)v.size();
    if (n_bounces) *n_bounces = nb;
    return FASES_OK;
}

int fases_ga_energy(const fases_ga *g, float *e_spec, float *e_diff) {
    if (!g) return FASES_E_ARG;
    if (e_spec) std::copy(g->e_spec.begin(), g->e_spec.end(), e_spec);
    if (e_diff) std::copy(g->e_diff.begin(), g->e_diff.end(), e_diff);
    return FASES_OK;
}

int fases_ga_band_irs(const fases_ga *g, int band, float *p_hf, float *p_rt) {
    if (!g || band < 0 || band >= kBands || !p_hf) return FASES_E_ARG;
    const int Nt = g->Nt, Nr = g->n_rec;
    const int Nw = 32;                              // windowed-sinc fractional delay, 33 taps
    std::fill(p_hf, p_hf + (size_t)Nt * Nr, 0.f);
    if (p_rt) std::fill(p_rt, p_rt + (size_t)Nt * Nr, 0.f);
    const float fs = g->fs, c = g->c;
    const float mair = g->m_air[(size_t)band];
    auto sinc = [](float x) -> float { return x == 0.0 ? 1.0 : std::sin(x) / x; };
    auto deposit = [&](float tau, float gain, int rr) {
        const int tt = (int)tau;
        for (int n0 = -(int)(0.5 * Nw); n0 <= (int)(0.5 * Nw); n0++) {
            const int nn = tt + n0;
            if (nn < 0 || nn >= Nt) continue;
            const float xx = nn - tau;
            float &cell = p_hf[(size_t)rr * Nt + nn];
            cell += gain * 0.5 * (1.0 + std::cos(2.0 * M_PI * xx / Nw)) * sinc(M_PI * xx);
        }
    };
    for (int r = 0; r < Nr; ++r) {
        // one random sign per time sample and receiver, the same in every band
        auto asign = [&](unsigned k) { return hash_sign(k * 2654435761u + (unsigned)r * 40503u); };
        for (int s = 0; s < g->n_src; ++s) {
            const float d0 = g->direct[(size_t)s * Nr + r];
            if (d0 > 0.0f) {
                const float tau0 = fs * d0 / c;
                if ((int)tau0 < Nt) deposit(tau0, g->src_p[(size_t)s] / d0 * std::exp(-mair * d0), r);
            }
            for (const IsmImage &im : g->images[(size_t)s * Nr + r]) {
                const float d = im.distance;
                const float tau = fs * d / c;
                if ((int)tau >= Nt) continue;
                float P = 1.0;
                for (size_t k = 0; k < im.angle.size(); ++k) {
                    const int gk = im.group[k];
                    const float al = (gk >= 0 && gk < g->n_groups) ? g->alpha[(size_t)gk * kBands + band] : 0.f;
                    float R = refl_at(refl_of(al), std::cos(im.angle[k]));
                    if (gk >= 0 && gk < g->n_groups) R *= std::sqrt(std::max(0.0f, 1.0f - g->sctr[(size_t)gk]));
                    P *= R;
                }
                deposit(tau, g->src_p[(size_t)s] * im.amp * P / d * std::exp(-mair * d), r);
            }
        }
        // ray-traced energy -> pressure with a random sign (p = ps/d convention: x 4 pi)
        for (int t = 1; t < Nt; ++t) {
            float e = 0.0;
            e += g->e_diff[((size_t)band * Nr + r) * Nt + t];
            e += g->e_spec[((size_t)band * Nr + r) * Nt + t];
            if (e <= 0.0) continue;
            const double v = 4.0 * M_PI * asign((unsigned)t) * (float)std::sqrt(e);
            p_hf[(size_t)r * Nt + t] += v;
            if (p_rt) p_rt[(size_t)r * Nt + t] += v;
        }
    }
    return FASES_OK;
}

int fases_ga_arrivals(const fases_ga *g, int r, int *n, float *buf) {
    if (!g || !n || r < 0 || r >= g->n_rec) return FASES_E_ARG;
    const float fs = g->fs, c = g->c;
    const int Nt = g->Nt, Nr = g->n_rec;
    int cnt = 0;
    auto push = [&](float delay, const float dir[3], const float gain[8]) {
        if (buf) {
            float *o = buf + (size_t)cnt * 12;
            o[0] = delay; o[1] = dir[0]; o[2] = dir[1]; o[3] = dir[2];
            for (int b = 0; b < kBands; ++b) o[4 + b] = gain[b];
        }
        ++cnt;
    };
    const F3 rr = g->rec_raw[(size_t)r];
    for (int s = 0; s < g->n_src; ++s) {
        const float d0 = g->direct[(size_t)s * Nr + r];
        if (d0 > 0.0f && (int)(fs * d0 / c) < Nt) {
            float dir[3] = {g->src_raw[(size_t)s].x - rr.x, g->src_raw[(size_t)s].y - rr.y, g->src_raw[(size_t)s].z - rr.z};
            const float nrm = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
            if (nrm > 1e-12f) { dir[0] /= nrm; dir[1] /= nrm; dir[2] /= nrm; }
            float gain[8];
            for (int b = 0; b < kBands; ++b) gain[b] = g->src_p[(size_t)s] / d0 * std::exp(-g->m_air[(size_t)b] * d0);
            push((float)(d0 / c), dir, gain);
        }
        for (const IsmImage &im : g->images[(size_t)s * Nr + r]) {
            const float d = im.distance;
            if ((int)(fs * d / c) >= Nt) continue;
            float gain[8];
            for (int b = 0; b < kBands; ++b) {
                float P = 1.0f;
                for (size_t k = 0; k < im.angle.size(); ++k) {
                    const int gk = im.group[k];
                    const float al = (gk >= 0 && gk < g->n_groups) ? g->alpha[(size_t)gk * kBands + b] : 0.f;
                    float R = refl_at(refl_of(al), std::cos(im.angle[k]));
                    if (gk >= 0 && gk < g->n_groups) R *= std::sqrt(std::max(0.0f, 1.0f - g->sctr[(size_t)gk]));
                    P *= R;
                }
                gain[b] = g->src_p[(size_t)s] * im.amp * P / d * std::exp(-g->m_air[(size_t)b] * d);
            }
            push((float)(d / c), im.doa, gain);
        }
    }
    *n = cnt;
    return FASES_OK;
}

int fases_ga_dump_rays(const fases_scene *scene, int n_src, const float *src_xyz, int n_rays, int max_order,
                       const char *path) {
    if (!scene || !path || n_src < 0 || (n_src > 0 && !src_xyz) || n_rays < 1 || max_order < 1)
        return ga_fail(FASES_E_ARG, "fases_ga_dump_rays: bad argument");
    const Scene &sc = *reinterpret_cast<const Scene *>(scene);
    const Mesh &M = sc.mesh;
    std::FILE *f = std::fopen(path, "w");
    if (!f) return ga_fail(FASES_E_IO, std::string("cannot write ") + path);
    for (size_t g = 0; g < M.groups.size(); ++g) std::fprintf(f, "# group %d %s\n", (int)g, M.groups[g].c_str());
    const RtScene S = rt_scene(M);
    std::vector<F3> srcs;
    for (int q = 0; q < n_src; ++q) srcs.push_back(f3(src_xyz[3 * q], src_xyz[3 * q + 1], src_xyz[3 * q + 2]));
    if (srcs.empty()) srcs.push_back(f3(0, 0, 0));
    if (S.nodes.empty()) { std::fclose(f); return FASES_OK; }
    // escaped rays get a stub of a quarter of the scene diagonal (order and group -1)
    const float stub = 0.25f * sqrtf(len2(sub(S.nodes[0].bmax, S.nodes[0].bmin)));
    auto world = [&](const F3 &p) { return f3(p.x / 1.f - M.offset.x, p.y / 1.f - M.offset.y, p.z / 1.f - M.offset.z); };
    int uid = 0;
    for (size_t si = 0; si < srcs.size(); ++si) {
        const F3 o0 = placed_point(M, &srcs[si].x);
        for (int ri = 0; ri < n_rays; ++ri, ++uid) {
            F3 o = o0, d = fib_dir(ri, n_rays);
            for (int b = 0; b < max_order; ++b) {
                float th;
                int tri;
                const bool hit = closest_hit(S.nodes.data(), S.a.data(), S.b.data(), S.c.data(), o, d, 1e-4f, th, tri);
                const F3 p = hit ? add(o, scl(d, th)) : add(o, scl(d, stub));
                const F3 ow = world(o), pw = world(p);
                std::fprintf(f, "%d %d %.6g %.6g %.6g %.6g %.6g %.6g %d %d\n", (int)si, uid, ow.x, ow.y, ow.z, pw.x, pw.y,
                             pw.z, hit ? b : -1, hit ? S.group[(size_t)tri] : -1);
                if (!hit) break;
                const F3 ref = reflected(d, S.nrm[(size_t)tri]);
                o = f3(p.x + ref.x * 1e-4f, p.y + ref.y * 1e-4f, p.z + ref.z * 1e-4f);
                d = ref;
            }
        }
    }
    std::fclose(f);
    return FASES_OK;
}

}  // extern "C"
