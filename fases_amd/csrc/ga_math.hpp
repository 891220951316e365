// Float geometry of the geometrical-acoustics part, one definition for the host
// (host_ga.cpp, g++) and the HIP kernels (fvm_ga.hip): point-triangle distance (Ericson,
// Real-Time Collision Detection, 2005, sec. 5.1.5), ray-triangle crossing (Moller and
// Trumbore, 1997), mirror images, the oblique reflection factor and the ray tracer's hash.
// Operand order follows 3D-solver/include/parsing3D.h, ism3D.h and raytrace3D.h; both
// translation units are compiled without FMA contraction, so each function returns the
// same bits on either side.
#pragma once

#include "ga.hpp"

#include <cmath>

#ifdef __HIPCC__
#define GA_FN __host__ __device__ inline
#else
#define GA_FN inline
#endif

namespace fases {
namespace ga {

GA_FN F3 f3(float x, float y, float z) { F3 r; r.x = x; r.y = y; r.z = z; return r; }
GA_FN F3 add(const F3 &a, const F3 &b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
GA_FN F3 sub(const F3 &a, const F3 &b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
GA_FN F3 scl(const F3 &a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
GA_FN float dot(const F3 &a, const F3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
GA_FN F3 cross(const F3 &a, const F3 &b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
GA_FN float len2(const F3 &a) { return dot(a, a); }

// unit vector, or 0 for a vanishing one
GA_FN F3 unit(const F3 &v) {
    const float l = sqrtf(len2(v));
    return (l > 1e-20f) ? scl(v, 1.f / l) : f3(0.f, 0.f, 0.f);
}

// squared distance from p to triangle (a, b, c): walk of the Voronoi regions
GA_FN float tri_dist2(const F3 &p, const F3 &a, const F3 &b, const F3 &c) {
    const F3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.f && d2 <= 0.f) return len2(ap);
    const F3 bp = sub(p, b);
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.f && d4 <= d3) return len2(bp);
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) return len2(sub(p, add(a, scl(ab, d1 / (d1 - d3)))));
    const F3 cp = sub(p, c);
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.f && d5 <= d6) return len2(cp);
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) return len2(sub(p, add(a, scl(ac, d2 / (d2 - d6)))));
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f)
        return len2(sub(p, add(b, scl(sub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6))))));
    const float den = 1.f / (va + vb + vc);
    const float v = vb * den, w = vc * den;
    return len2(sub(p, add(a, add(scl(ab, v), scl(ac, w)))));
}

// ray o + t d against triangle (a, b, c), edges included; a crossing with t > tmin
GA_FN bool ray_tri(const F3 &o, const F3 &d, const F3 &a, const F3 &b, const F3 &c, float tmin, float &t_out) {
    const F3 e1 = sub(b, a), e2 = sub(c, a);
    const F3 pv = cross(d, e2);
    const float det = dot(e1, pv);
    if (det > -1e-8f && det < 1e-8f) return false;
    const float inv = 1.f / det;
    const F3 tv = sub(o, a);
    const float u = dot(tv, pv) * inv;
    if (u < 0.f || u > 1.f) return false;
    const F3 qv = cross(tv, e1);
    const float v = dot(d, qv) * inv;
    if (v < 0.f || u + v > 1.f) return false;
    const float t = dot(e2, qv) * inv;
    if (t > tmin) { t_out = t; return true; }
    return false;
}

// mirror image of p in the plane n.x = d
GA_FN F3 mirror(const F3 &p, const F3 &n, float d) { return sub(p, scl(n, 2.f * (dot(n, p) - d))); }

// Oblique-incidence pressure reflection factor of a locally reacting surface with random-
// incidence absorption alpha (parsing3D.h:1178-1186): xi = (1 + R0)/(1 - R0), R0 = sqrt(1 -
// alpha), R = (xi cos - 1)/(xi cos + 1); anechoic above alpha 0.99, rigid-like above R0
// 0.99999. The angle-free part per (group, band) is kept as Refl.
GA_FN Refl refl_of(float alpha) {
    Refl r;
    r.kind = 2; r.r0 = 0.f; r.xi = 0.f;
    if (alpha > 0.99f) { r.kind = 0; return r; }
    const float R = sqrtf((float)(1.0 - (double)alpha));
    if ((double)R > 0.99999) { r.kind = 1; r.r0 = R; return r; }
    r.xi = (float)((1.0 + (double)R) / (1.0 - (double)R));
    return r;
}
GA_FN float refl_at(const Refl &r, float cos_theta) {
    if (r.kind == 0) return 0.0f;
    if (r.kind == 1) return r.r0;
    const float xc = r.xi * cos_theta;
    return (float)(((double)xc - 1.0) / ((double)xc + 1.0));
}

// slab test of a box against the ray segment [tmin, tmax] (inv = 1 / direction); the
// comparisons keep the reference's std::max / std::min order, so a 0 * inf NaN drops out
GA_FN bool slab(const F3 &bmin, const F3 &bmax, const F3 &ro, const F3 &inv, float tmin, float tmax) {
    float t0 = tmin, t1 = tmax, a, b;
    a = (bmin.x - ro.x) * inv.x; b = (bmax.x - ro.x) * inv.x;
    if (a > b) { const float s = a; a = b; b = s; }
    t0 = (t0 < a) ? a : t0; t1 = (b < t1) ? b : t1;
    a = (bmin.y - ro.y) * inv.y; b = (bmax.y - ro.y) * inv.y;
    if (a > b) { const float s = a; a = b; b = s; }
    t0 = (t0 < a) ? a : t0; t1 = (b < t1) ? b : t1;
    a = (bmin.z - ro.z) * inv.z; b = (bmax.z - ro.z) * inv.z;
    if (a > b) { const float s = a; a = b; b = s; }
    t0 = (t0 < a) ? a : t0; t1 = (b < t1) ? b : t1;
    return t1 >= t0;
}

// closest triangle hit beyond tmin over a BVH (right child popped first, as the
// reference traverses, so equal distances resolve to the same triangle)
GA_FN bool closest_hit(const BvhNode *nodes, const F3 *A, const F3 *B, const F3 *C, const F3 &ro, const F3 &rd,
                       float tmin, float &t_out, int &tri_out) {
    const F3 inv = f3(1.f / rd.x, 1.f / rd.y, 1.f / rd.z);
    float best = 3.402823466e+38f;
    int bi = -1;
    int stack[kBvhStack];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const BvhNode nd = nodes[stack[--sp]];
        if (!slab(nd.bmin, nd.bmax, ro, inv, tmin, best)) continue;
        if (nd.count > 0) {
            for (int k = 0; k < nd.count; ++k) {
                const int p = nd.first + k;
                float tt;
                if (ray_tri(ro, rd, A[p], B[p], C[p], tmin, tt) && tt < best) { best = tt; bi = p; }
            }
        } else if (sp < kBvhStack - 2) {
            stack[sp++] = nd.left;
            stack[sp++] = nd.right;
        }
    }
    if (bi < 0) return false;
    t_out = best;
    tri_out = bi;
    return true;
}

// mirror direction of a unit incident direction about N, renormalised
GA_FN F3 reflected(const F3 &inc, const F3 &N) {
    const float k = 2.f * dot(inc, N);
    F3 ref = f3(inc.x - k * N.x, inc.y - k * N.y, inc.z - k * N.z);
    const float l = sqrtf(len2(ref));
    if (l > 1e-20f) ref = scl(ref, 1.f / l);
    return ref;
}

// integer hash of the ray tracer (raytrace3D.h rt_sign / rt_u01)
GA_FN unsigned mix32(unsigned s) {
    s ^= s >> 16; s *= 0x7feb352du; s ^= s >> 15; s *= 0x846ca68bu; s ^= s >> 16;
    return s;
}
GA_FN float hash_sign(unsigned s) { return (mix32(s) & 1u) ? 1.f : -1.f; }
GA_FN float hash_u01(unsigned s) { return (float)(mix32(s) >> 8) * (1.f / 16777216.f); }

}  // namespace ga
}  // namespace fases
