// Geometrical acoustics on the GPU: validity of every image-source path (one thread per
// image and receiver: reflection points, in-facet tests, occlusion against every
// triangle) and the ray tracer (one thread per ray: BVH traversal, detection spheres,
// diffuse rain, on-off scattering), whose energy echograms are accumulated with float
// atomics.
//
// Follows 3D-solver/include/ism3D.h (ism_validate_path :424-503, ism_occluded :235-247,
// ism_hit_tri :228-234) and 3D-solver/include/raytrace3D.h (rt_slab :264-272,
// rt_closest_hit :276-299, rt_point_seg_dist2 :310-318, rt_trace_late :335-568,
// rt_lambert_dir :81-92, rt_compute_reflected :60-66). Compiled without FMA contraction;
// geometry decisions equal the host's (ga_math.hpp). Energies differ from a sequential
// sum only by the order of the float additions (and transcendental functions by an ulp).

#include "ga.hpp"
#include "ga_math.hpp"
#include "../../include/fases_core.h"

#include <hip/hip_runtime.h>
#include <string>
#include <vector>

namespace fases {
namespace ga {
namespace {

// ---------------------------------------------------------------------------------
// image-source validity
// ---------------------------------------------------------------------------------
struct IsmDev {
    const F3 *ta, *tb, *tc;
    const int *tri2plane;
    int ntri;
    const F3 *pn;
    const float *pd, *pband;
    const int *poff, *ptri;
    const Node *nodes;
    int nleaf;                       // nodes 1 .. nleaf
    const F3 *recs;
    int nrec;
    F3 src;
    float eps;
};

__device__ bool occluded_dev(const IsmDev &g, const F3 &p0, const F3 &p1, int ignA, int ignB) {
    const F3 dir = sub(p1, p0);
    for (int t = 0; t < g.ntri; ++t) {
        const int pl = g.tri2plane[t];
        if (pl >= 0 && (pl == ignA || pl == ignB)) continue;
        float tt;
        if (ray_tri(p0, dir, g.ta[t], g.tb[t], g.tc[t], g.eps, tt) && tt < 1.f - g.eps) return true;
    }
    return false;
}

__global__ void __launch_bounds__(64) k_ism_valid(IsmDev g, uint8_t *__restrict__ valid) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long)g.nleaf * g.nrec) return;
    const int leaf = 1 + (int)(gid / g.nrec), r = (int)(gid % g.nrec);
    const F3 receiver = g.recs[r];
    F3 pts[kMaxIsmOrder];
    int seq[kMaxIsmOrder];
    int m = 0;
    F3 point = receiver;
    int node = leaf;
    bool ok = true;
    while (g.nodes[node].plane >= 0) {
        const Node nd = g.nodes[node];
        const int pl = nd.plane;
        const F3 n = g.pn[pl];
        const float d = g.pd[pl];
        const F3 dir = sub(nd.image, point);
        const float denom = dot(n, dir);
        if (fabsf(denom) < 1e-12f) { ok = false; break; }
        const float t = (d - dot(n, point)) / denom;
        if (t < -g.eps || t >= 1.f - g.eps) { ok = false; break; }
        const F3 refl = add(point, scl(dir, t));
        const float b2 = g.pband[pl] * g.pband[pl];
        bool in = false;
        for (int q = g.poff[pl]; q < g.poff[pl + 1] && !in; ++q) {
            const int tr = g.ptri[q];
            in = tri_dist2(refl, g.ta[tr], g.tb[tr], g.tc[tr]) <= b2;
        }
        if (!in) { ok = false; break; }
        seq[m] = pl;
        pts[m] = refl;
        ++m;
        point = refl;
        node = nd.parent;
    }
    if (ok) ok = !occluded_dev(g, receiver, m > 0 ? pts[0] : g.src, m > 0 ? seq[0] : -1, -1);
    for (int s = 0; ok && s < m; ++s)
        ok = !occluded_dev(g, pts[s], (s + 1 < m) ? pts[s + 1] : g.src, seq[s], (s + 1 < m) ? seq[s + 1] : -1);
    valid[gid] = ok ? 1 : 0;
}

// ---------------------------------------------------------------------------------
// ray tracer
// ---------------------------------------------------------------------------------
struct RtDev {
    const BvhNode *nodes;
    const F3 *a, *b, *c, *nrm;
    const int *group;
    int ngroups;
    const Refl *refl;                // [ngroups][8], absorption clamped to [0, 0.999]
    const float *sctr;               // [ngroups], clamped to [0, 1]
    const F3 *dirs, *src, *rec;
    float *e_spec, *e_diff;          // [(b n_rec + r) Nt + t]
    float m_air[kBands];
};

__device__ inline bool closest_hit(const RtDev &S, const F3 &ro, const F3 &rd, float tmin, float &t_out,
                                   int &tri_out) {
    return fases::ga::closest_hit(S.nodes, S.a, S.b, S.c, ro, rd, tmin, t_out, tri_out);
}

__device__ inline float seg_dist2(const F3 &q, const F3 &p0, const F3 &p1, float &s_out) {
    const F3 d = sub(p1, p0);
    const float L2 = len2(d);
    float t = (L2 > 1e-20f) ? dot(sub(q, p0), d) / L2 : 0.f;
    t = fmaxf(0.f, fminf(1.f, t));
    s_out = t * sqrtf(L2);
    return len2(sub(q, add(p0, scl(d, t))));
}

__device__ inline F3 lambert(const F3 &N, float u1, float u2) {
    const float r = sqrtf(u1), phi = 6.283185307f * u2;
    const float x = r * cosf(phi), y = r * sinf(phi), z = sqrtf(fmaxf(0.f, 1.f - u1));
    const F3 a = (fabsf(N.z) < 0.999f) ? f3(0, 0, 1) : f3(1, 0, 0);
    F3 t = cross(a, N);
    const float tl = sqrtf(len2(t));
    t = (tl > 1e-8f) ? scl(t, 1.f / tl) : f3(1, 0, 0);
    const F3 bt = cross(N, t);
    const F3 dir = f3(x * t.x + y * bt.x + z * N.x, x * t.y + y * bt.y + z * N.y, x * t.z + y * bt.z + z * N.z);
    const float dl = sqrtf(len2(dir));
    return (dl > 1e-8f) ? scl(dir, 1.f / dl) : N;
}

__global__ void __launch_bounds__(64) k_rt(RtDev S, RtParams P) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long)P.n_src * P.n_rays) return;
    const int s = (int)(gid / P.n_rays), ri = (int)(gid % P.n_rays);
    const int Nr = P.n_rec, Nt = P.Nt;
    const float R2 = P.recv_radius * P.recv_radius;
    const float t_end = (float)Nt / P.fs;
    const float Kdiff = P.Kdiff[s], Kspec = P.Kspec[s];
    const float inv_pi = 0.318309886183791f;
    F3 o = S.src[s], d = S.dirs[ri];
    float Rp[kBands];
    for (int b = 0; b < kBands; ++b) Rp[b] = 1.f;
    float d_travel = 0.f;
    bool scattered = false, after_diffuse = false;
    for (int bounce = 0; bounce <= P.max_order; ++bounce) {
        float th;
        int tri;
        const bool hit = closest_hit(S, o, d, 1e-4f, th, tri);
        const float seg = hit ? th : 1e4f;
        const F3 p1 = add(o, scl(d, seg));
        // detection sphere: pure specular paths above the image-source order go to e_spec,
        // scattered ones to e_diff; not the segment right after a diffuse draw (the rain's)
        float *e_buf = scattered ? S.e_diff : S.e_spec;
        const bool ism_owned = !scattered && bounce <= P.ism_order;
        for (int rr = 0; !after_diffuse && !ism_owned && rr < Nr; ++rr) {
            float s_ca;
            if (seg_dist2(S.rec[rr], o, p1, s_ca) <= R2) {
                const float d_run = d_travel + s_ca;
                const float tt = d_run / P.c;
                if (tt < t_end && d_run > 1e-4f) {
                    const int n0 = (int)(tt * P.fs);
                    if (n0 >= 0 && n0 < Nt)
                        for (int b = 0; b < kBands; ++b) {
                            const float air2 = expf(-2.f * S.m_air[b] * d_run);
                            atomicAdd(&e_buf[((long long)b * Nr + rr) * Nt + n0], Kspec * Rp[b] * Rp[b] * air2);
                        }
                }
            }
        }
        if (!hit) break;
        d_travel += seg;
        if (d_travel / P.c >= t_end) break;
        const int gid_t = S.group[tri];
        const F3 N = S.nrm[tri];
        const float ct = fabsf(dot(d, N));
        float Rrefl[kBands];
        for (int b = 0; b < kBands; ++b) {
            const float R = (gid_t >= 0 && gid_t < S.ngroups) ? refl_at(S.refl[gid_t * kBands + b], ct) : 1.f;
            Rrefl[b] = Rp[b] * fabsf(R);
        }
        const float sg = (gid_t >= 0 && gid_t < S.ngroups) ? S.sctr[gid_t] : 0.f;
        // diffuse rain: this bounce as a Lambert secondary source seen by each receiver
        if (sg > 0.f) {
            const F3 No = (dot(N, d) > 0.f) ? scl(N, -1.f) : N;
            for (int rr = 0; rr < Nr; ++rr) {
                const F3 br = sub(S.rec[rr], p1);
                const float d2 = len2(br);
                if (d2 < 1e-8f) continue;
                const float drec = sqrtf(d2);
                const F3 u = scl(br, 1.f / drec);
                const float cosr = dot(No, u);
                if (cosr <= 0.f) continue;
                const F3 so = f3(p1.x + No.x * 1e-4f, p1.y + No.y * 1e-4f, p1.z + No.z * 1e-4f);
                float sh_t;
                int sh_tri;
                if (closest_hit(S, so, u, 1e-4f, sh_t, sh_tri) && sh_t < drec - 1e-3f) continue;
                const float d_run = d_travel + drec;
                if (d_run / P.c >= t_end) continue;
                const int n0 = (int)((d_run / P.c) * P.fs);
                if (n0 < 0 || n0 >= Nt) continue;
                const float geo = cosr * inv_pi / fmaxf(d2, R2);
                for (int b = 0; b < kBands; ++b) {
                    const float air2 = expf(-2.f * S.m_air[b] * d_run);
                    atomicAdd(&S.e_diff[((long long)b * Nr + rr) * Nt + n0],
                              Kdiff * sg * Rrefl[b] * Rrefl[b] * air2 * geo);
                }
            }
        }
        float alive = 0.f;
        for (int b = 0; b < kBands; ++b) {
            Rp[b] = Rrefl[b];
            alive = fmaxf(alive, Rp[b]);
        }
        if (alive < 1e-4f) break;
        // on-off scattering: with probability s a Lambert direction, else the mirror one
        const unsigned seed = ((unsigned)(((P.s_base + s) * P.n_rays + ri)) * 2654435761u) ^ ((unsigned)bounce * 40503u);
        const bool diffuse = (sg > 0.f && hash_u01(seed) < sg);
        F3 ref;
        if (diffuse) {
            const F3 No = (dot(N, d) > 0.f) ? scl(N, -1.f) : N;
            ref = lambert(No, hash_u01(seed ^ 0x9e3779b9u), hash_u01(seed ^ 0x85ebca6bu));
            scattered = true;
        } else {
            ref = reflected(d, N);
        }
        after_diffuse = diffuse;
        o = f3(p1.x + ref.x * 1e-4f, p1.y + ref.y * 1e-4f, p1.z + ref.z * 1e-4f);
        d = ref;
    }
}

// device copy of a host vector (released by the owner's list)
struct DevList {
    std::vector<void *> p;
    ~DevList() {
        for (void *q : p) (void)hipFree(q);
    }
    template <typename T> hipError_t up(const std::vector<T> &v, const T **out) {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(1, v.size()) * sizeof(T));
        if (e != hipSuccess) return e;
        p.push_back(q);
        if (!v.empty()) e = hipMemcpy(q, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
        *out = static_cast<const T *>(q);
        return e;
    }
};

#define GA_CHK(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            err = std::string(#expr) + ": " + hipGetErrorString(e_);                  \
            return FASES_E_HIP;                                                       \
        }                                                                             \
    } while (0)

}  // namespace

int ism_valid_gpu(int device, const IsmGeom &g, const std::vector<Node> &nodes, const F3 &src,
                  const std::vector<F3> &recs, float seg_eps, std::vector<uint8_t> &valid, std::string &err) {
    GA_CHK(hipSetDevice(device));
    DevList L;
    IsmDev D;
    GA_CHK(L.up(g.ta, &D.ta));
    GA_CHK(L.up(g.tb, &D.tb));
    GA_CHK(L.up(g.tc, &D.tc));
    GA_CHK(L.up(g.tri2plane, &D.tri2plane));
    GA_CHK(L.up(g.pn, &D.pn));
    GA_CHK(L.up(g.pd, &D.pd));
    GA_CHK(L.up(g.pband, &D.pband));
    GA_CHK(L.up(g.poff, &D.poff));
    GA_CHK(L.up(g.ptri, &D.ptri));
    GA_CHK(L.up(nodes, &D.nodes));
    GA_CHK(L.up(recs, &D.recs));
    D.ntri = (int)g.ta.size();
    D.nleaf = (int)nodes.size() - 1;
    D.nrec = (int)recs.size();
    D.src = src;
    D.eps = seg_eps;
    const long long n = (long long)D.nleaf * D.nrec;
    valid.assign((size_t)n, 0);
    if (n == 0) return FASES_OK;
    uint8_t *dv = nullptr;
    GA_CHK(hipMalloc(&dv, (size_t)n));
    L.p.push_back(dv);
    const long long blocks = (n + 63) / 64;
    hipLaunchKernelGGL(k_ism_valid, dim3((unsigned)blocks), dim3(64), 0, 0, D, dv);
    GA_CHK(hipGetLastError());
    GA_CHK(hipMemcpy(valid.data(), dv, (size_t)n, hipMemcpyDeviceToHost));
    return FASES_OK;
}

int rt_trace_gpu(int device, const RtScene &S, const RtParams &P, const std::vector<F3> &src,
                 const std::vector<F3> &rec, const std::vector<F3> &dirs, const std::vector<Refl> &refl,
                 const std::vector<float> &sctr, const std::vector<float> &m_air, float *e_spec, float *e_diff,
                 double *ms, std::string &err) {
    GA_CHK(hipSetDevice(device));
    DevList L;
    RtDev D;
    GA_CHK(L.up(S.nodes, &D.nodes));
    GA_CHK(L.up(S.a, &D.a));
    GA_CHK(L.up(S.b, &D.b));
    GA_CHK(L.up(S.c, &D.c));
    GA_CHK(L.up(S.nrm, &D.nrm));
    GA_CHK(L.up(S.group, &D.group));
    GA_CHK(L.up(refl, &D.refl));
    GA_CHK(L.up(sctr, &D.sctr));
    GA_CHK(L.up(dirs, &D.dirs));
    GA_CHK(L.up(src, &D.src));
    GA_CHK(L.up(rec, &D.rec));
    D.ngroups = S.ngroups;
    for (int b = 0; b < kBands; ++b) D.m_air[b] = b < (int)m_air.size() ? m_air[(size_t)b] : 0.f;
    const size_t ne = (size_t)kBands * P.n_rec * P.Nt;
    float *es = nullptr, *ed = nullptr;
    GA_CHK(hipMalloc(&es, ne * sizeof(float)));
    L.p.push_back(es);
    GA_CHK(hipMalloc(&ed, ne * sizeof(float)));
    L.p.push_back(ed);
    // previous source chunks accumulate: start from the host buffers
    GA_CHK(hipMemcpy(es, e_spec, ne * sizeof(float), hipMemcpyHostToDevice));
    GA_CHK(hipMemcpy(ed, e_diff, ne * sizeof(float), hipMemcpyHostToDevice));
    D.e_spec = es;
    D.e_diff = ed;
    const long long n = (long long)P.n_src * P.n_rays;
    hipEvent_t e0, e1;
    GA_CHK(hipEventCreate(&e0));
    GA_CHK(hipEventCreate(&e1));
    GA_CHK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL(k_rt, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, 0, D, P);
    GA_CHK(hipGetLastError());
    GA_CHK(hipEventRecord(e1, 0));
    GA_CHK(hipEventSynchronize(e1));
    float t = 0.f;
    GA_CHK(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = t;
    GA_CHK(hipMemcpy(e_spec, es, ne * sizeof(float), hipMemcpyDeviceToHost));
    GA_CHK(hipMemcpy(e_diff, ed, ne * sizeof(float), hipMemcpyDeviceToHost));
    return FASES_OK;
}

}  // namespace ga
}  // namespace fases
